/*
 * mndo_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C FP64 restatement of the reference MNDO SCF path
 * (LuisAlfredoNu/semiGPU, the src/common sources).  It is the checker for the HIP
 * product path and the CPU baseline in bench.py; nothing in the product links
 * it.  Each function cites the reference function it restates.
 *
 * Units: coordinates in Angstrom at the interface (converted to bohr with the
 * reference's 1.8897259886), matrices and energies in eV, matrices as packed
 * lower triangles (index r(r+1)/2 + c), eigenvectors column-major.
 *
 * Pinning (what checks each function; see tests/test_oracle_golden.py):
 *  - or_pair_blocks, or_hcore, or_overlap, or_fock_ref, or_energy, or_core_core,
 *    or_guess_density, or_density: bit-for-bit against the compiled reference
 *    (oracle/_ref/ref_dump) on every shipped geometry, via committed fixtures.
 *  - or_fock_nddo: against or_fock_ref (relative 1e-12) and the fixtures.
 *  - or_eigh: against numpy/LAPACK eigh, F C = C e and C^T C = I.
 *  - or_scf: converged energies against the fixtures (1e-6 eV) and the README.
 */
#ifndef MNDO_ORACLE_H
#define MNDO_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Number of AOs for a molecule (1 for H, 4 for C/N/O); -1 on unsupported Z. */
int or_n_ao(int n_atoms, const int* Z);

/* Number of doubles in the reference's jagged pair-block array
 * (src/common/twocenterintegral.cpp:55-107): 1 (H-H), 10 (H-X), 100 (X-X)
 * for every atom pair i >= j, one-center blocks included. */
long or_pair_block_count(int n_atoms, const int* Z);

/* All two-center integral blocks in the reference order [i][j<=i][k][l]
 * (ComputeAllTwoCenterIntegral, twocenterintegral.cpp:229-279). */
int or_pair_blocks(int n_atoms, const int* Z, const double* xyz_ang, double* out);

/* One integral (mu nu | lambda sigma) by global AO indices, looked up in the
 * blocks exactly like TwoCenterIntegral::GetValueFromArray (:285-386). */
double or_get_value(int n_atoms, const int* Z, const double* blocks, int a, int b, int c, int d);

/* STO-6G overlap for AOs a, b (Overlap::ComputeOverlap, overlap.cpp:68-134). */
double or_overlap(int n_atoms, const int* Z, const double* xyz_ang, int a, int b);

/* Packed Hcore (Hcore::ComputeElementMatrix, hcore.cpp:25-99). */
int or_hcore(int n_atoms, const int* Z, const double* xyz_ang, const double* blocks, double* h);

/* Guess density (densitymatrix.cpp:26-36). */
int or_guess_density(int n_atoms, const int* Z, double* p);

/* Density from column-major eigenvectors C (densitymatrix.cpp:38-45). */
int or_density(int n_ao, int n_occ, const double* C, double* p);

/* Valence electrons from core charges (densitymatrix.cpp:99-107). */
int or_n_electrons(int n_atoms, const int* Z);

/* Fock matrix, reference O(N^4) element loop (fockmatrix.cpp:20-26,
 * twoelectronmatrix.cpp:14-35). */
int or_fock_ref(int n_atoms, const int* Z, const double* blocks, const double* h,
                const double* p, double* f);

/* Fock matrix, same sum restricted to its non-zero atom-pair blocks
 * (O(N_atoms^2)); differs from or_fock_ref only by summation order. */
int or_fock_nddo(int n_atoms, const int* Z, const double* blocks, const double* h,
                 const double* p, double* f);

/* E_el = 1/2 sum P (H + F) (electronicenergy.cpp:12-36). */
double or_energy(int n_ao, const double* h, const double* f, const double* p);

/* Core-core repulsion (corecorerepulsion.cpp:12-65). */
double or_core_core(int n_atoms, const int* Z, const double* xyz_ang, const double* blocks);

/* Dense symmetric eigensolver standing in for LAPACK dsyev('V','L')
 * (scfcalculation.cpp:171-203): Householder tridiagonalisation followed by the
 * implicit QL method with Wilkinson shifts (Golub & Van Loan, Matrix
 * Computations, 4th ed., 8.3).  a is n*n column-major (only the lower
 * triangle is read); on return it holds the eigenvectors as columns, w the
 * ascending eigenvalues.  Returns 0 or the index of a non-converged value. */
int or_eigh(int n, double* a, double* w);

/* SCF outputs. */
typedef struct {
  double e_el;        /* final electronic energy, eV */
  double e_cc;        /* core-core repulsion, eV */
  double e_tot;       /* e_el + e_cc */
  int steps;          /* SCF iterations done */
  int converged;      /* |E_k - E_(k-1)| < threshold */
  int status;         /* 0 ok, >0 eigensolver failure, <0 bad input */
} or_scf_result;

/* Full SCF (SCFCalculation::ComputeSCF, scfcalculation.cpp:66-165).
 * fock_mode 0 = reference O(N^4) Fock, 1 = pair-blocked Fock.
 * energies (optional, size max_steps+1) gets E_0 .. E_steps.
 * p_out / f_out / eig_out (optional) get the final packed P, F and eigenvalues. */
int or_scf(int n_atoms, const int* Z, const double* xyz_ang, double threshold, int max_steps,
           int fock_mode, or_scf_result* res, double* energies, double* p_out, double* f_out,
           double* eig_out);

#ifdef __cplusplus
}
#endif
#endif
