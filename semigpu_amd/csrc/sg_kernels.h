// sg_kernels.h -- device batch layout and kernel launchers (internal).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sg_common.h"

namespace sg {

// Largest basis the eigensolver handles in one workgroup (LDS-resident
// tridiagonal, 20 x 64-column row chunks in the reduction).
constexpr int kMaxBasis = 1280;
// Largest basis of the multi-workgroup eigensolver (bisection keeps d and
// e^2 of one tridiagonal in LDS: 2 x 8192 doubles).
constexpr int kMaxLarge = 8192;
// Smallest basis the multi-workgroup eigensolver accepts (its panels and
// tables share the last N^2 doubles of the molecule's workspace).
constexpr int kMinLarge = 128;
constexpr int kLargeNb = 32;  // reflectors per panel of the multi-workgroup solver
constexpr int kDiisMax = 8;   // largest DIIS depth
// Molecules up to this basis size take the Fock build with one workgroup per
// molecule (k_fock_mol, every pair block read once); larger ones the build
// with one wavefront per atom (k_fock).
constexpr int kFockMolMaxN = 512;
constexpr int kFockMolWaves = 4;
// Dynamic LDS of k_fock_mol for a molecule of N basis functions and na atoms:
// the Coulomb weights and kFockMolWaves copies of the Coulomb sums (10 entries
// per heavy atom, 1 per hydrogen: 3 N - 2 na), the one-center exchange blocks
// (16 or 1: 5 N - 4 na) and one energy share per atom.
__host__ __device__ inline int64_t fock_mol_lds(int n, int na) {
  return (int64_t)sizeof(double) *
         ((int64_t)(kFockMolWaves + 1) * (3 * n - 2 * na) + (5 * n - 4 * na) + na);
}

// Device view of a batch of molecules.  All matrices are dense N x N per
// molecule, row-major (symmetric ones are stored in full); eigenvectors are
// column-major (C[k*N + i] = component i of eigenvector k), which is the
// layout the reference hands to dsyev (scfcalculation.cpp:175-179).
// Size buckets of the FP64 reduction of the one-workgroup eigensolver
// (k_tridiag_reg): largest first, bucket s takes the molecules with
// kTriBucketMin[s] < N (<= the previous bucket's minimum); the first bucket
// also takes those above kRegRows after k_tridiag's first N - kRegRows columns.
constexpr int kTriBuckets = 8;
constexpr int kRegRows = 256;
constexpr int kTriBucketMin[kTriBuckets] = {224, 192, 160, 128, 96, 64, 32, 0};
// Launch shapes a bucket can take (wavefronts per molecule, register tiles per
// wavefront), all with the same arithmetic: the autotuner picks one per
// bucket and device (sg_ctx_autotune), index 0 by default.
constexpr int kTriCandidates = 3;

struct BatchDev {
  int n_mol = 0, n_atoms = 0;
  int64_t n_mat = 0;     // sum N^2
  int64_t n_vec = 0;     // sum N
  int64_t n_blocks = 0;  // sum of pair-block doubles
  MolDesc* mols = nullptr;
  AtomDesc* atoms = nullptr;
  // per molecule, over its atom range: copies of the descriptors of its heavy
  // atoms, then its hydrogens (the Fock build's partner order; a copy rather
  // than an index saves the Fock wavefronts one dependent load)
  AtomDesc* partners = nullptr;
  // type-sorted atom-pair lists (row atom, column atom), row local > column local
  int2* pairs_xx = nullptr; int n_xx = 0;
  int2* pairs_hx = nullptr; int n_hx = 0;
  int2* pairs_hh = nullptr; int n_hh = 0;
  int2* pairs_all = nullptr; int n_pairs_all = 0;
  int2* diag_tasks = nullptr; int n_diag_tasks = 0;  // (atom, slot)
  double* blocks = nullptr;  // pair integrals, reference jagged order per molecule
  float* blocks32 = nullptr; // FP32 copy of the pair integrals (mixed-precision steps)
  double* H = nullptr;
  double* F = nullptr;
  double* P = nullptr;
  double* C = nullptr;       // eigenvectors
  double* W = nullptr;       // eigensolver workspace (5 * n_mat)
  double* eig = nullptr;     // eigenvalues (n_vec; the selected ones first)
  // how far each selected eigenvalue moved at its last bisection (n_vec; or
  // null): the next SCF step brackets it by that change
  double* dlam = nullptr;
  double* tau = nullptr;     // Householder scalars (n_vec)
  double* dT = nullptr;      // tridiagonal diagonal (n_vec)
  double* eT = nullptr;      // tridiagonal off-diagonal (n_vec)
  double* tnorm = nullptr;   // Gershgorin bound of each tridiagonal (n_mol)
  double* e_cc = nullptr;    // per molecule
  double* e_el = nullptr;    // current electronic energy
  double* e_new = nullptr;   // energy of the step just computed
  // opt-in per-iteration record (sg_scf_record_energies): E_el of the guess
  // and of every step k < e_hist_cap at e_hist[m * e_hist_cap + k]
  double* e_hist = nullptr;
  int e_hist_cap = 0;
  double* e_atom = nullptr;  // per atom: its share of sum P (H + F), written by the Fock build
  int* steps = nullptr;
  int* active = nullptr;     // 1 while the SCF of the molecule iterates
  int* converged = nullptr;
  int* status = nullptr;     // 0 ok, >0 eigensolver failure
  int* n_active = nullptr;   // device counter
  int* prec = nullptr;       // precision of the next SCF step: 1 = FP32 contractions, 0 = FP64
  int* prev_prec = nullptr;  // precision of the step just done
  int* steps32 = nullptr;    // FP32 steps of the current SCF per molecule (null: not counted)
  double* de_prev = nullptr; // |dE| of the molecule's previous SCF step
  // per atom, 10 entries: the Coulomb weights w(l) of its diagonal density
  // block (P_l, doubled off the diagonal), gathered before each Fock build
  double* wdiag = nullptr;
  // The SCF runs a mixed-precision policy: only then can a step's molecules
  // be at reduced precision (prec 1: FP32 Fock and density; 2: FP32 Fock and
  // split-BF16 density).
  int mixed = 0;
  // The reduced steps also run the one-workgroup eigensolver's reduction and
  // back-transform in FP32 (SG_POLICY_MIXED_EIG32): only then do the
  // eigensolver launches take the step precision.
  int eig32 = 0;
  // Density damping of the SCF (opt-in, 0 = off): each step's density is
  // beta P_old + (1 - beta) P_new, with P_old kept in C meanwhile.
  double damping = 0.0;
  // Pulay DIIS of the SCF (opt-in, depth 0 = off): the last diis_m Fock and
  // error matrices of each molecule (slot j at diis_F / diis_E + j n_mat +
  // mat0), the B matrices (kDiisMax^2 per molecule), the dot products'
  // partial sums, the coefficients (kDiisMax per molecule), the number of
  // vectors stored in the current SCF and their precision (1 = FP32 steps).
  int diis_m = 0;
  int64_t diis_stride = 0;  // doubles between history slots (>= the molecules' matrix range)
  double* diis_F = nullptr;
  double* diis_E = nullptr;
  double* diis_B = nullptr;
  double* diis_part = nullptr;
  double* diis_c = nullptr;
  int* diis_k = nullptr;
  int* diis_p = nullptr;
  int max_n = 0;             // largest N in the batch
  // dynamic LDS of the one-workgroup-per-molecule Fock build: the largest
  // fock_mol_lds over the molecules up to kFockMolMaxN (0: none)
  int fock_lds = 0;
  // rounds of the multi-workgroup core-core sum: the largest core_rounds over
  // the molecules (0: none takes it)
  int core_rounds = 0;
  // Molecule indices by decreasing N (or null for 0..n_mol-1): the per-molecule
  // eigensolver and density launches take their workgroups in this order, so
  // the longest jobs start first and the launch tail is short.
  const int* order = nullptr;
  // A view over some of a batch's molecules: the kernels that index molecules
  // directly (energies, SCF control, DIIS) take molecule sel[i] for i < n_mol
  // (null: molecule i).
  const int* sel = nullptr;
  // Molecules with N >= large_min take the multi-workgroup eigensolver (one
  // molecule at a time over the whole GPU).  They are the first n_large
  // entries of `order`; large_idx / large_mols are host copies of their
  // indices and descriptors.  max_n_small is the largest N of the others.
  int large_min = kMaxBasis + 1;
  int n_large = 0;
  const int* large_idx = nullptr;
  const MolDesc* large_mols = nullptr;
  int max_n_small = 0;
  // The molecules of the one-workgroup part (the tail of `order` after the
  // n_large ones) by reduction bucket: positions [tri_start[s], tri_start[s + 1])
  // of that part are bucket s; its first tri_head have N > kRegRows.
  int tri_start[kTriBuckets + 1] = {};
  int tri_head = 0;
  // launch shape of each bucket (an index into its candidates)
  int tri_variant[kTriBuckets] = {};
};

// Candidates of bucket s: *count of them, their (wavefronts, register tiles
// per wavefront) into nw[] / nr[] (kTriCandidates entries).
void tri_candidates(int s, int* count, int* nw, int* nr);
// Times every candidate of every bucket on a synthetic batch of random
// symmetric matrices of the bucket's largest N on stream s (best of `reps`
// launches, milliseconds into ms[s * kTriCandidates + c], -1 where a bucket
// has fewer candidates) and writes the fastest into choice[s].
hipError_t tri_autotune(hipStream_t s, int molecules, int reps, double* ms, int* choice);

// tri_start / tri_head from the N of the one-workgroup part's molecules in
// launch order (non-increasing; free server slots count as N = 0).
void set_tri_buckets(BatchDev& d, const int* n_in_order, int count);

// Molecule m takes part in a launch: every molecule (active null), or one
// still iterating without an eigensolver failure.
__device__ __forceinline__ bool mol_live(const int* active, const int* status, int m) {
  return !active || (active[m] && !status[m]);
}

// Selected eigenvalue range: all, or the smaller of occupied / virtual.
__host__ __device__ inline void selection(const MolDesc& M, int want_all, int& lo, int& k) {
  const int n = M.n_ao, no = M.n_occ;
  if (want_all) {
    lo = 0;
    k = n;
  } else if (no <= n - no) {
    lo = 0;
    k = no;
  } else {
    lo = no;
    k = n - no;
  }
}

// Workspace of a molecule on the multi-workgroup eigensolver (offsets in
// doubles from the molecule's 5 N^2 block):
//   0       A: the reduced matrix, full symmetric; reflector c in the upper part of row c
//   N^2     Y: selected eigenvectors, N x k row-major (as the one-workgroup solver)
//   2 N^2   inverse-iteration factors, 3 N x L
//   4 N^2   V and W panels (kLargeNb x N each), a spare vector (N), the
//           compact-WY T factors (16 x 16 per group of 16 reflectors), the
//           inverse-iteration shifts and cluster tables; during the
//           tridiagonalisation the T, shift and cluster rooms hold the column
//           step's partial sums and the current column;
//   4 N^2 + kLargePyOff N: the product's partial rows, one vector per column
//           chunk (kLargeChunks x N), then their v.py partial dots
//           ((N / 32 + 1) x kLargeChunks)
constexpr int kLargeChunkCols = 1024;  // columns per workgroup of the product
constexpr int kLargeChunks = (kMaxLarge + kLargeChunkCols - 1) / kLargeChunkCols;
constexpr int kLargePyOff = 100;
struct LargeWork {
  double *A, *Y, *fac, *V, *W, *py, *T, *sh;
  int *cstart, *clist;  // clist[k + 1] holds the number of clusters
  double *part, *part2, *xv, *pyp, *partvp;
  __host__ __device__ LargeWork(double* work, const MolDesc& M) {
    const int64_t n = M.n_ao, nn = n * n;
    A = work + 5 * M.mat0;
    Y = A + nn;
    fac = A + 2 * nn;
    V = A + 4 * nn;
    W = V + kLargeNb * n;
    py = W + kLargeNb * n;
    T = py + n;
    sh = T + 16 * n + 256;
    cstart = (int*)(sh + n);
    clist = cstart + n;
    part = T;                              // <= (N/32 + 1) x (2 kLargeNb + 1)
    part2 = sh;                            // <= 4 (N/256 + 1): one per column-kernel wavefront
    xv = (double*)cstart;                  // N (cstart and clist: N + 1 doubles)
    pyp = V + kLargePyOff * n;             // after clist (n >= 128 here)
    partvp = pyp + kLargeChunks * n;
  }
  // vectors whose factors are in flight at once (3 N x L doubles of factors)
  static __host__ __device__ int lanes(int n, int k) {
    int l = (2 * n / 3) & ~63;
    return k < l ? k : l;
  }
};

// `mask` (optional, per molecule) restricts a launch to the molecules whose
// entry is non-zero; it overrides the SCF's own `active` flags.
void launch_pair_integrals(const BatchDev& b, hipStream_t s, const int* mask = nullptr);
void launch_hcore(const BatchDev& b, hipStream_t s);
void launch_core_core(const BatchDev& b, hipStream_t s, const int* mask = nullptr);
void launch_guess_density(const BatchDev& b, hipStream_t s, const int* mask = nullptr);
void launch_fock(const BatchDev& b, hipStream_t s, bool all, const int* mask = nullptr);
// Rounds of kCoreChunk atom pairs of the multi-workgroup core-core sum of a
// molecule with n_atoms atoms (0: its sum runs in one workgroup).
int core_rounds(int n_atoms);
// fock_lds and core_rounds from the host descriptors of the molecules a launch
// can hold (used: null, or non-zero for the slots that hold a molecule)
void set_launch_shapes(BatchDev& d, const MolDesc* mols, int n, const int* used = nullptr);
void launch_energy(const BatchDev& b, hipStream_t s, double* out, bool all,
                   const int* mask = nullptr);
// E_el of the Fock matrices just built by launch_fock (same all / mask), from
// the per-atom shares the Fock build leaves; launch_energy reads H, F and P.
void launch_energy_fock(const BatchDev& b, hipStream_t s, double* out, bool all,
                        const int* mask = nullptr);
void launch_scf_init(const BatchDev& b, hipStream_t s, int max_steps, int policy,
                     const int* mask = nullptr);
// Zeroes the device counter of molecules still iterating (a kernel rather than
// a memset, so that the SCF iteration can be captured in a HIP graph).
void launch_reset_count(const BatchDev& b, hipStream_t s);
void launch_scf_update(const BatchDev& b, hipStream_t s, double threshold, int max_steps,
                       double switch_ev, int fp32_cap);
void launch_blocks_to_fp32(const BatchDev& b, hipStream_t s);
void launch_set_prec(const BatchDev& b, hipStream_t s, int v, const int* mask = nullptr);
void launch_clear_hf(const BatchDev& b, hipStream_t s, const int* mask);
// Density damping around launch_density (active molecules): keep P in C, then
// P <- damping C + (1 - damping) P.
void launch_damp_keep(const BatchDev& b, hipStream_t s);
void launch_damp_mix(const BatchDev& b, hipStream_t s);
// Pulay DIIS after the convergence test (molecules still iterating; the
// history restarts when the step precision changes): store F and F P - P F,
// F <- the extrapolated Fock matrix.
void launch_diis(const BatchDev& b, hipStream_t s);
// At the start of an SCF: every molecule, or the masked ones.
void launch_diis_reset(const BatchDev& b, hipStream_t s, const int* mask = nullptr);
int64_t diis_part_doubles(int n_mol);
// Eigensolver of F (all = every eigenpair, else the selected subspace of the
// active molecules): tridiagonalisation, bisection, inverse iteration,
// back-transformation; launch_eigh runs the four in order.  The reduction's
// launches: which = 0 the on-chip FP64 ones (k_tridiag_reg) and the
// multi-workgroup molecules, 2 the head of the molecules above kRegRows
// (k_tridiag, their first N - kRegRows columns; it must precede 0), 1 the FP32
// one of the mixed_eig32 policy (tridiag_fp32_launch: whether there is one),
// -1 all.
bool tridiag_fp32_launch(const BatchDev& b, bool all);
void launch_eigh_tridiag(const BatchDev& b, hipStream_t s, bool all, int which);
void launch_eigh_bisect(const BatchDev& b, hipStream_t s, bool all);
void launch_eigh_invit(const BatchDev& b, hipStream_t s, bool all);
void launch_eigh_back(const BatchDev& b, hipStream_t s, bool all);
void launch_eigh(const BatchDev& b, hipStream_t s, bool all);
// Multi-workgroup stages for molecule m (descriptor M, host copy); `active`
// null for every molecule, else only while active[m] and !status[m].
void launch_tridiag_large(const BatchDev& b, hipStream_t s, int m, const MolDesc& M,
                          const int* active);
void launch_back_large(const BatchDev& b, hipStream_t s, int m, const MolDesc& M,
                       const int* active, bool all);
void launch_density(const BatchDev& b, hipStream_t s, bool all);
// Analytic nuclear gradient at the current density: grad[3 atom + axis] in
// eV / Angstrom (the pair derivatives pass through each molecule's
// eigensolver workspace).
void launch_gradient(const BatchDev& b, hipStream_t s, double* grad);

}  // namespace sg
