// sg_kernels.h -- device batch layout and kernel launchers (internal).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sg_common.h"

namespace sg {

// Largest basis the eigensolver handles in one workgroup (LDS-resident
// tridiagonal, 20 x 64-column row chunks in the reduction).
constexpr int kMaxBasis = 1280;

// Device view of a batch of molecules.  All matrices are dense N x N per
// molecule, row-major (symmetric ones are stored in full); eigenvectors are
// column-major (C[k*N + i] = component i of eigenvector k), which is the
// layout the reference hands to dsyev (scfcalculation.cpp:175-179).
struct BatchDev {
  int n_mol = 0, n_atoms = 0;
  int64_t n_mat = 0;     // sum N^2
  int64_t n_vec = 0;     // sum N
  int64_t n_blocks = 0;  // sum of pair-block doubles
  MolDesc* mols = nullptr;
  AtomDesc* atoms = nullptr;
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
  double* tau = nullptr;     // Householder scalars (n_vec)
  double* dT = nullptr;      // tridiagonal diagonal (n_vec)
  double* eT = nullptr;      // tridiagonal off-diagonal (n_vec)
  double* tnorm = nullptr;   // Gershgorin bound of each tridiagonal (n_mol)
  double* e_cc = nullptr;    // per molecule
  double* e_el = nullptr;    // current electronic energy
  double* e_new = nullptr;   // energy of the step just computed
  int* steps = nullptr;
  int* active = nullptr;     // 1 while the SCF of the molecule iterates
  int* converged = nullptr;
  int* status = nullptr;     // 0 ok, >0 eigensolver failure
  int* n_active = nullptr;   // device counter
  int* prec = nullptr;       // precision of the next SCF step: 1 = FP32 contractions, 0 = FP64
  int* prev_prec = nullptr;  // precision of the step just done
  int max_n = 0;             // largest N in the batch
  // Molecule indices by decreasing N (or null for 0..n_mol-1): the per-molecule
  // eigensolver and density launches take their workgroups in this order, so
  // the longest jobs start first and the launch tail is short.
  const int* order = nullptr;
};

// `mask` (optional, per molecule) restricts a launch to the molecules whose
// entry is non-zero; it overrides the SCF's own `active` flags.
void launch_pair_integrals(const BatchDev& b, hipStream_t s, const int* mask = nullptr);
void launch_hcore(const BatchDev& b, hipStream_t s);
void launch_core_core(const BatchDev& b, hipStream_t s, const int* mask = nullptr);
void launch_guess_density(const BatchDev& b, hipStream_t s, const int* mask = nullptr);
void launch_fock(const BatchDev& b, hipStream_t s, bool all, const int* mask = nullptr);
void launch_energy(const BatchDev& b, hipStream_t s, double* out, bool all,
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
// Eigensolver of F (all = every eigenpair, else the selected subspace of the
// active molecules): tridiagonalisation, bisection, inverse iteration,
// back-transformation; launch_eigh runs the four in order.
void launch_eigh_tridiag(const BatchDev& b, hipStream_t s, bool all);
void launch_eigh_bisect(const BatchDev& b, hipStream_t s, bool all);
void launch_eigh_invit(const BatchDev& b, hipStream_t s, bool all);
void launch_eigh_back(const BatchDev& b, hipStream_t s, bool all);
void launch_eigh(const BatchDev& b, hipStream_t s, bool all);
void launch_density(const BatchDev& b, hipStream_t s, bool all);

}  // namespace sg
