// sg_gradient.hip -- analytic nuclear gradient of the MNDO total energy.
//
// The reference stops at the SCF energy; its README lists gradients and a
// geometry optimiser as to do (README.md:109-110).  In NDDO the SCF basis is
// orthonormal by assumption (no overlap matrix in the eigenproblem), so the
// converged density P is stationary under every allowed variation and the
// gradient is the derivative of the energy at fixed P.  Split by atom pairs,
//
//   E_AC(P) = sum_kl D_kl (k|l)_AC                      two-electron and core attraction
//           + sum_{mu in A, nu in C} P_mu,nu (beta_mu + beta_nu) S_mu,nu   resonance
//           + Z_A Z_C (ss|ss)_AC f(R_AC)                core-core
//
// with (k|l) the pair's rotated integrals (slot k of A's MOPAC pairs, slot l
// of C's, as stored in the pair block), S the STO-6G overlaps and
//
//   D_kl = w_A(k) w_C(l) - 1/2 X_kl - Z_C w_A(k) [l = ss] - Z_A w_C(l) [k = ss],
//   w_A(k) = P_ab (2 - [a = b]) for slot k = (a, b) of A (the Coulomb weights
//            of the Fock build, fockmatrix.cpp / twoelectronmatrix.cpp:14-35),
//   X_kl  = sum over the orderings (mu, lambda) of slot k and (nu, sigma) of
//           slot l of P_mu,nu P_lambda,sigma (the exchange of the off-diagonal
//           Fock blocks),
//
// the core attraction being the (mu nu | s_C s_C) terms of the Hcore diagonal
// (hcore.cpp:71-92).  dE/dR_C of every pair comes from the same pair math as
// the integral and Hcore kernels (sg_pair_math.h), run on dual numbers seeded
// with one coordinate of C (one lane per pair, the three directions in turn);
// dE/dR_A = -dE/dR_C.  The pair derivatives go to the molecule's eigensolver
// workspace, and each atom adds its pairs' in a fixed order.  The pair
// functions are in sg_gradient_pair.h.
#include "sg_common.h"
#include "sg_kernels.h"
#include "sg_pair_math.h"
#include "sg_gradient_pair.h"

namespace sg {

constexpr double kAngToBohr = 1.8897259886;  // atomicOrbitals.cpp:40-44

using namespace grad;

// Hydrogen pairs (KIND 1: one H, 2: both): one lane per pair, the three
// directions in turn.
template <int KIND>
__global__ void __launch_bounds__(64) k_grad_pairs(const int2* __restrict__ pairs, int n,
                                                    const AtomDesc* __restrict__ atoms,
                                                    const MolDesc* __restrict__ mols,
                                                    const double* __restrict__ Pm,
                                                    double* __restrict__ work) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int2 pr = pairs[i];
  const AtomDesc A = atoms[pr.x], C = atoms[pr.y];
  const MolDesc M = mols[A.mol];
  const Dens P{Pm + M.mat0, M.n_ao};
  double* g = work + 5 * M.mat0 + 3 * pair_slot(A.local, C.local);
#pragma unroll 1
  for (int dir = 0; dir < 3; ++dir) g[dir] = pair_grad<KIND>(A, C, P, dir);
}

// Heavy pairs: one wavefront per pair.  Per direction every lane forms the
// pair's local integrals and frame on dual numbers (lane 0 leaves the frame
// in LDS), lane l takes the entries l and l + 64 of the 100, weighted by their
// density coefficients, the lanes' sums are added by a fixed butterfly, and
// lane 0 adds the resonance and core-core terms.
__global__ void __launch_bounds__(256) k_grad_xx(const int2* __restrict__ pairs, int n,
                                                 const AtomDesc* __restrict__ atoms,
                                                 const MolDesc* __restrict__ mols,
                                                 const double* __restrict__ Pm,
                                                 double* __restrict__ work) {
  __shared__ double frame[4][18];  // per wave: the frame's 9 values and derivatives
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + wave;
  if (i >= n) return;  // the whole wavefront
  const int2 pr = pairs[i];
  const AtomDesc A = atoms[pr.x], C = atoms[pr.y];
  const MolDesc M = mols[A.mol];
  const Dens P{Pm + M.mat0, M.n_ao};
  double* g = work + 5 * M.mat0 + 3 * pair_slot(A.local, C.local);
  double* fr = frame[wave];
  const auto m = [fr](int r, int c) { return Dual(fr[2 * (3 * r + c)], fr[2 * (3 * r + c) + 1]); };
#pragma unroll 1
  for (int dir = 0; dir < 3; ++dir) {
    const Dual ac[3] = {Dual(A.x), Dual(A.y), Dual(A.z)};
    const Dual cc[3] = {Dual(C.x, dir == 0 ? 1.0 : 0.0), Dual(C.y, dir == 1 ? 1.0 : 0.0),
                        Dual(C.z, dir == 2 ? 1.0 : 0.0)};
    const Dual R = dist3(ac[0], ac[1], ac[2], cc[0], cc[1], cc[2]);
    {
      const FrameT<Dual> F = local_frame_xyz(ac[0], ac[1], ac[2], cc[0], cc[1], cc[2], R);
      if (lane == 0)
        for (int q = 0; q < 9; ++q) {
          fr[2 * q] = F.m[q / 3][q % 3].v;
          fr[2 * q + 1] = F.m[q / 3][q % 3].d;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const LocalXXT<Dual> v = local_xx(R, A.zn, C.zn);
    double s = 0.0;
    for (int kl = lane; kl < 100; kl += 64) {
      const int k = kl / 10, l = kl % 10;
      s += coef(P, A, k, C, l) * xx_entry_rt<Dual>(v, m, k, l).d;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) {
      Dual ab = 0.0;
      const Dual d0 = ac[0] - cc[0], d1 = ac[1] - cc[1], d2 = ac[2] - cc[2];
      ab += d0 * d0;
      ab += d1 * d1;
      ab += d2 * d2;
      s += resonance(P, A, ac, C, cc, ab).d;
      s += core_term(A, C, R).d;
      g[dir] = s;
    }
    // every lane has read the frame before the next direction rewrites it
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// Atom i: minus its pairs as row atom (columns in order), plus its pairs as
// column atom (rows in order), in eV / Angstrom.
__global__ void k_grad_atoms(int n_atoms, const AtomDesc* __restrict__ atoms,
                             const MolDesc* __restrict__ mols, const double* __restrict__ work,
                             double* __restrict__ grad) {
  const int gi = blockIdx.x * blockDim.x + threadIdx.x;
  if (gi >= n_atoms) return;
  const AtomDesc A = atoms[gi];
  const MolDesc M = mols[A.mol];
  const double* g = work + 5 * M.mat0;
  const int a = A.local;
  double s[3] = {0.0, 0.0, 0.0};
  for (int j = 0; j < a; ++j)
    for (int k = 0; k < 3; ++k) s[k] -= g[3 * pair_slot(a, j) + k];
  for (int r = a + 1; r < M.n_atoms; ++r)
    for (int k = 0; k < 3; ++k) s[k] += g[3 * pair_slot(r, a) + k];
  for (int k = 0; k < 3; ++k) grad[3 * (int64_t)gi + k] = s[k] * kAngToBohr;
}

static inline int blocks_for(int64_t n, int t) { return (int)((n + t - 1) / t); }

void launch_gradient(const BatchDev& b, hipStream_t s, double* grad) {
  if (b.n_xx)
    k_grad_xx<<<blocks_for(b.n_xx, 4), 256, 0, s>>>(b.pairs_xx, b.n_xx, b.atoms, b.mols, b.P, b.W);
  if (b.n_hx)
    k_grad_pairs<1><<<blocks_for(b.n_hx, 64), 64, 0, s>>>(b.pairs_hx, b.n_hx, b.atoms, b.mols, b.P,
                                                         b.W);
  if (b.n_hh)
    k_grad_pairs<2><<<blocks_for(b.n_hh, 64), 64, 0, s>>>(b.pairs_hh, b.n_hh, b.atoms, b.mols, b.P,
                                                         b.W);
  k_grad_atoms<<<blocks_for(b.n_atoms, 128), 128, 0, s>>>(b.n_atoms, b.atoms, b.mols, b.W, grad);
}

}  // namespace sg
