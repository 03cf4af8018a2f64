// sg_fock.hip -- Fock build, guess density, electronic energy and SCF control.
//
// Fock build (FockMatrix::ComputeMatrix, fockmatrix.cpp:20-53 with
// TwoElectronMatrix::ComputeGMatrix_UV, twoelectronmatrix.cpp:14-35).  The
// reference sums P_ij [(uv|ij) - 1/2 (ui|vj)] over all N^2 AO pairs for every
// element; every term with u,v or i,j on different atoms is zero
// (GetValueFromArray returns 0, twocenterintegral.cpp:383-385), so the same
// sum is evaluated here per atom pair:
//   u,v on A:        G_uv = sum_B sum_{l in B} (uv|l) w_B(l) - 1/2 sum_{i,j in A} P_ij (ui|vj)
//   u on A, v on B:  G_uv = -1/2 sum_{i in A, j in B} P_ij (ui|vj)
// with w_B(l) = P_l (x2 for the off-diagonal AO pairs of B).
//
// k_fock: one wavefront per atom A, lanes over the partner atoms B (heavy
// atoms first, then hydrogens, so that a pass of 64 lanes mostly runs one
// block shape).  Each lane
// reads block (A,B) once, adds its Coulomb term to a per-lane J_A, and -- for
// B below A in the molecule -- forms the exchange block K_AB and writes F_AB and
// F_BA.  J_A is reduced across the wave with a fixed butterfly, so every run
// (alone or inside any batch) adds the same numbers in the same order: no
// floating-point atomics anywhere.
#include <algorithm>
#include <type_traits>

#include "sg_common.h"
#include "sg_kernels.h"

namespace sg {

constexpr int kEnergyChunk = 16384;  // entries of a large molecule's E_el per workgroup
constexpr double kFp32Floor = 9.5367431640625e-07;  // 2^-20 = 8 ulps of FP32 at 1
// The same for the split-BF16 density (about 16 bits per product): 2^-14
constexpr double kBf16Floor = 6.103515625e-05;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int K>
struct Slot {
  static constexpr int a = mop_a(K), b = mop_b(K);
};

// Entry (A slot ka | B slot kb) of the block shared by A and B.
template <bool HA, bool HB, bool ROW, typename S>
__device__ __forceinline__ S blk_entry(const S* __restrict__ b, int ka, int kb) {
  if constexpr (HA && HB) return ROW ? b[ka * 10 + kb] : b[kb * 10 + ka];
  else if constexpr (HA) return b[ka];
  else if constexpr (HB) return b[kb];
  else return b[0];
}

// One partner B for atom A.  ROW: A is the block's row atom (B.local <= A.local),
// which is also when this lane owns the exchange block K_AB.  T is the
// arithmetic type of the contraction (double, or float in the FP32 phase of the
// mixed-precision policy, where S = float reads the FP32 copy of the blocks).
template <bool HA, bool HB, bool ROW, typename T, typename S>
__device__ __forceinline__ void partner(const S* __restrict__ b, const double* __restrict__ p,
                                        const double* __restrict__ wb, int N, int ao_a, int ao_b,
                                        T (&J)[10], T (&K)[4][4]) {
  constexpr int NA = HA ? 4 : 1, NB = HB ? 4 : 1;
  constexpr int SA = HA ? 10 : 1, SB = HB ? 10 : 1;
  T w[SB];
#pragma unroll
  for (int l = 0; l < SB; ++l) w[l] = (T)wb[l];
  T pab[NA][NB];
  if constexpr (ROW) {
#pragma unroll
    for (int x = 0; x < NA; ++x)
#pragma unroll
      for (int y = 0; y < NB; ++y) pab[x][y] = (T)p[(int64_t)(ao_a + x) * N + ao_b + y];
  }
#pragma unroll
  for (int k = 0; k < SA; ++k) {
    const int mu = mop_a(k), nu = mop_b(k);
#pragma unroll
    for (int l = 0; l < SB; ++l) {
      const int b1 = mop_a(l), b2 = mop_b(l);
      const T val = (T)blk_entry<HA, HB, ROW>(b, k, l);
      J[k] += val * w[l];
      if constexpr (ROW) {
        K[mu][b1] += val * pab[nu][b2];
        if (mu != nu) K[nu][b1] += val * pab[mu][b2];
        if (b1 != b2) K[mu][b2] += val * pab[nu][b1];
        if (mu != nu && b1 != b2) K[nu][b2] += val * pab[mu][b1];
      }
    }
  }
}

// Writes F_AB and F_BA = H - K/2 and adds the block's share of the
// electronic energy, 2 sum P_uv (H_uv + F_uv) (both triangles), to e.
template <bool HA, bool HB, bool ROW, typename T>
__device__ __forceinline__ void write_offdiag(const double* __restrict__ h,
                                              const double* __restrict__ p, double* __restrict__ f,
                                              int N, int ao_a, int ao_b, const T (&K)[4][4],
                                              double& e) {
  constexpr int NA = HA ? 4 : 1, NB = HB ? 4 : 1;
#pragma unroll
  for (int x = 0; x < NA; ++x)
#pragma unroll
    for (int y = 0; y < NB; ++y) {
      const int64_t r = ao_a + x, c = ao_b + y;
      const double hv = h[r * N + c];
      const double v = hv - 0.5 * (double)K[x][y];
      f[r * N + c] = v;
      f[c * N + r] = v;
      e += 2.0 * p[r * N + c] * (hv + v);
    }
}

// Returns (in every lane) atom A's share of sum P (H + F): the blocks (A, B)
// with B before A counted twice, (A, A) once.
// Partners are taken in the molecule's partner order (heavy atoms, then
// hydrogens), so the lanes of one pass mostly share a block shape.
// kaa: 16 doubles of LDS for this wave (the exchange block K_AA, kept out of
// registers while the other partners run).
template <bool HA, typename T, typename S>
__device__ double fock_atom(const AtomDesc& A, const MolDesc& M, const AtomDesc* __restrict__ atoms,
                            const AtomDesc* __restrict__ partners, const double* __restrict__ wdiag,
                            const S* __restrict__ blk,
                            const double* __restrict__ h, const double* __restrict__ p,
                            double* __restrict__ f, double* __restrict__ kaa, int lane) {
  const int N = M.n_ao;
  T J[10];
#pragma unroll
  for (int k = 0; k < 10; ++k) J[k] = (T)0;
  bool own_diag = false;
  double e = 0.0;
  for (int b0 = 0; b0 < M.n_atoms; b0 += 64) {
    const int bl = b0 + lane;
    if (bl >= M.n_atoms) continue;
    const AtomDesc B = partners[M.atom0 + bl];
    const double* wb = wdiag + (int64_t)(M.atom0 + B.local) * 10;
    const bool row = B.local <= A.local;
    const S* b = blk + (row ? block_offset(A, B) : block_offset(B, A));
    if (row) {
      T K[4][4];
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) K[x][y] = (T)0;
      if (B.zn > 1) partner<HA, true, true>(b, p, wb, N, A.ao0, B.ao0, J, K);
      else partner<HA, false, true>(b, p, wb, N, A.ao0, B.ao0, J, K);
      if (B.local == A.local) {
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
          for (int y = 0; y < 4; ++y) kaa[x * 4 + y] = (double)K[x][y];
        own_diag = true;
      } else if (B.zn > 1) {
        write_offdiag<HA, true, true>(h, p, f, N, A.ao0, B.ao0, K, e);
      } else {
        write_offdiag<HA, false, true>(h, p, f, N, A.ao0, B.ao0, K, e);
      }
    } else {
      T K[4][4];
      if (B.zn > 1) partner<HA, true, false>(b, p, wb, N, A.ao0, B.ao0, J, K);
      else partner<HA, false, false>(b, p, wb, N, A.ao0, B.ao0, J, K);
    }
  }
  constexpr int SA = HA ? 10 : 1;
  double Jd[10];
#pragma unroll
  for (int k = 0; k < SA; ++k) Jd[k] = wave_sum((double)J[k]);
  if (own_diag) {
    constexpr int NA = HA ? 4 : 1;
#pragma unroll
    for (int x = 0; x < NA; ++x)
#pragma unroll
      for (int y = 0; y < NA; ++y) {
        const int64_t r = A.ao0 + x, c = A.ao0 + y;
        const double g = Jd[mop_idx(x, y)] - 0.5 * kaa[x * 4 + y];
        const double hv = h[r * N + c], fv = hv + g;
        f[r * N + c] = fv;
        e += p[r * N + c] * (hv + fv);
      }
  }
  return wave_sum(e);
}

// Coulomb weights of each atom's diagonal density block, in the Fock build's
// order: w(l) = P_l, doubled for the off-diagonal AO pairs (exact in FP32 too).
__global__ void k_wdiag(int n_atoms, const AtomDesc* __restrict__ atoms,
                        const MolDesc* __restrict__ mols, const double* __restrict__ P,
                        double* __restrict__ wdiag, const int* __restrict__ active,
                        const int* __restrict__ status) {
  const int ia = blockIdx.x * blockDim.x + threadIdx.x;
  if (ia >= n_atoms) return;
  const AtomDesc A = atoms[ia];
  if (active && (!active[A.mol] || status[A.mol])) return;
  const MolDesc M = mols[A.mol];
  const int N = M.n_ao;
  if (N <= kFockMolMaxN) return;  // k_fock_mol's molecules
  const double* p = P + M.mat0;
  const int sb = A.zn > 1 ? 10 : 1;
  for (int l = 0; l < sb; ++l) {
    const int x = mop_a(l), y = mop_b(l);
    const double v = p[(int64_t)(A.ao0 + x) * N + A.ao0 + y];
    wdiag[(int64_t)ia * 10 + l] = x == y ? v : 2.0 * v;
  }
}

__global__ void __launch_bounds__(256) k_fock(int n_atoms, const AtomDesc* __restrict__ atoms,
                                               const AtomDesc* __restrict__ partners,
                                               const double* __restrict__ wdiag,
                                               const MolDesc* __restrict__ mols,
                                               const double* __restrict__ blocks,
                                               const float* __restrict__ blocks32,
                                               const double* __restrict__ H,
                                               const double* __restrict__ P, double* __restrict__ F,
                                               double* __restrict__ e_atom,
                                               const int* __restrict__ active,
                                               const int* __restrict__ status,
                                               const int* __restrict__ prec) {
  __shared__ double kaa_lds[4][16];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ia = blockIdx.x * 4 + wave;
  if (ia >= n_atoms) return;
  const AtomDesc A = atoms[ia];
  if (active && (!active[A.mol] || status[A.mol])) return;
  const MolDesc M = mols[A.mol];
  if (M.n_ao <= kFockMolMaxN) return;  // k_fock_mol's molecules
  const double* h = H + M.mat0;
  const double* p = P + M.mat0;
  double* f = F + M.mat0;
  double* kaa = kaa_lds[wave];
  double e;
  if (prec && prec[A.mol]) {
    const float* blk = blocks32 + M.blk0;
    if (A.zn > 1) e = fock_atom<true, float>(A, M, atoms, partners, wdiag, blk, h, p, f, kaa, lane);
    else e = fock_atom<false, float>(A, M, atoms, partners, wdiag, blk, h, p, f, kaa, lane);
  } else {
    const double* blk = blocks + M.blk0;
    if (A.zn > 1) e = fock_atom<true, double>(A, M, atoms, partners, wdiag, blk, h, p, f, kaa, lane);
    else e = fock_atom<false, double>(A, M, atoms, partners, wdiag, blk, h, p, f, kaa, lane);
  }
  if (lane == 0) e_atom[ia] = e;
}

// ---- Fock build with one workgroup per molecule (N <= kFockMolMaxN) ----
//
// Every stored block (A, B), B.local <= A.local, is read once: by the lane
// that takes partner B in the pass of row atom A.  That lane
//   * adds sum_l (k|l) w_B(l) to its Coulomb accumulators of A (summed over
//     the wavefront by a fixed butterfly once the row is done),
//   * adds sum_k (k|l) w_A(k) to B's Coulomb sums in its wavefront's own copy
//     in LDS (one lane per partner: no two lanes of a pass touch one entry),
//   * forms the exchange block K_AB and writes F_AB and F_BA with their share
//     of the electronic energy (K_AA of the one-center block goes to LDS).
// Row atoms go to the wavefronts round robin, heavy partners first, then
// hydrogens (each pass runs one block shape).  When every row is done the
// wavefronts' copies are added in wavefront order and the diagonal blocks
// written, so each sum runs in a fixed order and a molecule's result does not
// depend on the rest of its launch.  The weights w (P_l, doubled for the
// off-diagonal AO pairs of an atom) are gathered into LDS first.
//
// LDS slots: atom A's Coulomb entries start at 9 hv_before + local (10 for a
// heavy atom, 1 for a hydrogen), its K_AA at 15 hv_before + local (16 or 1).
__device__ __forceinline__ int jslot(const AtomDesc& A) { return 9 * A.hv_before + A.local; }
__device__ __forceinline__ int kslot(const AtomDesc& A) { return 15 * A.hv_before + A.local; }

// vector accesses of 8-byte (double) and 4-byte (float) aligned data
typedef double dbl2u __attribute__((ext_vector_type(2), aligned(8)));
typedef float flt4u __attribute__((ext_vector_type(4), aligned(4)));
typedef float flt2u __attribute__((ext_vector_type(2), aligned(4)));

template <int C, typename T>
__device__ __forceinline__ void load_run(const double* __restrict__ q, T (&v)[C]) {
  if constexpr (C == 1) {
    v[0] = (T)q[0];
  } else {
#pragma unroll
    for (int i = 0; i < C / 2; ++i) {
      const dbl2u u = *(const dbl2u*)(q + 2 * i);
      v[2 * i] = (T)u.x;
      v[2 * i + 1] = (T)u.y;
    }
  }
}
template <int C, typename T>
__device__ __forceinline__ void load_run(const float* __restrict__ q, T (&v)[C]) {
  if constexpr (C == 1) {
    v[0] = (T)q[0];
  } else {
    static_assert(C == 10, "runs of 1 or 10");
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const flt4u u = *(const flt4u*)(q + 4 * i);
      v[4 * i] = (T)u.x;
      v[4 * i + 1] = (T)u.y;
      v[4 * i + 2] = (T)u.z;
      v[4 * i + 3] = (T)u.w;
    }
    const flt2u u = *(const flt2u*)(q + 8);
    v[8] = (T)u.x;
    v[9] = (T)u.y;
  }
}

__device__ __forceinline__ float uniform(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ double uniform(double v) {
  const int64_t u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane((int)(u & 0xffffffff));
  const int hi = __builtin_amdgcn_readfirstlane((int)(u >> 32));
  return __longlong_as_double((int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

// One block (A, B) of row atom A, B.local <= A.local (see above).
template <bool HA, bool HB, typename T, typename S>
__device__ __forceinline__ void fock_pair(const AtomDesc& A, const AtomDesc& B, int N,
                                          const S* __restrict__ blk, const double* __restrict__ h,
                                          const double* __restrict__ p, double* __restrict__ f,
                                          const double* wl, double* jw, double* kaa,
                                          const T (&wA)[HA ? 10 : 1], T (&jA)[HA ? 10 : 1],
                                          double& e) {
  constexpr int NA = HA ? 4 : 1, NB = HB ? 4 : 1;
  constexpr int SA = HA ? 10 : 1, SB = HB ? 10 : 1;
  const S* b = blk + block_offset(A, B);
  const int sb = jslot(B);
  T pab[NA][NB];
#pragma unroll
  for (int x = 0; x < NA; ++x) load_run<NB>(p + (int64_t)(A.ao0 + x) * N + B.ao0, pab[x]);
  T K[NA][NB], jB[SB];
#pragma unroll
  for (int x = 0; x < NA; ++x)
#pragma unroll
    for (int y = 0; y < NB; ++y) K[x][y] = (T)0;
#pragma unroll
  for (int l = 0; l < SB; ++l) jB[l] = (T)0;
  // the block is SA x SB entries, read in runs of 10 (one run for X-H and
  // H-X blocks)
  constexpr int RUN = SA * SB == 1 ? 1 : 10;
  T run[RUN];
  // X-X blocks are ten rows of ten, each row group in a basic block of its
  // own (an opaque branch), so that the compiler keeps one group of rows in
  // registers at a time instead of issuing every load first.  In FP64 a group
  // is two rows, whose loads are in flight together, and B's weights are read
  // from LDS per group (an opaque offset keeps them out of registers across
  // the groups); in FP32 (fewer registers per wave) one row.
  if constexpr (HA && HB && sizeof(T) == 8) {
#pragma unroll
    for (int k2 = 0; k2 < 10; k2 += 2) {
      int go = 1;
      asm volatile("" : "+s"(go));
      if (!go) continue;
      T r2[2][10];
      load_run<10>(b + 10 * k2, r2[0]);
      load_run<10>(b + 10 * (k2 + 1), r2[1]);
      int o = sb;
      asm volatile("" : "+v"(o));
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int k = k2 + h;
        const int mu = mop_a(k), nu = mop_b(k);
#pragma unroll
        for (int l = 0; l < 10; ++l) {
          const int b1 = mop_a(l), b2 = mop_b(l);
          const T val = r2[h][l];
          jA[k] += val * (T)wl[o + l];
          jB[l] += val * wA[k];
          K[mu][b1] += val * pab[nu][b2];
          if (mu != nu) K[nu][b1] += val * pab[mu][b2];
          if (b1 != b2) K[mu][b2] += val * pab[nu][b1];
          if (mu != nu && b1 != b2) K[nu][b2] += val * pab[mu][b1];
        }
      }
    }
  } else {
    T wB[SB];
#pragma unroll
    for (int l = 0; l < SB; ++l) wB[l] = (T)wl[sb + l];
#pragma unroll
    for (int k = 0; k < SA; ++k) {
      int go = 1;
      if constexpr (HB) asm volatile("" : "+s"(go));
      if (!go) continue;
      const int mu = mop_a(k), nu = mop_b(k);
      if constexpr (HB) load_run<10>(b + 10 * k, run);
      else if (k == 0) load_run<RUN>(b, run);
#pragma unroll
      for (int l = 0; l < SB; ++l) {
        const int b1 = mop_a(l), b2 = mop_b(l);
        const T val = HB ? run[l] : run[k];
        jA[k] += val * wB[l];
        jB[l] += val * wA[k];
        K[mu][b1] += val * pab[nu][b2];
        if (mu != nu) K[nu][b1] += val * pab[mu][b2];
        if (b1 != b2) K[mu][b2] += val * pab[nu][b1];
        if (mu != nu && b1 != b2) K[nu][b2] += val * pab[mu][b1];
      }
    }
  }
  if (B.local == A.local) {
    // the one-center block: K_AA for the diagonal block, written at the end
#pragma unroll
    for (int x = 0; x < NA; ++x)
#pragma unroll
      for (int y = 0; y < NB; ++y) kaa[kslot(A) + x * NB + y] = (double)K[x][y];
    return;
  }
  // F_AB = H_AB - K_AB / 2 and its share 2 sum P (H + F) of the energy
  double v[NA][NB];
#pragma unroll
  for (int x = 0; x < NA; ++x) {
    const int64_t r = A.ao0 + x;
    double hv[NB], pv[NB];
    load_run<NB>(h + r * N + B.ao0, hv);
    load_run<NB>(p + r * N + B.ao0, pv);
#pragma unroll
    for (int y = 0; y < NB; ++y) {
      v[x][y] = hv[y] - 0.5 * (double)K[x][y];
      e += 2.0 * pv[y] * (hv[y] + v[x][y]);
    }
    if constexpr (NB == 4) {
      *(dbl2u*)(f + r * N + B.ao0) = dbl2u{v[x][0], v[x][1]};
      *(dbl2u*)(f + r * N + B.ao0 + 2) = dbl2u{v[x][2], v[x][3]};
    } else {
      f[r * N + B.ao0] = v[x][0];
    }
  }
  // F_BA = F_AB^T: row B.ao0 + y holds columns A.ao0 .. A.ao0 + NA - 1
#pragma unroll
  for (int y = 0; y < NB; ++y) {
    double* q = f + (int64_t)(B.ao0 + y) * N + A.ao0;
    if constexpr (NA == 4) {
      *(dbl2u*)q = dbl2u{v[0][y], v[1][y]};
      *(dbl2u*)(q + 2) = dbl2u{v[2][y], v[3][y]};
    } else {
      q[0] = v[0][y];
    }
  }
#pragma unroll
  for (int l = 0; l < SB; ++l) jw[sb + l] += (double)jB[l];
}

// Row atom A of a molecule: its blocks with every partner B.local <= A.local,
// heavy partners first (the first hv_before + HA entries of the molecule's
// partner list), then hydrogens (the list's hydrogen part, from n_h).
template <bool HA, typename T, typename S>
__device__ __forceinline__ void fock_row(const AtomDesc& A, const MolDesc& M, int n_h,
                                         const AtomDesc* __restrict__ partners,
                                         const S* __restrict__ blk, const double* __restrict__ h,
                                         const double* __restrict__ p, double* __restrict__ f,
                                         const double* wl, double* jw, double* kaa, double* eacc,
                                         int lane) {
  constexpr int SA = HA ? 10 : 1;
  const int sa = jslot(A), N = M.n_ao;
  T wA[SA], jA[SA];
#pragma unroll
  for (int k = 0; k < SA; ++k) {
    wA[k] = uniform((T)wl[sa + k]);  // the same in every lane: scalar registers
    jA[k] = (T)0;
  }
  double e = 0.0;
  const int nhp = A.hv_before + (HA ? 1 : 0), nHp = A.local + 1 - nhp;
  const AtomDesc* pl = partners + M.atom0;
  for (int c0 = 0; c0 < nhp; c0 += 64) {
    const int i = c0 + lane;
    if (i < nhp) fock_pair<HA, true, T>(A, pl[i], N, blk, h, p, f, wl, jw, kaa, wA, jA, e);
  }
  for (int c0 = 0; c0 < nHp; c0 += 64) {
    const int i = c0 + lane;
    if (i < nHp) fock_pair<HA, false, T>(A, pl[n_h + i], N, blk, h, p, f, wl, jw, kaa, wA, jA, e);
  }
#pragma unroll
  for (int k = 0; k < SA; ++k) {
    const double s = wave_sum((double)jA[k]);
    if (lane == 0) jw[sa + k] += s;
  }
  e = wave_sum(e);
  if (lane == 0) eacc[A.local] = e;
}

template <typename T, typename S>
__device__ __forceinline__ void fock_rows(const MolDesc& M, int n_h,
                                          const AtomDesc* __restrict__ atoms,
                                          const AtomDesc* __restrict__ partners,
                                          const S* __restrict__ blk, const double* __restrict__ h,
                                          const double* __restrict__ p, double* __restrict__ f,
                                          const double* wl, double* jw, double* kaa, double* eacc,
                                          int wave, int lane) {
  for (int a = wave; a < M.n_atoms; a += kFockMolWaves) {
    const AtomDesc A = atoms[M.atom0 + a];
    if (A.zn > 1) fock_row<true, T>(A, M, n_h, partners, blk, h, p, f, wl, jw, kaa, eacc, lane);
    else fock_row<false, T>(A, M, n_h, partners, blk, h, p, f, wl, jw, kaa, eacc, lane);
  }
}

// Diagonal block of atom A once every row is done: J_A = the wavefronts'
// copies added in order, F_AA = H_AA + J_A - K_AA / 2; returns its share
// sum P (H + F) of the energy.
template <bool HA>
__device__ __forceinline__ double fock_diag(const AtomDesc& A, int N, int S, const double* jall,
                                            const double* kaa, const double* __restrict__ h,
                                            const double* __restrict__ p, double* __restrict__ f) {
  constexpr int NA = HA ? 4 : 1, SA = HA ? 10 : 1;
  const int s = jslot(A), ks = kslot(A);
  double J[SA];
#pragma unroll
  for (int k = 0; k < SA; ++k) {
    double t = jall[s + k];
#pragma unroll
    for (int w = 1; w < kFockMolWaves; ++w) t += jall[w * S + s + k];
    J[k] = t;
  }
  double ed = 0.0;
#pragma unroll
  for (int x = 0; x < NA; ++x)
#pragma unroll
    for (int y = 0; y < NA; ++y) {
      const int64_t r = A.ao0 + x, c = A.ao0 + y;
      const double hv = h[r * N + c];
      const double fv = hv + (J[mop_idx(x, y)] - 0.5 * kaa[ks + x * NA + y]);
      f[r * N + c] = fv;
      ed += p[r * N + c] * (hv + fv);
    }
  return ed;
}

// T: the contraction's type.  The FP64 and FP32 builds are separate launches
// (each skips the other's molecules), so that the FP32 one keeps its smaller
// register count.
template <typename T>
__global__ void __launch_bounds__(64 * kFockMolWaves)
__attribute__((amdgpu_waves_per_eu(std::is_same<T, float>::value ? 4 : 3)))
k_fock_mol(int n_mol, const int* __restrict__ order, const MolDesc* __restrict__ mols,
           const AtomDesc* __restrict__ atoms, const AtomDesc* __restrict__ partners,
           const double* __restrict__ blocks, const float* __restrict__ blocks32,
           const double* __restrict__ H, const double* __restrict__ P, double* __restrict__ F,
           double* __restrict__ e_atom, const int* __restrict__ active,
           const int* __restrict__ status, const int* __restrict__ prec) {
  extern __shared__ double lds[];
  const int m = order ? order[blockIdx.x] : blockIdx.x;
  if (!mol_live(active, status, m)) return;
  // FP32 contractions for the molecules whose step runs at reduced precision
  if ((prec && prec[m]) != std::is_same<T, float>::value) return;
  const MolDesc M = mols[m];
  const int N = M.n_ao, na = M.n_atoms;
  if (N > kFockMolMaxN || na == 0) return;  // k_fock's molecules
  const int n_h = (N - na) / 3;             // heavy atoms: N = 4 n_h + (na - n_h)
  const int S = 3 * N - 2 * na, KS = 5 * N - 4 * na;
  double* wl = lds;
  double* jall = wl + S;
  double* kaa = jall + kFockMolWaves * S;
  double* eacc = kaa + KS;
  const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const double* h = H + M.mat0;
  const double* p = P + M.mat0;
  double* f = F + M.mat0;
  // the Coulomb weights of every atom; the Coulomb sums start at zero
  for (int a = tid; a < na; a += 64 * kFockMolWaves) {
    const AtomDesc A = atoms[M.atom0 + a];
    const int s = jslot(A), sb = A.zn > 1 ? 10 : 1;
    for (int l = 0; l < sb; ++l) {
      const int x = mop_a(l), y = mop_b(l);
      const double v = p[(int64_t)(A.ao0 + x) * N + A.ao0 + y];
      wl[s + l] = x == y ? v : 2.0 * v;
    }
  }
  for (int i = tid; i < kFockMolWaves * S; i += 64 * kFockMolWaves) jall[i] = 0.0;
  __syncthreads();
  double* jw = jall + wave * S;
  if constexpr (std::is_same<T, float>::value)
    fock_rows<float>(M, n_h, atoms, partners, blocks32 + M.blk0, h, p, f, wl, jw, kaa, eacc, wave,
                     lane);
  else
    fock_rows<double>(M, n_h, atoms, partners, blocks + M.blk0, h, p, f, wl, jw, kaa, eacc, wave,
                      lane);
  __syncthreads();
  // diagonal blocks: F_AA = H_AA + J_A - K_AA / 2, and each atom's energy share
  for (int a = tid; a < na; a += 64 * kFockMolWaves) {
    const AtomDesc A = atoms[M.atom0 + a];
    const double ed = A.zn > 1 ? fock_diag<true>(A, N, S, jall, kaa, h, p, f)
                               : fock_diag<false>(A, N, S, jall, kaa, h, p, f);
    e_atom[M.atom0 + a] = eacc[a] + ed;
  }
}
// (ElectronicEnergy::ComputeEnergy, electronicenergy.cpp:12-36, of the F
// just built).
__global__ void k_energy_atoms(int n_mol, const int* __restrict__ sel,
                               const MolDesc* __restrict__ mols,
                               const double* __restrict__ e_atom, double* __restrict__ out,
                               const int* __restrict__ active, const int* __restrict__ status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_mol) return;
  const int m = sel ? sel[i] : i;
  if (active && (!active[m] || status[m])) return;
  const MolDesc M = mols[m];
  const double* ea = e_atom + M.atom0;
  double s = 0.0;
  for (int a = 0; a < M.n_atoms; ++a) s += ea[a];
  out[m] = 0.5 * s;
}

__global__ void k_blocks_to_fp32(int64_t n, const double* __restrict__ src, float* __restrict__ dst) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = (float)src[i];
}

// Guess density (densitymatrix.cpp:26-36): P_ii = Z_core / n_AO(atom), 0 elsewhere.
__global__ void k_guess_density(const MolDesc* __restrict__ mols, const AtomDesc* __restrict__ atoms,
                                double* __restrict__ P, const int* __restrict__ mask) {
  if (mask && !mask[blockIdx.x]) return;
  const MolDesc M = mols[blockIdx.x];
  const int N = M.n_ao;
  double* p = P + M.mat0;
  for (int64_t t = threadIdx.x; t < (int64_t)N * N; t += blockDim.x) p[t] = 0.0;
  __syncthreads();
  for (int a = threadIdx.x; a < M.n_atoms; a += blockDim.x) {
    const AtomDesc A = atoms[M.atom0 + a];
    const int na = n_ao_of(A.zn);
    const double v = (double)core_charge(A.zn) / na;
    for (int u = 0; u < na; ++u) p[(int64_t)(A.ao0 + u) * N + A.ao0 + u] = v;
  }
}

// E_el = 1/2 sum_ij P_ij (H_ij + F_ij) (ElectronicEnergy::ComputeEnergy,
// electronicenergy.cpp:12-36), one workgroup per molecule, fixed-order tree.
__global__ void __launch_bounds__(256) k_energy(const MolDesc* __restrict__ mols,
                                                 const double* __restrict__ H,
                                                 const double* __restrict__ F,
                                                 const double* __restrict__ P,
                                                 double* __restrict__ out,
                                                 const int* __restrict__ active,
                                                 const int* __restrict__ status,
                                                 const int* __restrict__ order) {
  const int m = order ? order[blockIdx.x] : blockIdx.x;
  if (active && (!active[m] || status[m])) return;
  const MolDesc M = mols[m];
  const int64_t nn = (int64_t)M.n_ao * M.n_ao;
  const double* h = H + M.mat0;
  const double* f = F + M.mat0;
  const double* p = P + M.mat0;
  double s = 0.0;
  for (int64_t t = threadIdx.x; t < nn; t += blockDim.x) s += p[t] * (h[t] + f[t]);
  __shared__ double red[4];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[m] = 0.5 * ((red[0] + red[1]) + (red[2] + red[3]));
}

// E_el of one large molecule: workgroup g sums entries [g C, (g+1) C) into
// part[g]; k_energy_total adds the parts in order.
__global__ void __launch_bounds__(256) k_energy_chunks(const double* __restrict__ h,
                                                        const double* __restrict__ f,
                                                        const double* __restrict__ p, int64_t nn,
                                                        double* __restrict__ part,
                                                        const int* __restrict__ active,
                                                        const int* __restrict__ status, int m) {
  if (active && (!active[m] || status[m])) return;
  const int64_t t0 = (int64_t)blockIdx.x * kEnergyChunk;
  const int64_t t1 = t0 + kEnergyChunk < nn ? t0 + kEnergyChunk : nn;
  double s = 0.0;
  for (int64_t t = t0 + threadIdx.x; t < t1; t += blockDim.x) s += p[t] * (h[t] + f[t]);
  __shared__ double red[4];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ void __launch_bounds__(256) k_energy_total(const double* __restrict__ part, int chunks,
                                                       double* __restrict__ out,
                                                       const int* __restrict__ active,
                                                       const int* __restrict__ status, int m) {
  if (active && (!active[m] || status[m])) return;
  double s = 0.0;
  for (int g = threadIdx.x; g < chunks; g += blockDim.x) s += part[g];
  __shared__ double red[4];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[m] = 0.5 * ((red[0] + red[1]) + (red[2] + red[3]));
}

__global__ void k_scf_init(int n_mol, int max_steps, int policy, const double* __restrict__ e0,
                           double* __restrict__ e_el, int* __restrict__ steps,
                           int* __restrict__ active, int* __restrict__ converged,
                           int* __restrict__ status, int* __restrict__ n_active,
                           int* __restrict__ prec, int* __restrict__ prev_prec,
                           int* __restrict__ steps32, double* __restrict__ de_prev,
                           const int* __restrict__ mask, double* __restrict__ e_hist,
                           int e_hist_cap) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m == 0 && !mask) *n_active = 0;
  if (m >= n_mol) return;
  if (mask && !mask[m]) return;
  e_el[m] = e0[m];
  if (e_hist) {
    double* h = e_hist + (int64_t)m * e_hist_cap;
    h[0] = e0[m];
    for (int k = 1; k < e_hist_cap; ++k) h[k] = 0.0;
  }
  steps[m] = 0;
  if (steps32) steps32[m] = 0;
  de_prev[m] = INFINITY;
  active[m] = max_steps > 0 ? 1 : 0;
  converged[m] = 0;
  status[m] = 0;
  prev_prec[m] = prec[m];
  (void)policy;
}

__global__ void k_set_prec(int n_mol, int v, int* __restrict__ prec, const int* __restrict__ mask) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < n_mol && (!mask || mask[m])) prec[m] = v;
}

// Bookkeeping after one SCF step (scfcalculation.cpp:82-146): converged when
// |E_k - E_(k-1)| < threshold (sameReal, mymath.h:68-75); a failed eigensolve
// stops the molecule with its previous energy.  Mixed precision: a step run
// with FP32 contractions never counts as converged; the molecule switches to
// FP64 once |dE| < switch_ev (or after fp32_cap FP32 steps), and convergence
// is tested only between two FP64 steps.
__global__ void k_scf_update(int n_mol, const int* __restrict__ sel, double threshold,
                             int max_steps, double switch_ev, int fp32_cap,
                             const double* __restrict__ e_new,
                             double* __restrict__ e_el, int* __restrict__ steps,
                             int* __restrict__ active, int* __restrict__ converged,
                             const int* __restrict__ status, int* __restrict__ n_active,
                             int* __restrict__ prec, int* __restrict__ prev_prec,
                             int* __restrict__ steps32, double* __restrict__ de_prev,
                             double* __restrict__ e_hist, int e_hist_cap) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_mol) return;
  const int m = sel ? sel[i] : i;
  if (!active[m]) return;
  if (status[m]) {
    active[m] = 0;
    return;
  }
  const double old = e_el[m], e = e_new[m];
  e_el[m] = e;
  const int st = steps[m] + 1;
  steps[m] = st;
  if (e_hist && st < e_hist_cap) e_hist[(int64_t)m * e_hist_cap + st] = e;
  const double de = fabs(old - e);
  const int step_prec = prec[m];
  if (steps32 && step_prec) steps32[m] += 1;
  const int conv = step_prec == 0 && prev_prec[m] == 0 && de < threshold;
  prev_prec[m] = step_prec;
  // Reduced steps end at |dE| < switch_ev, or once |dE| is below what their
  // arithmetic resolves on this energy (kFp32Floor |E_el|, 8 FP32 ulps;
  // kBf16Floor |E_el| for the split-BF16 density), or
  // stops falling by half per step (FP32 rounding, or a slow SCF, sets the
  // pace), or after fp32_cap steps
  const double de_last = de_prev[m];
  de_prev[m] = de;
  const double floor = step_prec == 2 ? kBf16Floor : kFp32Floor;
  if (step_prec && (de < switch_ev || de < floor * fabs(e) || de > 0.5 * de_last ||
                    st >= fp32_cap))
    prec[m] = 0;
  converged[m] = conv;
  const int still = !conv && st < max_steps;
  active[m] = still;
  if (still) atomicAdd(n_active, 1);
}

// Zeroes H and F of the masked molecules, so that a molecule placed in ranges
// an earlier molecule used starts from the same state as in a fresh batch.
__global__ void k_clear_hf(const MolDesc* __restrict__ mols, double* __restrict__ H,
                           double* __restrict__ F, const int* __restrict__ mask) {
  if (!mask[blockIdx.y]) return;
  const MolDesc M = mols[blockIdx.y];
  const int64_t nn = (int64_t)M.n_ao * M.n_ao;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nn;
       t += (int64_t)gridDim.x * blockDim.x) {
    H[M.mat0 + t] = 0.0;
    F[M.mat0 + t] = 0.0;
  }
}

static inline int blocks_for(int64_t n, int t) { return (int)((n + t - 1) / t); }

// Density damping: mode 0 keeps P in C, mode 1 mixes P <- beta C + (1 - beta) P
// (full N x N matrices of the active molecules).
__global__ void k_damp(const MolDesc* __restrict__ mols, double* __restrict__ P,
                       double* __restrict__ C, double beta, int mode,
                       const int* __restrict__ active, const int* __restrict__ status) {
  const int m = blockIdx.y;
  if (!mol_live(active, status, m)) return;
  const MolDesc M = mols[m];
  const int64_t nn = (int64_t)M.n_ao * M.n_ao;
  double* p = P + M.mat0;
  double* c = C + M.mat0;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nn;
       t += (int64_t)gridDim.x * blockDim.x) {
    if (mode == 0) c[t] = p[t];
    else p[t] = beta * c[t] + (1.0 - beta) * p[t];
  }
}

void launch_damp_keep(const BatchDev& b, hipStream_t s) {
  if (b.n_mol) k_damp<<<dim3(16, b.n_mol), 256, 0, s>>>(b.mols, b.P, b.C, b.damping, 0, b.active, b.status);
}

void launch_damp_mix(const BatchDev& b, hipStream_t s) {
  if (b.n_mol) k_damp<<<dim3(16, b.n_mol), 256, 0, s>>>(b.mols, b.P, b.C, b.damping, 1, b.active, b.status);
}

__global__ void k_reset_count(int* __restrict__ n) { *n = 0; }

void launch_reset_count(const BatchDev& b, hipStream_t s) { k_reset_count<<<1, 64, 0, s>>>(b.n_active); }

void launch_clear_hf(const BatchDev& b, hipStream_t s, const int* mask) {
  if (b.n_mol) k_clear_hf<<<dim3(8, b.n_mol), 256, 0, s>>>(b.mols, b.H, b.F, mask);
}

void launch_guess_density(const BatchDev& b, hipStream_t s, const int* mask) {
  k_guess_density<<<b.n_mol, 256, 0, s>>>(b.mols, b.atoms, b.P, mask);
}

void launch_fock(const BatchDev& b, hipStream_t s, bool all, const int* mask) {
  const int* act = mask ? mask : (all ? nullptr : b.active);
  const int* prec = all ? nullptr : b.prec;
  if (b.fock_lds > 0 && b.n_mol > 0) {
    k_fock_mol<double><<<b.n_mol, 64 * kFockMolWaves, b.fock_lds, s>>>(
        b.n_mol, b.order, b.mols, b.atoms, b.partners, b.blocks, b.blocks32, b.H, b.P, b.F,
        b.e_atom, act, b.status, prec);
    if (prec && b.mixed)
      k_fock_mol<float><<<b.n_mol, 64 * kFockMolWaves, b.fock_lds, s>>>(
          b.n_mol, b.order, b.mols, b.atoms, b.partners, b.blocks, b.blocks32, b.H, b.P, b.F,
          b.e_atom, act, b.status, prec);
  }
  if (b.max_n > kFockMolMaxN) {
    k_wdiag<<<blocks_for(b.n_atoms, 256), 256, 0, s>>>(b.n_atoms, b.atoms, b.mols, b.P, b.wdiag,
                                                       act, b.status);
    k_fock<<<blocks_for(b.n_atoms, 4), 256, 0, s>>>(b.n_atoms, b.atoms, b.partners, b.wdiag,
                                                    b.mols, b.blocks, b.blocks32, b.H, b.P, b.F,
                                                    b.e_atom, act, b.status, prec);
  }
}

void set_launch_shapes(BatchDev& d, const MolDesc* mols, int n, const int* used) {
  int64_t lds = 0;
  int rounds = 0;
  for (int m = 0; m < n; ++m) {
    if ((used && !used[m]) || mols[m].n_atoms == 0) continue;
    if (mols[m].n_ao <= kFockMolMaxN)
      lds = std::max(lds, fock_mol_lds(mols[m].n_ao, mols[m].n_atoms));
    rounds = std::max(rounds, core_rounds(mols[m].n_atoms));
  }
  d.fock_lds = (int)lds;
  d.core_rounds = rounds;
}

void launch_energy_fock(const BatchDev& b, hipStream_t s, double* out, bool all, const int* mask) {
  k_energy_atoms<<<blocks_for(b.n_mol, 64), 64, 0, s>>>(b.n_mol, b.sel, b.mols, b.e_atom, out,
                                                        mask ? mask : (all ? nullptr : b.active),
                                                        b.status);
}

void launch_blocks_to_fp32(const BatchDev& b, hipStream_t s) {
  k_blocks_to_fp32<<<2048, 256, 0, s>>>(b.n_blocks, b.blocks, b.blocks32);
}

void launch_energy(const BatchDev& b, hipStream_t s, double* out, bool all, const int* mask) {
  const int* act = mask ? mask : (all ? nullptr : b.active);
  const int n_small = b.n_mol - b.n_large;
  if (n_small > 0)
    k_energy<<<n_small, 256, 0, s>>>(b.mols, b.H, b.F, b.P, out, act, b.status,
                                     b.n_large ? b.order + b.n_large : nullptr);
  // molecules on the multi-workgroup eigensolver: chunks of kEnergyChunk
  // entries per workgroup, partial sums in the molecule's workspace (its T
  // factors are dead by now), then one ordered sum
  for (int i = 0; i < b.n_large; ++i) {
    const MolDesc& M = b.large_mols[i];
    const int m = b.large_idx[i];
    const int64_t nn = (int64_t)M.n_ao * M.n_ao;
    const int chunks = (int)((nn + kEnergyChunk - 1) / kEnergyChunk);
    LargeWork L(b.W, M);
    k_energy_chunks<<<chunks, 256, 0, s>>>(b.H + M.mat0, b.F + M.mat0, b.P + M.mat0, nn, L.T, act,
                                           b.status, m);
    k_energy_total<<<1, 256, 0, s>>>(L.T, chunks, out, act, b.status, m);
  }
}

void launch_scf_init(const BatchDev& b, hipStream_t s, int max_steps, int policy,
                     const int* mask) {
  k_scf_init<<<blocks_for(b.n_mol, 256), 256, 0, s>>>(b.n_mol, max_steps, policy, b.e_new, b.e_el,
                                                      b.steps, b.active, b.converged, b.status,
                                                      b.n_active, b.prec, b.prev_prec, b.steps32,
                                                      b.de_prev, mask, b.e_hist, b.e_hist_cap);
}

void launch_set_prec(const BatchDev& b, hipStream_t s, int v, const int* mask) {
  k_set_prec<<<blocks_for(b.n_mol, 256), 256, 0, s>>>(b.n_mol, v, b.prec, mask);
}

void launch_scf_update(const BatchDev& b, hipStream_t s, double threshold, int max_steps,
                       double switch_ev, int fp32_cap) {
  k_scf_update<<<blocks_for(b.n_mol, 256), 256, 0, s>>>(
      b.n_mol, b.sel, threshold, max_steps, switch_ev, fp32_cap, b.e_new, b.e_el, b.steps,
      b.active, b.converged, b.status, b.n_active, b.prec, b.prev_prec, b.steps32, b.de_prev,
      b.e_hist, b.e_hist_cap);
}

}  // namespace sg
