// sg_integrals.hip -- two-center multipole integrals, STO-6G overlaps, Hcore and
// core-core repulsion for a batch of molecules.
//
// This file is compiled with -ffp-contract=off: every expression keeps the
// operation order of the reference CPU code (no fused multiply-adds), so the
// pair blocks and Hcore reproduce the reference's FP64 values bit for bit
// where the math library agrees (sqrt and division are correctly rounded on
// gfx950; exp may differ in the last bit, which only touches overlaps).
//
// Kernels
//   k_pair_xx / k_pair_hx / k_pair_hh   one lane per atom pair (type-sorted lists):
//        the 21 distinct local-frame multipole interactions of the pair are
//        evaluated once (multipole.cpp:20-498), the local frame
//        (twocenterintegral.cpp:468-490) is built once, and the 100 / 10 / 1
//        rotated entries (twocenterintegral.cpp:515-736) are produced by
//        compile-time-unrolled formulas.  The reference recomputes the local
//        multipoles for each of the 100 entries (:388-466).
//   k_one_center                        one lane per atom: one-center blocks from
//        the G parameters (twocenterintegral.cpp:738-778).
//   k_hcore_offdiag                     one lane per atom pair: beta * S / 2 blocks
//        (hcore.cpp:38-69) with the STO-6G overlap (overlap.cpp:68-192).
//   k_hcore_diag                        one lane per (atom, AO pair): U + core
//        attraction summed over the other atoms in atom order (hcore.cpp:71-99).
//   k_core_core                         one workgroup per molecule: pair terms
//        evaluated in parallel, summed in the reference order
//        (corecorerepulsion.cpp:12-65).
#include "sg_common.h"
#include "sg_kernels.h"
#include "sg_pair_math.h"

namespace sg {

constexpr int kCoreChunk = 2048;  // core-core pair terms per round of k_core_core
// Molecules with more atom pairs than this sum their core-core terms in a fixed
// tree instead of the reference's serial order (one lane adding 4.5 M terms
// of the 1000-water box took 117 ms).
constexpr int64_t kCoreSerialMax = 1 << 17;

template <int E>
struct XXFill {
  __device__ static __forceinline__ void run(const LocalXX& v, const Frame& F, double* out,
                                             float* out32) {
    XXFill<E - 1>::run(v, F, out, out32);
    const double x = xx_entry<(E - 1) / 10, (E - 1) % 10>(v, F);
    out[E - 1] = x;
    if (out32) out32[E - 1] = (float)x;
  }
};
template <>
struct XXFill<0> {
  __device__ static __forceinline__ void run(const LocalXX&, const Frame&, double*, float*) {}
};

// ---------------------------------------------------------------------------
// Pair kernels.  pairs[i] = (row atom, column atom) with row local index > column.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) k_pair_xx(const int2* __restrict__ pairs, int n,
                                                 const AtomDesc* __restrict__ atoms,
                                                 const MolDesc* __restrict__ mols,
                                                 double* __restrict__ blocks,
                                                 float* __restrict__ blocks32) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int2 pr = pairs[i];
  const AtomDesc A = atoms[pr.x], C = atoms[pr.y];
  const double R = dist3(A.x, A.y, A.z, C.x, C.y, C.z);
  const Frame F = local_frame(A, C, R);
  const LocalXX v = local_xx(R, A.zn, C.zn);
  const int64_t o = mols[A.mol].blk0 + block_offset(A, C);
  XXFill<100>::run(v, F, blocks + o, blocks32 ? blocks32 + o : nullptr);
}

// (H H | X pair) blocks: the H orbital is always the first pair and the frame
// runs from H to X (ComputePair_HX / ComputePair_XH, :397-442).
__global__ void __launch_bounds__(256) k_pair_hx(const int2* __restrict__ pairs, int n,
                                                  const AtomDesc* __restrict__ atoms,
                                                  const MolDesc* __restrict__ mols,
                                                  double* __restrict__ blocks,
                                                  float* __restrict__ blocks32) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int2 pr = pairs[i];
  const AtomDesc Ar = atoms[pr.x], Ac = atoms[pr.y];
  const AtomDesc& H = Ar.zn == 1 ? Ar : Ac;
  const AtomDesc& X = Ar.zn == 1 ? Ac : Ar;
  const double R = dist3(H.x, H.y, H.z, X.x, X.y, X.z);
  const Frame F = local_frame(H, X, R);
  const double(&m)[3][3] = F.m;
  const int h = H.zn, x = X.zn;
  const double qq = mp_qq(R, h, x);
  const double qUz = mp_qUz(R, h, x);
  const double vpi = qq + mp_qQpipi(R, h, x);
  const double vz = qq + mp_qQzz(R, h, x);
  const int64_t o = mols[Ar.mol].blk0 + block_offset(Ar, Ac);
  double* out = blocks + o;
  float* o32 = blocks32 ? blocks32 + o : nullptr;
  out[0] = qq * kHartreeToEv;
  if (o32) o32[0] = (float)out[0];
#pragma unroll
  for (int k = 1; k < 10; ++k) {
    const int t2 = mop_a(k), t3 = mop_b(k);
    double r;
    if (t2 == 0) {
      r = qUz * m[2][t3 - 1];
    } else {
      double a = vpi * m[0][t2 - 1] * m[0][t3 - 1] + vpi * m[1][t2 - 1] * m[1][t3 - 1];
      double b = vz * m[2][t2 - 1] * m[2][t3 - 1];
      r = a + b;
    }
    out[k] = r * kHartreeToEv;
    if (o32) o32[k] = (float)out[k];
  }
}

__global__ void __launch_bounds__(256) k_pair_hh(const int2* __restrict__ pairs, int n,
                                                  const AtomDesc* __restrict__ atoms,
                                                  const MolDesc* __restrict__ mols,
                                                  double* __restrict__ blocks,
                                                  float* __restrict__ blocks32) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int2 pr = pairs[i];
  const AtomDesc A = atoms[pr.x], C = atoms[pr.y];
  const double R = dist3(A.x, A.y, A.z, C.x, C.y, C.z);
  const int64_t o = mols[A.mol].blk0 + block_offset(A, C);
  const double v = mp_qq(R, A.zn, C.zn) * kHartreeToEv;
  blocks[o] = v;
  if (blocks32) blocks32[o] = (float)v;
}

// One-center blocks (SelfIntegral*, :738-778); zero where the reference
// returns 0 (mixed SS|SP, SP|PP, ...).
__device__ double one_center(int z, int s0, int s1, int s2, int s3) {
  const int pa = (s0 > 0) + (s1 > 0), pb = (s2 > 0) + (s3 > 0);
  if (pa == pb) {
    if (pa == 0) return c_par.gss[z];
    if (pa == 1) return s1 == s3 ? c_par.hsp[z] : 0.0;
    if (s0 == s1 && s2 == s3) return s0 == s2 ? c_par.gpp[z] : c_par.gp2[z];
    if (s0 == s2 && s1 == s3) return 0.5 * (c_par.gpp[z] - c_par.gp2[z]);
    return 0.0;
  }
  if (pa == 0 && pb == 2) return s2 == s3 ? c_par.gsp[z] : 0.0;
  if (pa == 2 && pb == 0) return s0 == s1 ? c_par.gsp[z] : 0.0;
  return 0.0;
}

__global__ void k_one_center(int n_atoms, const AtomDesc* __restrict__ atoms,
                             const MolDesc* __restrict__ mols, double* __restrict__ blocks,
                             float* __restrict__ blocks32, const int* __restrict__ mask) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_atoms) return;
  const AtomDesc A = atoms[i];
  if (mask && !mask[A.mol]) return;
  const int64_t o = mols[A.mol].blk0 + block_offset(A, A);
  double* out = blocks + o;
  const int cnt = A.zn == 1 ? 1 : 100;
  for (int t = 0; t < cnt; ++t) {
    const int k = t / 10, l = t % 10;
    const double v = A.zn == 1 ? c_par.gss[1]
                               : one_center(A.zn, mop_a(k), mop_b(k), mop_a(l), mop_b(l));
    out[t] = v;
    if (blocks32) blocks32[o + t] = (float)v;
  }
}

__device__ __forceinline__ void hcore_store(const AtomDesc& A, int u, const AtomDesc& B, int v,
                                            double s, int N, double* h) {
  const double bu = u == 0 ? c_par.betaS[A.zn] : c_par.betaP[A.zn];
  const double bv = v == 0 ? c_par.betaS[B.zn] : c_par.betaP[B.zn];
  const double hv = bu + bv;
  const double val = hv * s * 0.5;
  const int r = A.ao0 + u, c = B.ao0 + v;
  h[(int64_t)r * N + c] = val;
  h[(int64_t)c * N + r] = val;
}

// XX: both heavy; HX: one of the two is hydrogen (either order); HH.
template <int KIND>
__global__ void __launch_bounds__(128) k_hcore_offdiag(const int2* __restrict__ pairs, int n,
                                                        const AtomDesc* __restrict__ atoms,
                                                        const MolDesc* __restrict__ mols,
                                                        double* __restrict__ H) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int2 pr = pairs[i];
  AtomDesc A = atoms[pr.x], B = atoms[pr.y];
  const MolDesc M = mols[A.mol];
  const int N = M.n_ao;
  double* h = H + M.mat0;
  double ab = 0.0;
  {
    const double d0 = A.x - B.x, d1 = A.y - B.y, d2 = A.z - B.z;
    ab += d0 * d0;
    ab += d1 * d1;
    ab += d2 * d2;
  }
  if (KIND == 2) {  // H-H
    double t[1] = {0.0};
    hcore_shell<1, 1>(A, 0, B, 0, ab, t);
    hcore_store(A, 0, B, 0, t[0], N, h);
  } else if (KIND == 1) {  // H-X: A is the heavy atom after the swap
    const bool swap = A.zn == 1;
    if (swap) {
      const AtomDesc T = A;
      A = B;
      B = T;
    }
    // overlaps are taken as S(AO of the pair's first atom, AO of its second)
    double ss[1] = {0.0}, ps[3] = {0.0, 0.0, 0.0};
    if (!swap) {
      hcore_shell<1, 1>(A, 0, B, 0, ab, ss);
      hcore_shell<3, 1>(A, 1, B, 0, ab, ps);
    } else {
      // first atom is the hydrogen: S(H s, X ao), the reference's argument order
      double sp[3] = {0.0, 0.0, 0.0};
      hcore_shell<1, 1>(B, 0, A, 0, ab, ss);
      hcore_shell<1, 3>(B, 0, A, 1, ab, sp);
      ps[0] = sp[0];
      ps[1] = sp[1];
      ps[2] = sp[2];
    }
    hcore_store(A, 0, B, 0, ss[0], N, h);
    for (int u = 0; u < 3; ++u) hcore_store(A, u + 1, B, 0, ps[u], N, h);
  } else {  // X-X
    double ss[1] = {0.0}, sp[3] = {0.0, 0.0, 0.0}, ps[3] = {0.0, 0.0, 0.0};
    double pp[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    hcore_shell<1, 1>(A, 0, B, 0, ab, ss);
    hcore_shell<1, 3>(A, 0, B, 1, ab, sp);
    hcore_shell<3, 1>(A, 1, B, 0, ab, ps);
    hcore_shell<3, 3>(A, 1, B, 1, ab, pp);
    hcore_store(A, 0, B, 0, ss[0], N, h);
    for (int v = 0; v < 3; ++v) hcore_store(A, 0, B, v + 1, sp[v], N, h);
    for (int u = 0; u < 3; ++u) hcore_store(A, u + 1, B, 0, ps[u], N, h);
    for (int u = 0; u < 3; ++u)
      for (int v = 0; v < 3; ++v) hcore_store(A, u + 1, B, v + 1, pp[u * 3 + v], N, h);
  }
}

// (mu nu | s_B s_B) with mu, nu on A (GetValueFromArray, :285-386).
__device__ __forceinline__ double value_uv_ss(const double* blk_mol, const AtomDesc& A,
                                              const AtomDesc& B, int slot) {
  if (A.local > B.local) {
    const double* b = blk_mol + block_offset(A, B);
    if (A.zn > 1) return B.zn > 1 ? b[slot * 10] : b[slot];
    return b[0];
  }
  const double* b = blk_mol + block_offset(B, A);
  if (A.zn > 1) return B.zn > 1 ? b[slot] : b[slot];
  return b[0];
}

// Diagonal Hcore blocks: tasks = (atom, slot) with slot < 10 for heavy atoms.
__global__ void k_hcore_diag(const int2* __restrict__ tasks, int n,
                             const AtomDesc* __restrict__ atoms, const MolDesc* __restrict__ mols,
                             const double* __restrict__ blocks, double* __restrict__ H) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int2 t = tasks[i];
  const AtomDesc A = atoms[t.x];
  const MolDesc M = mols[A.mol];
  const int slot = t.y;
  const int u = mop_a(slot), v = mop_b(slot);
  const double* blk = blocks + M.blk0;
  double s = 0.0;
  for (int b = 0; b < M.n_atoms; ++b) {
    if (b == A.local) continue;
    const AtomDesc B = atoms[M.atom0 + b];
    s += -(double)core_charge(B.zn) * value_uv_ss(blk, A, B, slot);
  }
  double val;
  if (u == v) {
    val = (u == 0 ? c_par.uss[A.zn] : c_par.upp[A.zn]) + s;
  } else {
    val = s;
  }
  const int N = M.n_ao;
  double* h = H + M.mat0;
  h[(int64_t)(A.ao0 + u) * N + (A.ao0 + v)] = val;
  h[(int64_t)(A.ao0 + v) * N + (A.ao0 + u)] = val;
}

// Core-core repulsion per molecule (CoreCoreRepulsion::ComputeRepulsion).
// Core-core term of atoms i < j (corecorerepulsion.cpp:12-65).
__device__ __forceinline__ double core_pair(const AtomDesc& A, const AtomDesc& B,
                                            const double* __restrict__ blk) {
  double t = (double)(core_charge(A.zn) * core_charge(B.zn));
  t *= blk[block_offset(B, A)];
  double r = dist3(A.x, A.y, A.z, B.x, B.y, B.z);
  r = r * kBohrToAng;
  double f;
  if (B.zn == 1 && (A.zn == 7 || A.zn == 8))
    f = 1.0 + r * exp(-c_par.alpha[A.zn] * r) + exp(-c_par.alpha[B.zn] * r);
  else if (A.zn == 1 && (B.zn == 7 || B.zn == 8))
    f = 1.0 + r * exp(-c_par.alpha[B.zn] * r) + exp(-c_par.alpha[A.zn] * r);
  else
    f = 1.0 + exp(-c_par.alpha[A.zn] * r) + exp(-c_par.alpha[B.zn] * r);
  return t * f;
}

// Pair (i, j > i) of index p in the reference order (row-major over i < j).
__device__ __forceinline__ void pair_of(int64_t p, int na, int& i, int& j) {
  // row i: the largest i with i (2 na - i - 1) / 2 <= p
  const double b = 2.0 * na - 1.0;
  i = (int)((b - sqrt(b * b - 8.0 * (double)p)) * 0.5);
  i = i < 0 ? 0 : (i > na - 2 ? na - 2 : i);
  while (i > 0 && (int64_t)i * (2 * na - i - 1) / 2 > p) --i;
  while ((int64_t)(i + 1) * (2 * na - i - 2) / 2 <= p) ++i;
  j = (int)(p - (int64_t)i * (2 * na - i - 1) / 2) + i + 1;
}

// One workgroup per molecule.  The pairs are taken in the reference order in
// rounds of kCoreChunk terms: waves 1-3 evaluate round r + 1 into one LDS
// buffer while lane 0 of wave 0 adds round r from the other to the running
// sum, in order, so the sum is the reference's serial sum bit for bit.
// Molecules above kCoreSerialMax pairs take k_core_rounds and k_core_sum
// instead.
__global__ void __launch_bounds__(256) k_core_core(const AtomDesc* __restrict__ atoms,
                                                    const MolDesc* __restrict__ mols,
                                                    const double* __restrict__ blocks,
                                                    double* __restrict__ e_cc,
                                                    const int* __restrict__ mask) {
  const int m = blockIdx.x;
  if (mask && !mask[m]) return;
  const MolDesc M = mols[m];
  const int na = M.n_atoms;
  const int64_t npair = (int64_t)na * (na - 1) / 2;
  const double* blk = blocks + M.blk0;
  const AtomDesc* at = atoms + M.atom0;
  __shared__ double term[2][kCoreChunk];
  const int nr = (int)((npair + kCoreChunk - 1) / kCoreChunk);
  if (npair > kCoreSerialMax) return;  // k_core_rounds and k_core_sum
  auto fill = [&](int r, int t0, int nt) {
    const int64_t p0 = (int64_t)r * kCoreChunk;
    const int cnt = (int)(npair - p0 < kCoreChunk ? npair - p0 : kCoreChunk);
    for (int q = t0; q < cnt; q += nt) {
      int i, j;
      pair_of(p0 + q, na, i, j);
      term[r & 1][q] = core_pair(at[i], at[j], blk);
    }
  };
  if (nr > 0) fill(0, threadIdx.x, blockDim.x);
  __syncthreads();
  double e = 0.0;
  for (int r = 0; r < nr; ++r) {
    if (threadIdx.x >= 64) {
      if (r + 1 < nr) fill(r + 1, threadIdx.x - 64, blockDim.x - 64);
    } else if (threadIdx.x == 0) {
      const int64_t p0 = (int64_t)r * kCoreChunk;
      const int cnt = (int)(npair - p0 < kCoreChunk ? npair - p0 : kCoreChunk);
      const double* tb = term[r & 1];
      double v[8], w[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = u < cnt ? tb[u] : 0.0;
      for (int q0 = 0; q0 < cnt; q0 += 8) {
#pragma unroll
        for (int u = 0; u < 8; ++u) w[u] = q0 + 8 + u < cnt ? tb[q0 + 8 + u] : 0.0;
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (q0 + u < cnt) e += v[u];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = w[u];
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) e_cc[m] = e;
}

// Molecules above kCoreSerialMax atom pairs: one workgroup per round of
// kCoreChunk terms (grid: rounds of the largest such molecule x molecules).
// Every thread evaluates and sums its share of the round (terms t, t + 256, ...
// in order), the 256 shares are added by a fixed tree, and k_core_sum adds
// the rounds in order: the same sum for the molecule wherever it is run, with
// a smaller rounding error than the serial one.  The round sums wait in the
// molecule's eigensolver workspace (unused before the SCF).
__global__ void __launch_bounds__(256) k_core_rounds(const AtomDesc* __restrict__ atoms,
                                                      const MolDesc* __restrict__ mols,
                                                      const double* __restrict__ blocks,
                                                      double* __restrict__ work,
                                                      const int* __restrict__ mask) {
  const int m = blockIdx.y, r = blockIdx.x;
  if (mask && !mask[m]) return;
  const MolDesc M = mols[m];
  const int na = M.n_atoms;
  const int64_t npair = (int64_t)na * (na - 1) / 2;
  if (npair <= kCoreSerialMax || (int64_t)r * kCoreChunk >= npair) return;
  const double* blk = blocks + M.blk0;
  const AtomDesc* at = atoms + M.atom0;
  __shared__ double sh[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t p0 = (int64_t)r * kCoreChunk;
  const int cnt = (int)(npair - p0 < kCoreChunk ? npair - p0 : kCoreChunk);
  double part = 0.0;
  for (int q = threadIdx.x; q < cnt; q += blockDim.x) {
    int i, j;
    pair_of(p0 + q, na, i, j);
    part += core_pair(at[i], at[j], blk);
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) part += __shfl_xor(part, o, 64);
  if (lane == 0) sh[wave] = part;
  __syncthreads();
  if (threadIdx.x == 0) work[5 * M.mat0 + r] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

__global__ void k_core_sum(int n_mol, const MolDesc* __restrict__ mols,
                           const double* __restrict__ work, double* __restrict__ e_cc,
                           const int* __restrict__ mask) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n_mol || (mask && !mask[m])) return;
  const MolDesc M = mols[m];
  const int64_t npair = (int64_t)M.n_atoms * (M.n_atoms - 1) / 2;
  if (npair <= kCoreSerialMax) return;
  const int nr = (int)((npair + kCoreChunk - 1) / kCoreChunk);
  const double* part = work + 5 * M.mat0;
  double e = 0.0;
  for (int r = 0; r < nr; ++r) e += part[r];
  e_cc[m] = e;
}

int core_rounds(int n_atoms) {
  const int64_t npair = (int64_t)n_atoms * (n_atoms - 1) / 2;
  return npair > kCoreSerialMax ? (int)((npair + kCoreChunk - 1) / kCoreChunk) : 0;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
static inline int blocks_for(int64_t n, int t) { return (int)((n + t - 1) / t); }

void launch_pair_integrals(const BatchDev& b, hipStream_t s, const int* mask) {
  if (b.n_xx)
    k_pair_xx<<<blocks_for(b.n_xx, 64), 64, 0, s>>>(b.pairs_xx, b.n_xx, b.atoms, b.mols, b.blocks,
                                                   b.blocks32);
  if (b.n_hx)
    k_pair_hx<<<blocks_for(b.n_hx, 256), 256, 0, s>>>(b.pairs_hx, b.n_hx, b.atoms, b.mols, b.blocks,
                                                     b.blocks32);
  if (b.n_hh)
    k_pair_hh<<<blocks_for(b.n_hh, 256), 256, 0, s>>>(b.pairs_hh, b.n_hh, b.atoms, b.mols, b.blocks,
                                                     b.blocks32);
  k_one_center<<<blocks_for(b.n_atoms, 128), 128, 0, s>>>(b.n_atoms, b.atoms, b.mols, b.blocks,
                                                          b.blocks32, mask);
}

void launch_hcore(const BatchDev& b, hipStream_t s) {
  if (b.n_xx)
    k_hcore_offdiag<0><<<blocks_for(b.n_xx, 128), 128, 0, s>>>(b.pairs_xx, b.n_xx, b.atoms, b.mols,
                                                               b.H);
  if (b.n_hx)
    k_hcore_offdiag<1><<<blocks_for(b.n_hx, 128), 128, 0, s>>>(b.pairs_hx, b.n_hx, b.atoms, b.mols,
                                                               b.H);
  if (b.n_hh)
    k_hcore_offdiag<2><<<blocks_for(b.n_hh, 128), 128, 0, s>>>(b.pairs_hh, b.n_hh, b.atoms, b.mols,
                                                               b.H);
  k_hcore_diag<<<blocks_for(b.n_diag_tasks, 128), 128, 0, s>>>(b.diag_tasks, b.n_diag_tasks, b.atoms,
                                                               b.mols, b.blocks, b.H);
}

void launch_core_core(const BatchDev& b, hipStream_t s, const int* mask) {
  k_core_core<<<b.n_mol, 256, 0, s>>>(b.atoms, b.mols, b.blocks, b.e_cc, mask);
  if (b.core_rounds > 0 && b.n_mol > 0) {
    k_core_rounds<<<dim3(b.core_rounds, b.n_mol), 256, 0, s>>>(b.atoms, b.mols, b.blocks, b.W,
                                                               mask);
    k_core_sum<<<blocks_for(b.n_mol, 64), 64, 0, s>>>(b.n_mol, b.mols, b.W, b.e_cc, mask);
  }
}

}  // namespace sg
