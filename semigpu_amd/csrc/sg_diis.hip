// sg_diis.hip -- Pulay DIIS extrapolation of the Fock matrix (opt-in SCF
// convergence aid; the reference's SCF, scfcalculation.cpp:82-146, takes each
// new Fock matrix as it is).
//
// Per molecule and SCF step, after the convergence test and only for the
// molecules that go on iterating:
//   1. k_diis_err   error matrix e = F P - P F (the commutator; the ZDO basis
//                   has S = 1), written with F into ring slot k mod m;
//                   16 x 16 output tiles of the lower triangle, e_ji = -e_ij;
//   2. k_diis_dot   <e_k, e_j> for the stored slots, 16 chunks per molecule,
//                   one partial sum per chunk and slot (no atomics: the sums
//                   run in a fixed order, so a molecule's result does not
//                   depend on its batch);
//   3. k_diis_solve the (n+1) x (n+1) system [B -1; -1 0] [c; l] = [0; -1],
//                   B scaled by its largest diagonal entry, partial pivoting;
//                   a singular or non-finite system takes the plain step
//                   (c = 1 on the newest slot);
//   4. k_diis_mix   F <- sum_j c_j F_j.
// The matrices are the full N x N row-major ones of the batch; history slot j
// of a molecule is at hist + j * n_mat + mat0.
#include "sg_kernels.h"

namespace sg {

namespace {

constexpr int kChunks = 16;

__device__ __forceinline__ int diis_slot(const int* k, int m, int depth) { return k[m] % depth; }

__global__ void __launch_bounds__(256) k_diis_err(const MolDesc* __restrict__ mols,
                                                  const double* __restrict__ F,
                                                  const double* __restrict__ P,
                                                  double* __restrict__ histF,
                                                  double* __restrict__ histE, int64_t n_mat,
                                                  const int* __restrict__ kcount, int depth,
                                                  const int* __restrict__ active,
                                                  const int* __restrict__ status,
                                                  const int* __restrict__ fp32_step) {
  const int m = blockIdx.y;
  if (!mol_live(active, status, m) || fp32_step[m]) return;
  const MolDesc M = mols[m];
  const int n = M.n_ao, nt = (n + 15) / 16;
  // tile t of the lower triangle: row tile ti, column tile tj <= ti
  const int t = blockIdx.x;
  if (t >= nt * (nt + 1) / 2) return;
  int ti = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while (ti * (ti + 1) / 2 > t) --ti;
  while ((ti + 1) * (ti + 2) / 2 <= t) ++ti;
  const int tj = t - ti * (ti + 1) / 2;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int i = ti * 16 + ty, j = tj * 16 + tx;
  const double* f = F + M.mat0;
  const double* p = P + M.mat0;
  __shared__ double Fi[16][17], Pi[16][17], Fj[16][17], Pj[16][17];
  double acc = 0.0;
  for (int k0 = 0; k0 < n; k0 += 16) {
    const int k = k0 + tx;
    const int ri = ti * 16 + ty, rj = tj * 16 + ty;
    const bool ki = ri < n && k < n, kj = rj < n && k < n;
    Fi[ty][tx] = ki ? f[(int64_t)ri * n + k] : 0.0;
    Pi[ty][tx] = ki ? p[(int64_t)ri * n + k] : 0.0;
    Fj[ty][tx] = kj ? f[(int64_t)rj * n + k] : 0.0;
    Pj[ty][tx] = kj ? p[(int64_t)rj * n + k] : 0.0;
    __syncthreads();
    // e_ij = sum_k F_ik P_kj - P_ik F_kj, with P_kj = P_jk and F_kj = F_jk
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      acc = fma(Fi[ty][q], Pj[tx][q], acc);
      acc = fma(-Pi[ty][q], Fj[tx][q], acc);
    }
    __syncthreads();
  }
  if (i >= n || j >= n || j > i) return;
  const int s = diis_slot(kcount, m, depth);
  double* e = histE + (int64_t)s * n_mat + M.mat0;
  double* h = histF + (int64_t)s * n_mat + M.mat0;
  const int64_t ij = (int64_t)i * n + j, ji = (int64_t)j * n + i;
  // one writer per entry: the thread of the lower-triangle position writes
  // both halves (the diagonal of a commutator of symmetric matrices is 0)
  e[ij] = i == j ? 0.0 : acc;
  e[ji] = i == j ? 0.0 : -acc;
  h[ij] = f[ij];
  h[ji] = f[ji];
}

// part[(m * kChunks + chunk) * kDiisMax + j] = chunk's share of <e_s, e_j>.
__global__ void __launch_bounds__(256) k_diis_dot(const MolDesc* __restrict__ mols,
                                                  const double* __restrict__ histE, int64_t n_mat,
                                                  const int* __restrict__ kcount, int depth,
                                                  double* __restrict__ part,
                                                  const int* __restrict__ active,
                                                  const int* __restrict__ status,
                                                  const int* __restrict__ fp32_step) {
  const int m = blockIdx.y;
  if (!mol_live(active, status, m) || fp32_step[m]) return;
  const MolDesc M = mols[m];
  const int64_t nn = (int64_t)M.n_ao * M.n_ao;
  const int s = diis_slot(kcount, m, depth);
  const int nv = min(kcount[m] + 1, depth);
  const double* es = histE + (int64_t)s * n_mat + M.mat0;
  double acc[kDiisMax];
#pragma unroll
  for (int j = 0; j < kDiisMax; ++j) acc[j] = 0.0;
  const int64_t per = (nn + kChunks - 1) / kChunks;
  const int64_t lo = blockIdx.x * per, hi = min(nn, lo + per);
  for (int64_t t = lo + threadIdx.x; t < hi; t += blockDim.x) {
    const double a = es[t];
#pragma unroll
    for (int j = 0; j < kDiisMax; ++j)
      if (j < nv) acc[j] = fma(a, histE[(int64_t)j * n_mat + M.mat0 + t], acc[j]);
  }
  __shared__ double red[kDiisMax][256];
#pragma unroll
  for (int j = 0; j < kDiisMax; ++j) red[j][threadIdx.x] = acc[j];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int j = 0; j < kDiisMax; ++j) red[j][threadIdx.x] += red[j][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < kDiisMax)
    part[((int64_t)m * kChunks + blockIdx.x) * kDiisMax + threadIdx.x] = red[threadIdx.x][0];
}

// One thread per molecule: B update, the bordered system, the coefficients.
__global__ void k_diis_solve(int n_mol, int* __restrict__ kcount, int depth,
                             const double* __restrict__ part, double* __restrict__ B,
                             double* __restrict__ coef, const int* __restrict__ active,
                             const int* __restrict__ status, const int* __restrict__ fp32_step) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n_mol || !mol_live(active, status, m)) return;
  if (fp32_step[m]) {  // history restarts at the first FP64 step
    kcount[m] = 0;
    return;
  }
  const int kc = kcount[m];
  const int s = kc % depth, nv = min(kc + 1, depth);
  double* b = B + (int64_t)m * kDiisMax * kDiisMax;
  for (int j = 0; j < nv; ++j) {
    double v = 0.0;
    for (int c = 0; c < kChunks; ++c) v += part[((int64_t)m * kChunks + c) * kDiisMax + j];
    b[s * kDiisMax + j] = v;
    b[j * kDiisMax + s] = v;
  }
  double* c = coef + (int64_t)m * kDiisMax;
  for (int j = 0; j < kDiisMax; ++j) c[j] = 0.0;
  bool ok = nv > 1;
  if (ok) {
    double scale = 0.0;
    for (int j = 0; j < nv; ++j) scale = fmax(scale, b[j * kDiisMax + j]);
    ok = scale > 0.0 && isfinite(scale);
    constexpr int W = kDiisMax + 2;
    double A[kDiisMax + 1][W];
    const int r = nv + 1;
    if (ok) {
      for (int i = 0; i < nv; ++i) {
        for (int j = 0; j < nv; ++j) A[i][j] = b[i * kDiisMax + j] / scale;
        A[i][nv] = -1.0;
        A[i][r] = 0.0;
      }
      for (int j = 0; j < nv; ++j) A[nv][j] = -1.0;
      A[nv][nv] = 0.0;
      A[nv][r] = -1.0;
      for (int col = 0; col < r && ok; ++col) {
        int piv = col;
        for (int i = col + 1; i < r; ++i)
          if (fabs(A[i][col]) > fabs(A[piv][col])) piv = i;
        if (!(fabs(A[piv][col]) > 1e-14)) {
          ok = false;
          break;
        }
        if (piv != col)
          for (int j = 0; j <= r; ++j) {
            const double tmp = A[col][j];
            A[col][j] = A[piv][j];
            A[piv][j] = tmp;
          }
        for (int i = col + 1; i < r; ++i) {
          const double l = A[i][col] / A[col][col];
          for (int j = col; j <= r; ++j) A[i][j] -= l * A[col][j];
        }
      }
      if (ok) {
        double x[kDiisMax + 1];
        for (int i = r - 1; i >= 0; --i) {
          double v = A[i][r];
          for (int j = i + 1; j < r; ++j) v -= A[i][j] * x[j];
          x[i] = v / A[i][i];
        }
        for (int j = 0; j < nv; ++j) ok = ok && isfinite(x[j]);
        if (ok)
          for (int j = 0; j < nv; ++j) c[j] = x[j];
      }
    }
  }
  if (!ok) {
    for (int j = 0; j < kDiisMax; ++j) c[j] = 0.0;
    c[s] = 1.0;
  }
  kcount[m] = kc + 1;
}

__global__ void __launch_bounds__(256) k_diis_mix(const MolDesc* __restrict__ mols,
                                                  double* __restrict__ F,
                                                  const double* __restrict__ histF, int64_t n_mat,
                                                  const double* __restrict__ coef,
                                                  const int* __restrict__ active,
                                                  const int* __restrict__ status,
                                                  const int* __restrict__ fp32_step) {
  const int m = blockIdx.y;
  if (!mol_live(active, status, m) || fp32_step[m]) return;
  const MolDesc M = mols[m];
  const int64_t nn = (int64_t)M.n_ao * M.n_ao;
  double c[kDiisMax];
#pragma unroll
  for (int j = 0; j < kDiisMax; ++j) c[j] = coef[(int64_t)m * kDiisMax + j];
  double* f = F + M.mat0;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nn;
       t += (int64_t)gridDim.x * blockDim.x) {
    double v = 0.0;
#pragma unroll
    for (int j = 0; j < kDiisMax; ++j)  // unused slots hold c = 0 and undefined matrices
      if (c[j] != 0.0) v = fma(c[j], histF[(int64_t)j * n_mat + M.mat0 + t], v);
    f[t] = v;
  }
}

__global__ void k_diis_reset(int n_mol, int* __restrict__ kcount) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < n_mol) kcount[m] = 0;
}

}  // namespace

void launch_diis_reset(const BatchDev& b, hipStream_t s) {
  if (b.n_mol && b.diis_m) k_diis_reset<<<(b.n_mol + 255) / 256, 256, 0, s>>>(b.n_mol, b.diis_k);
}

void launch_diis(const BatchDev& b, hipStream_t s) {
  if (!b.n_mol || !b.diis_m) return;
  const int nt = (b.max_n + 15) / 16;
  k_diis_err<<<dim3(nt * (nt + 1) / 2, b.n_mol), 256, 0, s>>>(
      b.mols, b.F, b.P, b.diis_F, b.diis_E, b.n_mat, b.diis_k, b.diis_m, b.active, b.status, b.prev_prec);
  k_diis_dot<<<dim3(kChunks, b.n_mol), 256, 0, s>>>(b.mols, b.diis_E, b.n_mat, b.diis_k, b.diis_m,
                                                    b.diis_part, b.active, b.status, b.prev_prec);
  k_diis_solve<<<(b.n_mol + 63) / 64, 64, 0, s>>>(b.n_mol, b.diis_k, b.diis_m, b.diis_part,
                                                  b.diis_B, b.diis_c, b.active, b.status,
                                                  b.prev_prec);
  k_diis_mix<<<dim3(kChunks, b.n_mol), 256, 0, s>>>(b.mols, b.F, b.diis_F, b.n_mat, b.diis_c,
                                                    b.active, b.status, b.prev_prec);
}

int64_t diis_part_doubles(int n_mol) { return (int64_t)n_mol * kChunks * kDiisMax; }

}  // namespace sg
