// sg_diis.hip -- Pulay DIIS extrapolation of the Fock matrix (opt-in SCF
// convergence aid; the reference's SCF, scfcalculation.cpp:82-146, takes each
// new Fock matrix as it is).
//
// Per molecule and SCF step, after the convergence test and only for the
// molecules that go on iterating:
//   1. k_diis_err   error matrix e = F P - P F (the commutator; the ZDO basis
//                   has S = 1), written with F into ring slot k mod m;
//                   64 x 64 blocks of the lower triangle on the FP64 matrix
//                   cores, e_ji = -e_ij;
//   2. k_diis_dot   <e_k, e_j> for the stored slots over the strict lower
//                   triangles (twice their sum), 16 row chunks per molecule,
//                   one partial sum per chunk and slot (no atomics: the sums
//                   run in a fixed order, so a molecule's result does not
//                   depend on its batch);
//   3. k_diis_solve the (n+1) x (n+1) system [B -1; -1 0] [c; l] = [0; -1],
//                   B scaled by its largest diagonal entry, partial pivoting;
//                   a singular or non-finite system takes the plain step
//                   (c = 1 on the newest slot);
//   4. k_diis_mix   F <- sum_j c_j F_j.
// The matrices are the full N x N row-major ones of the batch; history slot j
// of a molecule is at hist + j * stride + mat0 (stride: the batch's sum N^2,
// or the server's matrix pool).
#include "sg_kernels.h"

namespace sg {

namespace {

constexpr int kChunks = 16;

__device__ __forceinline__ int diis_slot(const int* k, int m, int depth) { return k[m] % depth; }

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kErrTile = 64;  // output block of one workgroup (lower triangle of 64 x 64 blocks)
constexpr int kErrK = 16;     // k columns staged through LDS per pass

// Four wavefronts per workgroup, each a 32 x 32 quarter of the block as 2 x 2
// v_mfma_f64_16x16x4f64 tiles (operand lane (r, q) = row r & 15, k column q;
// result lane (c, q) holds rows 4 i + q, column c):
//   e_ij = sum_k F_ik P_jk - P_ik F_jk   (P_kj = P_jk, F_kj = F_jk)
__global__ void __launch_bounds__(256) k_diis_err(const int* __restrict__ sel,
                                                  const MolDesc* __restrict__ mols,
                                                  const double* __restrict__ F,
                                                  const double* __restrict__ P,
                                                  double* __restrict__ histF,
                                                  double* __restrict__ histE, int64_t stride,
                                                  const int* __restrict__ kcount, int depth,
                                                  const int* __restrict__ active,
                                                  const int* __restrict__ status) {
  const int m = sel ? sel[blockIdx.y] : blockIdx.y;
  if (!mol_live(active, status, m)) return;
  const MolDesc M = mols[m];
  const int n = M.n_ao, nt = (n + kErrTile - 1) / kErrTile;
  const int t = blockIdx.x;
  if (t >= nt * (nt + 1) / 2) return;
  int bi = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
  while (bi * (bi + 1) / 2 > t) --bi;
  while ((bi + 1) * (bi + 2) / 2 <= t) ++bi;
  const int bj = t - bi * (bi + 1) / 2;
  const int I0 = bi * kErrTile, J0 = bj * kErrTile;
  const double* f = F + M.mat0;
  const double* p = P + M.mat0;
  __shared__ double sFi[kErrTile][kErrK + 1], sPi[kErrTile][kErrK + 1];
  __shared__ double sFj[kErrTile][kErrK + 1], sPj[kErrTile][kErrK + 1];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int lr = lane & 15, lk = lane >> 4;
  const int wr = wave >> 1, wc = wave & 1;
  v4d acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c) acc[a][c] = v4d{0.0, 0.0, 0.0, 0.0};
  const int lrow = tid >> 4, lcol = tid & 15;  // staging: 16 rows of 16 columns per pass
  for (int k0 = 0; k0 < n; k0 += kErrK) {
    const int k = k0 + lcol;
#pragma unroll
    for (int q = 0; q < kErrTile / 16; ++q) {
      const int rr = lrow + 16 * q, ri = I0 + rr, rj = J0 + rr;
      const bool ki = ri < n && k < n, kj = rj < n && k < n;
      sFi[rr][lcol] = ki ? f[(int64_t)ri * n + k] : 0.0;
      sPi[rr][lcol] = ki ? p[(int64_t)ri * n + k] : 0.0;
      sFj[rr][lcol] = kj ? f[(int64_t)rj * n + k] : 0.0;
      sPj[rr][lcol] = kj ? p[(int64_t)rj * n + k] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int st = 0; st < kErrK / 4; ++st) {
      const int kk = 4 * st + lk;
      double aF[2], aP[2], bF[2], bP[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const int r = 32 * wr + 16 * a + lr, c = 32 * wc + 16 * a + lr;
        aF[a] = sFi[r][kk];
        aP[a] = -sPi[r][kk];
        bP[a] = sPj[c][kk];
        bF[a] = sFj[c][kk];
      }
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          acc[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(aF[a], bP[c], acc[a][c], 0, 0, 0);
          acc[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(aP[a], bF[c], acc[a][c], 0, 0, 0);
        }
    }
    __syncthreads();
  }
  const int s = diis_slot(kcount, m, depth);
  double* e = histE + (int64_t)s * stride + M.mat0;
  double* h = histF + (int64_t)s * stride + M.mat0;
  // one writer per entry: the lane of the lower-triangle position writes both
  // halves (the diagonal of a commutator of symmetric matrices is 0)
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int i = I0 + 32 * wr + 16 * a + 4 * rr + lk, j = J0 + 32 * wc + 16 * c + lr;
        if (i >= n || j >= n || j > i) continue;
        const int64_t ij = (int64_t)i * n + j, ji = (int64_t)j * n + i;
        const double v = acc[a][c][rr];
        e[ij] = i == j ? 0.0 : v;
        e[ji] = i == j ? 0.0 : -v;
        h[ij] = f[ij];
        h[ji] = f[ji];
      }
}

// part[(m * kChunks + chunk) * kDiisMax + j] = chunk's share of <e_s, e_j>.
__global__ void __launch_bounds__(256) k_diis_dot(const int* __restrict__ sel,
                                                  const MolDesc* __restrict__ mols,
                                                  const double* __restrict__ histE, int64_t stride,
                                                  const int* __restrict__ kcount, int depth,
                                                  double* __restrict__ part,
                                                  const int* __restrict__ active,
                                                  const int* __restrict__ status) {
  const int m = sel ? sel[blockIdx.y] : blockIdx.y;
  if (!mol_live(active, status, m)) return;
  const MolDesc M = mols[m];
  const int n = M.n_ao;
  const int s = diis_slot(kcount, m, depth);
  const int nv = min(kcount[m] + 1, depth);
  const double* es = histE + (int64_t)s * stride + M.mat0;
  double acc[kDiisMax];
#pragma unroll
  for (int j = 0; j < kDiisMax; ++j) acc[j] = 0.0;
  // the error matrices are antisymmetric with a zero diagonal: <e_s, e_j> is
  // twice the sum over the strict lower triangle, rows chunk, chunk + 16, ...
  for (int i = blockIdx.x; i < n; i += kChunks)
    for (int c = threadIdx.x; c < i; c += blockDim.x) {
      const int64_t t = (int64_t)i * n + c;
      const double a = es[t];
#pragma unroll
      for (int j = 0; j < kDiisMax; ++j)
        if (j < nv) acc[j] = fma(a, histE[(int64_t)j * stride + M.mat0 + t], acc[j]);
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
    part[((int64_t)m * kChunks + blockIdx.x) * kDiisMax + threadIdx.x] = 2.0 * red[threadIdx.x][0];
}

// One thread per molecule: B update, the bordered system, the coefficients.
__global__ void k_diis_solve(int n_mol, const int* __restrict__ sel, int* __restrict__ kcount,
                             int depth,
                             const double* __restrict__ part, double* __restrict__ B,
                             double* __restrict__ coef, const int* __restrict__ active,
                             const int* __restrict__ status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_mol) return;
  const int m = sel ? sel[i] : i;
  if (!mol_live(active, status, m)) return;
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

__global__ void __launch_bounds__(256) k_diis_mix(const int* __restrict__ sel,
                                                  const MolDesc* __restrict__ mols,
                                                  double* __restrict__ F,
                                                  const double* __restrict__ histF, int64_t stride,
                                                  const double* __restrict__ coef,
                                                  const int* __restrict__ active,
                                                  const int* __restrict__ status) {
  const int m = sel ? sel[blockIdx.y] : blockIdx.y;
  if (!mol_live(active, status, m)) return;
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
      if (c[j] != 0.0) v = fma(c[j], histF[(int64_t)j * stride + M.mat0 + t], v);
    f[t] = v;
  }
}

__global__ void k_diis_reset(int n_mol, int* __restrict__ kcount, int* __restrict__ hist_prec,
                             const int* __restrict__ mask) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < n_mol && (!mask || mask[m])) {
    kcount[m] = 0;
    hist_prec[m] = -1;
  }
}

// The history holds Fock matrices of one precision: it restarts when the
// precision of the step just done differs from the stored ones'.
__global__ void k_diis_begin(int n_mol, const int* __restrict__ sel, int* __restrict__ kcount,
                             int* __restrict__ hist_prec, const int* __restrict__ step_prec,
                             const int* __restrict__ active, const int* __restrict__ status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_mol) return;
  const int m = sel ? sel[i] : i;
  if (!mol_live(active, status, m)) return;
  if (hist_prec[m] != step_prec[m]) {
    kcount[m] = 0;
    hist_prec[m] = step_prec[m];
  }
}

}  // namespace

void launch_diis_reset(const BatchDev& b, hipStream_t s, const int* mask) {
  if (b.n_mol && b.diis_m)
    k_diis_reset<<<(b.n_mol + 255) / 256, 256, 0, s>>>(b.n_mol, b.diis_k, b.diis_p, mask);
}

void launch_diis(const BatchDev& b, hipStream_t s) {
  if (!b.n_mol || !b.diis_m) return;
  const int nt = (b.max_n + kErrTile - 1) / kErrTile;
  k_diis_begin<<<(b.n_mol + 255) / 256, 256, 0, s>>>(b.n_mol, b.sel, b.diis_k, b.diis_p,
                                                     b.prev_prec, b.active, b.status);
  k_diis_err<<<dim3(nt * (nt + 1) / 2, b.n_mol), 256, 0, s>>>(b.sel, b.mols, b.F, b.P, b.diis_F,
                                                              b.diis_E, b.diis_stride, b.diis_k,
                                                              b.diis_m, b.active, b.status);
  k_diis_dot<<<dim3(kChunks, b.n_mol), 256, 0, s>>>(b.sel, b.mols, b.diis_E, b.diis_stride,
                                                    b.diis_k, b.diis_m, b.diis_part, b.active,
                                                    b.status);
  k_diis_solve<<<(b.n_mol + 63) / 64, 64, 0, s>>>(b.n_mol, b.sel, b.diis_k, b.diis_m,
                                                  b.diis_part, b.diis_B, b.diis_c, b.active,
                                                  b.status);
  k_diis_mix<<<dim3(kChunks, b.n_mol), 256, 0, s>>>(b.sel, b.mols, b.F, b.diis_F, b.diis_stride,
                                                    b.diis_c, b.active, b.status);
}

int64_t diis_part_doubles(int n_mol) { return (int64_t)n_mol * kChunks * kDiisMax; }

}  // namespace sg
