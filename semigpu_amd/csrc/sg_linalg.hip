// sg_linalg.hip -- density matrix and dense symmetric eigensolver, batched.
//
// k_density: P = 2 C_occ C_occ^T (DensityMatrix::ComputeElementMatrix,
//   densitymatrix.cpp:38-45); C column-major, n_occ = floor(n_e / 2).
//
// k_eigh: replaces LAPACK dsyev('V','L') on the Fock matrix
//   (SCFCalculation::GetEigenValVecOfFmatrix, scfcalculation.cpp:171-203).
//   One workgroup per molecule:
//     1. Householder tridiagonalisation of a copy of F (workspace W), the
//        reflector of column k kept in row k of W;
//     2. Q = H_0 ... H_(n-3) accumulated backward into the workspace Z;
//     3. implicit symmetric QR with Wilkinson shifts on (d, e) -- one lane
//        computes a sweep's Givens rotations into LDS, then every lane replays
//        the sweep on its own rows of Z (rows are independent, no barriers
//        inside a sweep);
//     4. eigenvalues sorted ascending, eigenvectors written to C.
//   Golub & Van Loan, Matrix Computations (4th ed.) 8.3.
#include "sg_common.h"
#include "sg_kernels.h"

namespace sg {

constexpr int kEigThreads = 256;
constexpr int kEigMaxN = 1024;

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block-wide sum, result returned to every thread (fixed order).
__device__ double block_sum(double v, double* red) {
  v = wsum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
  const int nw = blockDim.x >> 6;
  for (int i = 0; i < nw; ++i) t += red[i];
  return t;
}

__global__ void __launch_bounds__(kEigThreads) k_eigh(const MolDesc* __restrict__ mols,
                                                       const double* __restrict__ F,
                                                       double* __restrict__ Wbuf,
                                                       double* __restrict__ C,
                                                       double* __restrict__ eig,
                                                       const int* __restrict__ active,
                                                       int* __restrict__ status) {
  const int m = blockIdx.x;
  if (active && (!active[m] || status[m])) return;
  const MolDesc M = mols[m];
  const int n = M.n_ao;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int wave = tid >> 6, lane = tid & 63, nwave = nt >> 6;
  double* W = Wbuf + 2 * M.mat0;       // n x n, row-major
  double* Z = W + (int64_t)n * n;      // n x n, column-major
  const double* f = F + M.mat0;
  double* c = C + M.mat0;
  double* w_out = eig + M.vec0;

  __shared__ double d[kEigMaxN], e[kEigMaxN], vs[kEigMaxN], ps[kEigMaxN];
  __shared__ double tau_s[kEigMaxN];
  __shared__ double red[kEigThreads / 64];
  __shared__ int perm[kEigMaxN];
  __shared__ int s_lo, s_hi, s_stat;

  for (int64_t t = tid; t < (int64_t)n * n; t += nt) W[t] = f[t];
  __syncthreads();

  // 1. tridiagonalisation
  for (int k = 0; k + 2 < n; ++k) {
    const double* rowk = W + (int64_t)k * n;
    double s2 = 0.0;
    for (int i = k + 2 + tid; i < n; i += nt) s2 += rowk[i] * rowk[i];
    s2 = block_sum(s2, red);
    const double x0 = rowk[k + 1];
    if (s2 == 0.0) {
      if (tid == 0) {
        tau_s[k] = 0.0;
        e[k] = x0;
      }
      __syncthreads();
      continue;
    }
    const double nrm = sqrt(x0 * x0 + s2);
    const double beta = x0 >= 0.0 ? -nrm : nrm;
    const double tk = (beta - x0) / beta;
    const double sc = 1.0 / (x0 - beta);
    for (int i = k + 1 + tid; i < n; i += nt) {
      const double v = i == k + 1 ? 1.0 : rowk[i] * sc;
      vs[i] = v;
    }
    if (tid == 0) {
      tau_s[k] = tk;
      e[k] = beta;
    }
    __syncthreads();
    // store v in row k (read back in step 2)
    for (int i = k + 1 + tid; i < n; i += nt) W[(int64_t)k * n + i] = vs[i];
    // p = tau * W22 v : one wave per row
    for (int i = k + 1 + wave; i < n; i += nwave) {
      const double* ri = W + (int64_t)i * n;
      double s = 0.0;
      for (int j = k + 1 + lane; j < n; j += 64) s += ri[j] * vs[j];
      s = wsum(s);
      if (lane == 0) ps[i] = tk * s;
    }
    __syncthreads();
    double kk = 0.0;
    for (int i = k + 1 + tid; i < n; i += nt) kk += ps[i] * vs[i];
    kk = block_sum(kk, red) * (0.5 * tk);
    for (int i = k + 1 + tid; i < n; i += nt) ps[i] -= kk * vs[i];
    __syncthreads();
    const int mm = n - k - 1;
    for (int64_t t = tid; t < (int64_t)mm * mm; t += nt) {
      const int i = k + 1 + (int)(t / mm), j = k + 1 + (int)(t % mm);
      W[(int64_t)i * n + j] -= vs[i] * ps[j] + ps[i] * vs[j];
    }
    __syncthreads();
  }
  for (int i = tid; i < n; i += nt) d[i] = W[(int64_t)i * n + i];
  if (tid == 0) {
    if (n >= 2) e[n - 2] = W[(int64_t)(n - 1) * n + (n - 2)];
    if (n >= 1) e[n - 1] = 0.0;
    s_stat = 0;
  }
  // 2. Q accumulated backward into Z (column-major)
  for (int64_t t = tid; t < (int64_t)n * n; t += nt) {
    const int col = (int)(t / n), row = (int)(t % n);
    Z[t] = row == col ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int k = n - 3; k >= 0; --k) {
    const double tk = tau_s[k];
    if (tk == 0.0) continue;
    for (int i = k + 1 + tid; i < n; i += nt) vs[i] = W[(int64_t)k * n + i];
    __syncthreads();
    for (int j = k + 1 + wave; j < n; j += nwave) {
      double* zj = Z + (int64_t)j * n;
      double s = 0.0;
      for (int i = k + 1 + lane; i < n; i += 64) s += vs[i] * zj[i];
      s = wsum(s) * tk;
      for (int i = k + 1 + lane; i < n; i += 64) zj[i] -= s * vs[i];
    }
    __syncthreads();
  }
  // 3. implicit QR on (d, e); rotations of a sweep in vs (cos) / ps (sin)
  const double eps = 2.220446049250313e-16;
  int iters = 0;
  while (true) {
    if (tid == 0) {
      int hi = n - 1;
      for (int i = 0; i < hi; ++i)
        if (fabs(e[i]) <= eps * (fabs(d[i]) + fabs(d[i + 1]))) e[i] = 0.0;
      while (hi > 0 && e[hi - 1] == 0.0) --hi;
      int lo = hi;
      if (hi > 0) {
        lo = hi - 1;
        while (lo > 0 && e[lo - 1] != 0.0) --lo;
        if (++iters > 60 * n) {
          s_stat = hi + 1;
          hi = 0;
        } else {
          const double dl = (d[hi - 1] - d[hi]) * 0.5, b = e[hi - 1];
          const double den = dl + (dl >= 0.0 ? hypot(dl, b) : -hypot(dl, b));
          const double mu = d[hi] - b * b / den;
          double x = d[lo] - mu, y = e[lo];
          for (int k = lo; k < hi; ++k) {
            const double r = hypot(x, y);
            double cs = 1.0, sn = 0.0;
            if (r != 0.0) {
              cs = x / r;
              sn = y / r;
            }
            if (k > lo) e[k - 1] = r;
            const double dk = d[k], dk1 = d[k + 1], ek = e[k];
            d[k] = cs * cs * dk + 2.0 * cs * sn * ek + sn * sn * dk1;
            d[k + 1] = sn * sn * dk - 2.0 * cs * sn * ek + cs * cs * dk1;
            e[k] = cs * sn * (dk1 - dk) + (cs * cs - sn * sn) * ek;
            if (k + 1 < hi) {
              y = sn * e[k + 1];
              e[k + 1] = cs * e[k + 1];
              x = e[k];
            }
            vs[k] = cs;
            ps[k] = sn;
          }
        }
      }
      s_lo = lo;
      s_hi = hi;
    }
    __syncthreads();
    const int lo = s_lo, hi = s_hi;
    if (hi <= 0) break;
    for (int i = tid; i < n; i += nt) {
      double carry = Z[(int64_t)lo * n + i];
      for (int k = lo; k < hi; ++k) {
        const double nxt = Z[(int64_t)(k + 1) * n + i];
        const double cs = vs[k], sn = ps[k];
        Z[(int64_t)k * n + i] = cs * carry + sn * nxt;
        carry = -sn * carry + cs * nxt;
      }
      Z[(int64_t)hi * n + i] = carry;
    }
    __syncthreads();
  }
  // 4. ascending order (rank by value, ties by index)
  for (int i = tid; i < n; i += nt) {
    const double di = d[i];
    int r = 0;
    for (int j = 0; j < n; ++j) {
      const double dj = d[j];
      r += (dj < di) || (dj == di && j < i);
    }
    perm[i] = r;
  }
  __syncthreads();
  for (int64_t t = tid; t < (int64_t)n * n; t += nt) {
    const int col = (int)(t / n), row = (int)(t % n);
    c[(int64_t)perm[col] * n + row] = Z[t];
  }
  for (int i = tid; i < n; i += nt) w_out[perm[i]] = d[i];
  if (tid == 0 && s_stat) status[m] = s_stat;
}

// P = 2 C_occ C_occ^T: one workgroup per (molecule, 16-row strip).
__global__ void __launch_bounds__(256) k_density(const MolDesc* __restrict__ mols,
                                                  const double* __restrict__ C,
                                                  double* __restrict__ P,
                                                  const int* __restrict__ active,
                                                  const int* __restrict__ status, int strips) {
  const int m = blockIdx.y;
  if (active && (!active[m] || status[m])) return;
  const MolDesc M = mols[m];
  const int n = M.n_ao, no = M.n_occ;
  const double* c = C + M.mat0;
  double* p = P + M.mat0;
  for (int row = blockIdx.x; row < n; row += strips) {
    for (int col = threadIdx.x; col < n; col += blockDim.x) {
      double s = 0.0;
      for (int k = 0; k < no; ++k) s += c[(int64_t)k * n + row] * c[(int64_t)k * n + col];
      p[(int64_t)row * n + col] = s * 2.0;
    }
  }
}

void launch_eigh(const BatchDev& b, hipStream_t s, bool all) {
  k_eigh<<<b.n_mol, kEigThreads, 0, s>>>(b.mols, b.F, b.W, b.C, b.eig, all ? nullptr : b.active,
                                         b.status);
}

void launch_density(const BatchDev& b, hipStream_t s, bool all) {
  const int strips = 64;
  dim3 grid(strips, b.n_mol);
  k_density<<<grid, 256, 0, s>>>(b.mols, b.C, b.P, all ? nullptr : b.active, b.status, strips);
}

}  // namespace sg
