// sg_large.hip -- eigensolver for molecules too large for one workgroup
// (n_ao >= BatchDev::large_min, up to kMaxLarge basis functions).
//
// Same algorithm as the per-workgroup solver in sg_linalg.hip (blocked
// Householder tridiagonalisation as LAPACK dsytrd/dlatrd, bisection, inverse
// iteration, compact-WY back-transformation), with each stage spread over
// the whole GPU for one molecule:
//  * the work matrix is kept in full symmetric storage, so the product with
//    the trailing block is a row-wise dot that workgroups split by rows;
//  * one column step (finish the previous column's W, build the next
//    reflector) is a single-workgroup launch; the product is a grid launch;
//  * the trailing update of each panel is a grid of 16 x 16 matrix-core tiles;
//  * the back-transformation gives each workgroup a 16-column slice of Y and
//    walks the reflector groups with their T factors precomputed.
// Panels (V, W), the product vector and the T factors live in the molecule's
// eigensolver workspace.
#include "sg_common.h"
#include "sg_kernels.h"

namespace sg {

namespace {

constexpr int kLT = 1024;  // threads of the single-workgroup column step

typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4d mfma(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ double wred(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ double bsum(double v, double* red) {
  v = wred(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, nw = blockDim.x >> 6;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < nw; ++i) t += red[i];
  return t;
}

}  // namespace

__global__ void k_lt_copy(const double* __restrict__ F, double* __restrict__ A, int64_t nn,
                          const int* __restrict__ active, const int* __restrict__ status, int m) {
  if (!mol_live(active, status, m)) return;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nn;
       t += (int64_t)gridDim.x * blockDim.x)
    A[t] = F[t];
}

// Column step of panel k0: if fin, W column j-1 from the product py (the
// dlatrd update of column c-1); if col, reflector of column c = k0 + j.
__global__ void __launch_bounds__(kLT) k_lt_column(int n, double* __restrict__ A,
                                                    double* __restrict__ V,
                                                    double* __restrict__ W,
                                                    const double* __restrict__ py,
                                                    double* __restrict__ d, double* __restrict__ e,
                                                    double* __restrict__ tau, int k0, int j,
                                                    int fin, int col,
                                                    const int* __restrict__ active,
                                                    const int* __restrict__ status, int m) {
  if (!mol_live(active, status, m)) return;
  extern __shared__ double lds[];
  double* x = lds;           // n
  double* red = x + n;       // 16
  double* tv = red + 16;     // 2 kLargeNb
  const int tid = threadIdx.x, nt = blockDim.x;
  const int wave = tid >> 6, lane = tid & 63, nwave = nt >> 6;
  if (fin) {
    const int jj = j - 1, c1 = k0 + jj;
    const double tk = tau[c1];
    const double* vj = V + (int64_t)jj * n;
    for (int qq = wave; qq < 2 * jj; qq += nwave) {
      const double* cv = qq < jj ? W + (int64_t)qq * n : V + (int64_t)(qq - jj) * n;
      double s = 0.0;
      for (int i = c1 + 1 + lane; i < n; i += 64) s += cv[i] * vj[i];
      s = wred(s);
      if (lane == 0) tv[qq] = s;
    }
    __syncthreads();
    double pv = 0.0;
    for (int i = c1 + 1 + tid; i < n; i += nt) {
      double s = py[i];
      for (int q = 0; q < jj; ++q)
        s -= V[(int64_t)q * n + i] * tv[q] + W[(int64_t)q * n + i] * tv[jj + q];
      s *= tk;
      x[i] = s;
      pv += s * vj[i];
    }
    pv = bsum(pv, red);
    const double al = -0.5 * tk * pv;
    double* wj = W + (int64_t)jj * n;
    for (int i = tid; i < n; i += nt) wj[i] = i > c1 ? x[i] + al * vj[i] : 0.0;
    __threadfence_block();
    __syncthreads();
  }
  if (!col) return;
  const int c = k0 + j;
  // column c = row c of the symmetric matrix, less the panel's corrections
  for (int i = c + tid; i < n; i += nt) {
    double s = A[(int64_t)c * n + i];
    for (int q = 0; q < j; ++q)
      s -= V[(int64_t)q * n + i] * W[(int64_t)q * n + c] + W[(int64_t)q * n + i] * V[(int64_t)q * n + c];
    x[i] = s;
  }
  __syncthreads();
  double s2 = 0.0;
  for (int i = c + 2 + tid; i < n; i += nt) s2 += x[i] * x[i];
  s2 = bsum(s2, red);
  const double alpha = x[c + 1];
  double* vj = V + (int64_t)j * n;
  if (s2 == 0.0) {
    if (tid == 0) {
      d[c] = x[c];
      e[c] = alpha;
      tau[c] = 0.0;
    }
    for (int i = tid; i < n; i += nt) {
      vj[i] = 0.0;
      if (i > c) A[(int64_t)c * n + i] = 0.0;
    }
    return;
  }
  const double beta = -copysign(sqrt(alpha * alpha + s2), alpha);
  const double tk = (beta - alpha) / beta;
  const double sc = 1.0 / (alpha - beta);
  for (int i = tid; i < n; i += nt) {
    const double v = i <= c ? 0.0 : (i == c + 1 ? 1.0 : x[i] * sc);
    vj[i] = v;
    if (i > c) A[(int64_t)c * n + i] = v;
  }
  if (tid == 0) {
    d[c] = x[c];
    e[c] = beta;
    tau[c] = tk;
  }
}

// py[i] = sum_{l > c} A[i][l] v_l for i > c: a wave takes 4 rows, the 4 row
// sums are reduced together.
__global__ void __launch_bounds__(256) k_lt_symv(int n, const double* __restrict__ A,
                                                  const double* __restrict__ v,
                                                  double* __restrict__ py, int c,
                                                  const int* __restrict__ active,
                                                  const int* __restrict__ status, int m) {
  if (!mol_live(active, status, m)) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i0 = c + 1 + (blockIdx.x * 4 + wave) * 4;
  if (i0 >= n) return;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const double* ar[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) ar[r] = A + (int64_t)min(i0 + r, n - 1) * n;
  for (int l = c + 1 + lane; l < n; l += 64) {
    const double vl = v[l];
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[r] += ar[r][l] * vl;
  }
  // lanes < 32 keep rows 0,1 / >= 32 rows 2,3; then 16-lane halves
  double v2[2], v1;
  {
    const bool hi = lane & 32;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const double send = hi ? acc[k] : acc[k + 2];
      v2[k] = (hi ? acc[k + 2] : acc[k]) + __shfl_xor(send, 32, 64);
    }
    const bool h16 = lane & 16;
    const double send = h16 ? v2[0] : v2[1];
    v1 = (h16 ? v2[1] : v2[0]) + __shfl_xor(send, 16, 64);
  }
#pragma unroll
  for (int off = 8; off >= 1; off >>= 1) v1 += __shfl_xor(v1, off, 64);
  const int rho = ((lane >> 5) & 1) * 2 + ((lane >> 4) & 1);
  if ((lane & 15) == 0 && i0 + rho < n) py[i0 + rho] = v1;
}

// A[r:, r:] -= V W^T + W V^T (both triangles), one 16 x 16 tile per wave.
__global__ void __launch_bounds__(256) k_lt_update(int n, double* __restrict__ A,
                                                    const double* __restrict__ V,
                                                    const double* __restrict__ W, int r,
                                                    int nb, const int* __restrict__ active,
                                                    const int* __restrict__ status, int m) {
  if (!mol_live(active, status, m)) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lr = lane & 15, lk = lane >> 4;
  const int TT = (n - r + 15) >> 4;
  const int64_t t = (int64_t)blockIdx.x * 4 + wave;
  if (t >= (int64_t)TT * TT) return;
  const int I0 = r + 16 * (int)(t / TT), J0 = r + 16 * (int)(t % TT);
  const int col = J0 + lr;
  v4d acc;
#pragma unroll
  for (int rr = 0; rr < 4; ++rr) {
    const int row = I0 + 4 * rr + lk;
    acc[rr] = (row < n && col < n) ? A[(int64_t)row * n + col] : 0.0;
  }
  const int ra = min(I0 + lr, n - 1), cb = min(J0 + lr, n - 1);
  const int nsteps = (2 * nb + 3) >> 2;
  for (int st = 0; st < nsteps; ++st) {
    const int k = 4 * st + lk;
    double av = 0.0, bv = 0.0;
    if (k < nb) {
      av = -V[(int64_t)k * n + ra];
      bv = W[(int64_t)k * n + cb];
    } else if (k < 2 * nb) {
      av = -W[(int64_t)(k - nb) * n + ra];
      bv = V[(int64_t)(k - nb) * n + cb];
    }
    acc = mfma(av, bv, acc);
  }
#pragma unroll
  for (int rr = 0; rr < 4; ++rr) {
    const int row = I0 + 4 * rr + lk;
    if (row < n && col < n) A[(int64_t)row * n + col] = acc[rr];
  }
}

__global__ void k_lt_tail(int n, const double* __restrict__ A, double* __restrict__ d,
                          double* __restrict__ e, const int* __restrict__ active,
                          const int* __restrict__ status, int m) {
  if (threadIdx.x == 0 && mol_live(active, status, m)) {
    if (n >= 2) {
      d[n - 2] = A[(int64_t)(n - 2) * n + (n - 2)];
      e[n - 2] = A[(int64_t)(n - 1) * n + (n - 2)];
    }
    d[n - 1] = A[(int64_t)(n - 1) * n + (n - 1)];
    e[n - 1] = 0.0;
  }
}

// T factor (16 x 16, upper) of each group of 16 reflectors k0 = 16 g:
// H_k0 ... H_(k0+b-1) = I - V T V^T, V row q = reflector k0+q (upper part of
// row k0+q of A).  One workgroup per group; the Gram matrix V V^T is summed
// over row slices by the waves (fixed order).
__global__ void __launch_bounds__(256) k_bt_tfactor(int n, int nref, const double* __restrict__ A,
                                                     const double* __restrict__ tau,
                                                     double* __restrict__ Tall,
                                                     const int* __restrict__ active,
                                                     const int* __restrict__ status, int m) {
  if (!mol_live(active, status, m)) return;
  __shared__ double part[4][256];
  __shared__ double Gm[256];
  __shared__ double Tm[256];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lr = lane & 15, lk = lane >> 4;
  const int k0 = 16 * blockIdx.x;
  const int b = min(16, nref - k0);
  v4d acc = {0, 0, 0, 0};
  for (int i0 = k0 + 1 + 4 * wave; i0 < n; i0 += 16) {
    const int i = i0 + lk, c = k0 + lr;
    const double a = (lr < b && i < n && i > c) ? A[(int64_t)c * n + i] : 0.0;
    acc = mfma(a, a, acc);
  }
#pragma unroll
  for (int rr = 0; rr < 4; ++rr) part[wave][(4 * rr + lk) * 16 + lr] = acc[rr];
  Tm[threadIdx.x] = 0.0;
  __syncthreads();
  Gm[threadIdx.x] = ((part[0][threadIdx.x] + part[1][threadIdx.x]) + part[2][threadIdx.x]) +
                    part[3][threadIdx.x];
  __syncthreads();
  if (wave == 0) {
    for (int i = 0; i < b; ++i) {
      const double ti = tau[k0 + i];
      if (lane < i) {
        double s = 0.0;
        for (int q = lane; q < i; ++q) s += Tm[lane * 16 + q] * Gm[q * 16 + i];
        Tm[lane * 16 + i] = -ti * s;
      }
      if (lane == 0) Tm[i * 16 + i] = ti;
      __threadfence_block();
      __builtin_amdgcn_wave_barrier();
    }
  }
  __syncthreads();
  Tall[(int64_t)blockIdx.x * 256 + threadIdx.x] = Tm[threadIdx.x];
}

// Y (n x k, row-major) <- Q Y for the 16-column slice of this workgroup,
// groups from last to first: W1 = V Y (waves split the rows, partial tiles
// summed in order), W2 = T W1, Y -= V^T W2 (row tiles split over the waves).
__global__ void __launch_bounds__(256) k_bt_large(int n, int nref, int k,
                                                   const double* __restrict__ A,
                                                   const double* __restrict__ Tall,
                                                   double* __restrict__ Y,
                                                   const int* __restrict__ active,
                                                   const int* __restrict__ status, int m) {
  if (!mol_live(active, status, m)) return;
  __shared__ double part[4][256];
  __shared__ double W2[256];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lr = lane & 15, lk = lane >> 4;
  const int j = 16 * blockIdx.x + lr;
  const int ng = (nref + 15) / 16;
  for (int g = ng - 1; g >= 0; --g) {
    const int k0 = 16 * g, b = min(16, nref - k0);
    v4d w1 = {0, 0, 0, 0};
    for (int i0 = k0 + 1 + 4 * wave; i0 < n; i0 += 16) {
      const int i = i0 + lk, c = k0 + lr;
      const double a = (lr < b && i < n && i > c) ? A[(int64_t)c * n + i] : 0.0;
      const double y = (i < n && j < k) ? Y[(int64_t)i * k + j] : 0.0;
      w1 = mfma(a, y, w1);
    }
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) part[wave][(4 * rr + lk) * 16 + lr] = w1[rr];
    __syncthreads();
    if (wave == 0) {
      v4d s;
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int idx = (4 * rr + lk) * 16 + lr;
        s[rr] = ((part[0][idx] + part[1][idx]) + part[2][idx]) + part[3][idx];
      }
      const double* T = Tall + (int64_t)g * 256;
      v4d w2 = {0, 0, 0, 0};
#pragma unroll
      for (int st = 0; st < 4; ++st) w2 = mfma(T[lr * 16 + 4 * st + lk], s[st], w2);
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) W2[(4 * rr + lk) * 16 + lr] = w2[rr];
    }
    __syncthreads();
    const int IT = (n - (k0 + 1) + 15) >> 4;
    for (int it = wave; it < IT; it += 4) {
      const int I0 = k0 + 1 + 16 * it;
      v4d cur;
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int row = I0 + 4 * rr + lk;
        cur[rr] = (row < n && j < k) ? Y[(int64_t)row * k + j] : 0.0;
      }
#pragma unroll
      for (int st = 0; st < 4; ++st) {
        const int q = 4 * st + lk, i = I0 + lr, c = k0 + q;
        const double a = (q < b && i < n && i > c) ? -A[(int64_t)c * n + i] : 0.0;
        cur = mfma(a, W2[q * 16 + lr], cur);
      }
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int row = I0 + 4 * rr + lk;
        if (row < n && j < k) Y[(int64_t)row * k + j] = cur[rr];
      }
    }
    __syncthreads();
  }
}

void launch_tridiag_large(const BatchDev& b, hipStream_t s, int m, const MolDesc& M,
                          const int* act) {
  const int n = M.n_ao;
  LargeWork L(b.W, M);
  double* d = b.dT + M.vec0;
  double* e = b.eT + M.vec0;
  double* tau = b.tau + M.vec0;
  const int* st = b.status;
  k_lt_copy<<<1024, 256, 0, s>>>(b.F + M.mat0, L.A, (int64_t)n * n, act, st, m);
  const int nref = n > 2 ? n - 2 : 0;
  const size_t col_lds = sizeof(double) * ((size_t)n + 16 + 2 * kLargeNb);
  for (int k0 = 0; k0 < nref; k0 += kLargeNb) {
    const int nb = min(kLargeNb, nref - k0);
    for (int j = 0; j < nb; ++j) {
      const int c = k0 + j;
      k_lt_column<<<1, kLT, col_lds, s>>>(n, L.A, L.V, L.W, L.py, d, e, tau, k0, j, j > 0, 1, act,
                                          st, m);
      const int rows = n - c - 1;
      k_lt_symv<<<(rows + 15) / 16, 256, 0, s>>>(n, L.A, L.V + (int64_t)j * n, L.py, c, act, st, m);
    }
    k_lt_column<<<1, kLT, col_lds, s>>>(n, L.A, L.V, L.W, L.py, d, e, tau, k0, nb, 1, 0, act, st,
                                        m);
    const int r = k0 + nb;
    const int TT = (n - r + 15) >> 4;
    const int64_t tiles = (int64_t)TT * TT;
    k_lt_update<<<(unsigned)((tiles + 3) / 4), 256, 0, s>>>(n, L.A, L.V, L.W, r, nb, act, st, m);
  }
  k_lt_tail<<<1, 64, 0, s>>>(n, L.A, d, e, act, st, m);
}

void launch_back_large(const BatchDev& b, hipStream_t s, int m, const MolDesc& M,
                       const int* act, bool all) {
  const int n = M.n_ao;
  const int nref = n > 2 ? n - 2 : 0;
  int lo, k;
  selection(M, all ? 1 : 0, lo, k);
  if (nref == 0 || k == 0) return;
  LargeWork L(b.W, M);
  const int ng = (nref + 15) / 16;
  k_bt_tfactor<<<ng, 256, 0, s>>>(n, nref, L.A, b.tau + M.vec0, L.T, act, b.status, m);
  k_bt_large<<<(k + 15) / 16, 256, 0, s>>>(n, nref, k, L.A, L.T, L.Y, act, b.status, m);
}

}  // namespace sg
