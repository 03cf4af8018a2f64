// sg_large.hip -- eigensolver for molecules too large for one workgroup
// (n_ao >= BatchDev::large_min, up to kMaxLarge basis functions).
//
// Same algorithm as the per-workgroup solver in sg_linalg.hip (blocked
// Householder tridiagonalisation as LAPACK dsytrd/dlatrd, bisection, inverse
// iteration, compact-WY back-transformation), with each stage spread over
// the whole GPU for one molecule:
//  * the work matrix is kept in full symmetric storage, so the product with
//    the trailing block is a row-wise dot that workgroups split by rows;
//  * a column is two grid launches: the product (workgroups over blocks of
//    32 rows and chunks of kLargeChunkCols columns, so that the grid fills the
//    GPU down to a few thousand rows; it also forms the reflector from the
//    column and leaves per-workgroup partial rows and dots), then the
//    completion of W's column with the next column's corrections (one row per
//    thread), whose workgroups each add the partial rows and dots in order;
//  * the trailing update of each panel is a grid of 16 x 16 matrix-core tiles;
//  * the back-transformation gives each workgroup a 16-column slice of Y and
//    walks the reflector groups with their T factors precomputed.
// Panels (V, W), the product vector and the T factors live in the molecule's
// eigensolver workspace.
#include "sg_common.h"
#include "sg_kernels.h"

namespace sg {

namespace {

constexpr int kColRows = 256;                    // rows per workgroup of the column kernels
constexpr int kProdWaves = 8;                    // waves per product workgroup (4 rows each)
constexpr int kPartStride = 2 * kLargeNb + 1;    // partial dots per product workgroup

typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4d mfma(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ double wred(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

}  // namespace

__global__ void k_lt_copy(const double* __restrict__ F, double* __restrict__ A, int64_t nn,
                          const int* __restrict__ active, const int* __restrict__ status, int m) {
  if (!mol_live(active, status, m)) return;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < nn;
       t += (int64_t)gridDim.x * blockDim.x)
    A[t] = F[t];
}

// Sum of one value per thread over the workgroup, in a fixed order.
__device__ double block_total(double v, double* red) {
  v = wred(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, nw = blockDim.x >> 6;
  if (l == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < nw; ++i) t += red[i];
  __syncthreads();
  return t;
}

// Column c with no panel corrections (the first column of a panel): x = row
// c of A from the diagonal on, and per-wavefront partial sums of x_i^2, i >= c+2.
__global__ void __launch_bounds__(kColRows) k_lt_x0(int n, const double* __restrict__ A,
                                                     double* __restrict__ xv,
                                                     double* __restrict__ part2, int c,
                                                     const int* __restrict__ active,
                                                     const int* __restrict__ status, int m) {
  if (!mol_live(active, status, m)) return;
  const int i = c + blockIdx.x * kColRows + threadIdx.x;
  double sq = 0.0;
  if (i < n) {
    const double v = A[(int64_t)c * n + i];
    xv[i] = v;
    if (i >= c + 2) sq = v * v;
  }
  sq = wred(sq);
  if ((threadIdx.x & 63) == 0) part2[blockIdx.x * (kColRows / 64) + (threadIdx.x >> 6)] = sq;
}

// Reflector of column c (j-th of its panel) from x and the partial sums of
// x^2, then the product py = A v over rows i > c (4 rows per wave),
// workgroup (x, y) over rows c + 1 + 32 x .. and columns c + 1 +
// kLargeChunkCols y ..: its rows' partial sums into pyp[y] and their dot with
// v into partvp; the workgroups of chunk 0 also leave the partial dots W_q.v
// and V_q.v (q < j) of their rows and write v into V_j and into the upper
// part of row c of A; d, e, tau of column c.  v = sc x past row c + 1 and
// 1 at c + 1, so the product and the dots run on x and take the scale sc at
// the end: their loads do not wait for the norm's partial sums.
__global__ void __launch_bounds__(64 * kProdWaves) k_lt_prod(int n, double* __restrict__ A,
                                                              double* __restrict__ V,
                                                              const double* __restrict__ W,
                                                              const double* __restrict__ xv,
                                                              const double* __restrict__ part2,
                                                              double* __restrict__ part,
                                                              double* __restrict__ pyp,
                                                              double* __restrict__ partvp,
                                                              double* __restrict__ d,
                                                              double* __restrict__ e,
                                                              double* __restrict__ tau, int c, int j,
                                                              const int* __restrict__ active,
                                                              const int* __restrict__ status,
                                                              int m) {
  if (!mol_live(active, status, m)) return;
  constexpr int R = 4 * kProdWaves;  // rows per workgroup
  __shared__ double red[kProdWaves];
  __shared__ double rows_py[R], rows_v[R];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int cb = blockIdx.y;              // column chunk
  const int r0 = c + 1 + blockIdx.x * R;  // first row of this workgroup
  const int i0 = r0 + wave * 4;
  // x past row c + 1 (0 at c + 1 and past the matrix)
  auto xof = [&](int l) { return l > c + 1 && l < n ? xv[l] : 0.0; };
  // the panel dots over this workgroup's rows (chunk 0): with x, and the
  // entry at row c + 1 apart
  const int nd = cb == 0 ? 2 * j : 0;
  double pdx = 0.0, pd1 = 0.0;
  if (threadIdx.x < nd) {
    const int t = threadIdx.x;
    const double* src = t < j ? W + (int64_t)t * n : V + (int64_t)(t - j) * n;
    for (int b0 = 0; b0 < R; b0 += 16) {
      double x[16], y[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int i = r0 + b0 + r;
        x[r] = src[min(i, n - 1)];
        y[r] = xof(i);
      }
#pragma unroll
      for (int r = 0; r < 16; ++r) pdx += x[r] * y[r];
      if (b0 == 0 && r0 == c + 1) pd1 = x[0];
    }
  }
  // the norm's partial sums, one per wavefront of the column kernels (loaded
  // now, added after the product)
  const int nG = (n - c + kColRows - 1) / kColRows * (kColRows / 64);
  double s2 = 0.0;
  for (int g = threadIdx.x; g < nG; g += blockDim.x) s2 += part2[g];
  double acc[4] = {0.0, 0.0, 0.0, 0.0}, acc1[4] = {0.0, 0.0, 0.0, 0.0};
  const double* ar[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) ar[r] = A + (int64_t)min(i0 + r, n - 1) * n;
  const int lbeg = c + 1 + cb * kLargeChunkCols, lend = min(n, lbeg + kLargeChunkCols);
  if (i0 < n) {
    // two 64-column tiles per batch of loads; the second is clamped to the
    // chunk's last column and weighted 0 past it; two batches in flight
#pragma unroll 2
    for (int l = lbeg + lane; l < lend; l += 128) {
      const int l2 = l + 64 < lend ? l + 64 : lend - 1;
      const double x0 = xof(l), x1 = l + 64 < lend ? xof(l2) : 0.0;
      double a0[4], a1[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        a0[r] = ar[r][l];
        a1[r] = ar[r][l2];
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[r] += a0[r] * x0 + a1[r] * x1;
      // column c + 1 (the first of chunk 0, lane 0's first load) takes weight 1
      if (l == c + 1) {
#pragma unroll
        for (int r = 0; r < 4; ++r) acc1[r] = a0[r];
      }
    }
  }
  // reflector scalars (every workgroup computes the same numbers)
  s2 = block_total(s2, red);
  const double alpha = xv[c + 1];
  const bool zero = s2 == 0.0;
  double beta = alpha, tk = 0.0, sc = 0.0;
  if (!zero) {
    beta = -copysign(sqrt(alpha * alpha + s2), alpha);
    tk = (beta - alpha) / beta;
    sc = 1.0 / (alpha - beta);
  }
  if (blockIdx.x == 0 && cb == 0 && threadIdx.x == 0) {
    d[c] = xv[c];
    e[c] = beta;
    tau[c] = tk;
  }
  const double one = zero ? 0.0 : 1.0;
#pragma unroll
  for (int r = 0; r < 4; ++r) acc[r] = sc * acc[r] + one * acc1[r];
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
  if ((lane & 15) == 0) {
    const int i = i0 + rho;
    rows_py[wave * 4 + rho] = i < n ? v1 : 0.0;
    if (i < n) pyp[(int64_t)cb * n + i] = v1;
  }
  if (threadIdx.x < R) {
    const int i = r0 + threadIdx.x;
    const double v = i >= n ? 0.0 : (i == c + 1 ? one : xv[i] * sc);
    rows_v[threadIdx.x] = v;
    if (i < n && cb == 0) {
      V[(int64_t)j * n + i] = v;
      A[(int64_t)c * n + i] = v;
    }
  }
  if (threadIdx.x < nd) part[(int64_t)blockIdx.x * kPartStride + threadIdx.x] = sc * pdx + one * pd1;
  __syncthreads();
  // v.py over this workgroup's rows (of this chunk's columns), by wave 0
  if (wave == 0) {
    const double pr = lane < R ? rows_py[lane] * rows_v[lane] : 0.0;  // rows_v is 0 past the matrix
    const double sacc = wred(pr);
    if (lane == 0) partvp[(int64_t)blockIdx.x * gridDim.y + cb] = sacc;
  }
}

// Completes W_j (the dlatrd update of column c = k0 + j):
//   s = tau (py - V tv1 - W tv2),  W_j = s + a v,  a = -tau/2 (s.v),
// with tv1 = W^T v, tv2 = V^T v and s.v = tau (v.py - 2 tv1.tv2) from the
// product's partial dots (added up by every workgroup); then, when another
// column of the panel follows, x for column c+1 (row c+1 of A less the panel's
// corrections) and the partial sums of x^2.  One row per thread.
__global__ void __launch_bounds__(kColRows) k_lt_cols(int n, const double* __restrict__ A,
                                                       const double* __restrict__ V,
                                                       double* __restrict__ W,
                                                       const double* __restrict__ pyp,
                                                       const double* __restrict__ part,
                                                       const double* __restrict__ partvp,
                                                       const double* __restrict__ tau,
                                                       double* __restrict__ xv,
                                                       double* __restrict__ part2, int c, int j,
                                                       int next, const int* __restrict__ active,
                                                       const int* __restrict__ status, int m) {
  if (!mol_live(active, status, m)) return;
  __shared__ double tv[2 * kLargeNb + 1];
  __shared__ double red[kColRows / 64];
  const int nch = (n - c - 1 + kLargeChunkCols - 1) / kLargeChunkCols;  // the product's chunks
  const int i = c + 1 + blockIdx.x * kColRows + threadIdx.x, c1 = c + 1;
  const bool in = i < n;
  const int ir = in ? i : n - 1;  // loads stay inside the matrix
  // First the loads that do not need the product's dots: the product's rows
  // i and c + 1 (their chunks' partial sums) and, when another column
  // follows, row c + 1 of A at i less the panel's corrections.
  double pi[kLargeChunks], p1[kLargeChunks];
#pragma unroll
  for (int cb = 0; cb < kLargeChunks; ++cb) {
    pi[cb] = cb < nch ? pyp[(int64_t)cb * n + ir] : 0.0;
    p1[cb] = cb < nch ? pyp[(int64_t)cb * n + c1] : 0.0;
  }
  double xa = 0.0;
  if (next) {
    xa = A[(int64_t)c1 * n + ir];
    for (int q0 = 0; q0 < j; q0 += 8) {
      double a[8], b[8], ac[8], bc[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int q = min(q0 + u, j - 1);
        a[u] = V[(int64_t)q * n + ir];
        b[u] = W[(int64_t)q * n + ir];
        ac[u] = V[(int64_t)q * n + c1];
        bc[u] = W[(int64_t)q * n + c1];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (q0 + u < j) xa -= a[u] * bc[u] + b[u] * ac[u];
    }
  }
  {
    // the product's partial dots, summed in workgroup order: thread T takes
    // dot T % 64 (< 2 j) over the row blocks g = T / 64, T / 64 + 4, ...; the
    // four slices are then added in order.  v.py has an entry per (row
    // block, chunk): every thread sums a strided share, then the workgroup
    // adds the shares in a fixed order.  (Every workgroup forms the same sums.)
    __shared__ double sl[kColRows / 64][64];
    const int t = threadIdx.x & 63, k = threadIdx.x >> 6;
    const int G = (n - c - 1 + 4 * kProdWaves - 1) / (4 * kProdWaves), nd = 2 * j;
    const int GE = G * nch;  // entries of v.py
    double vacc = 0.0;
    for (int g0 = threadIdx.x; g0 < GE; g0 += 4 * kColRows) {
      double x[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int g = g0 + u * kColRows;
        x[u] = g < GE ? partvp[g] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) vacc += x[u];
    }
    double sacc = 0.0;
    if (t < nd) {
      constexpr int U = 12;  // loads in flight per thread
      for (int g0 = k; g0 < G; g0 += U * (kColRows / 64)) {
        double x[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int g = g0 + u * (kColRows / 64);
          x[u] = g < G ? part[(int64_t)g * kPartStride + t] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) sacc += x[u];
      }
    }
    sl[k][t] = sacc;
    const double vpy = block_total(vacc, red);  // its barriers also publish sl
    if (threadIdx.x < nd) {
      double v = 0.0;
#pragma unroll
      for (int q = 0; q < kColRows / 64; ++q) v += sl[q][threadIdx.x];
      tv[threadIdx.x] = v;
    }
    if (threadIdx.x == 0) tv[2 * j] = vpy;
  }
  __syncthreads();
  const double tk = tau[c];
  double cross = 0.0;
  for (int q = 0; q < j; ++q) cross += tv[q] * tv[j + q];
  const double al = -0.5 * tk * (tk * (tv[2 * j] - 2.0 * cross));
  const double* vj = V + (int64_t)j * n;
  // s = tau (py - V tv1 - W tv2) at rows i and c + 1 in one pass (loads of 8
  // panel columns issued before their products)
  double si = pi[0], s1 = p1[0];
#pragma unroll
  for (int cb = 1; cb < kLargeChunks; ++cb) {
    si += pi[cb];
    s1 += p1[cb];
  }
  for (int q0 = 0; q0 < j; q0 += 8) {
    double a[8], b[8], ac[8], bc[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int q = min(q0 + u, j - 1);
      a[u] = V[(int64_t)q * n + ir];
      b[u] = W[(int64_t)q * n + ir];
      ac[u] = V[(int64_t)q * n + c1];
      bc[u] = W[(int64_t)q * n + c1];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (q0 + u < j) {
        si -= a[u] * tv[q0 + u] + b[u] * tv[j + q0 + u];
        s1 -= ac[u] * tv[q0 + u] + bc[u] * tv[j + q0 + u];
      }
  }
  double w = 0.0;
  if (in) {
    w = tk * si + al * vj[i];
    W[(int64_t)j * n + i] = w;
  }
  if (!next) return;
  const double wc1 = tk * s1 + al * vj[c1];
  double sq = 0.0;
  if (in) {
    const double x = xa - (vj[i] * wc1 + w * vj[c1]);
    xv[i] = x;
    if (i >= c1 + 2) sq = x * x;
  }
  sq = wred(sq);
  if ((threadIdx.x & 63) == 0) part2[blockIdx.x * (kColRows / 64) + (threadIdx.x >> 6)] = sq;
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
  if (nb == kLargeNb) {
    // a full panel: the 16 k-steps' operands loaded first (V W^T in steps
    // 0..7, W V^T in 8..15), then the products in the same order
    constexpr int S = kLargeNb / 2;
    double av[S], bv[S];
#pragma unroll
    for (int st = 0; st < S; ++st) {
      const int k = 4 * st + lk, kk = st < S / 2 ? k : k - kLargeNb;
      const double* pa = st < S / 2 ? V : W;
      const double* pb = st < S / 2 ? W : V;
      av[st] = pa[(int64_t)kk * n + ra];
      bv[st] = pb[(int64_t)kk * n + cb];
    }
#pragma unroll
    for (int st = 0; st < S; ++st) acc = mfma(-av[st], bv[st], acc);
  } else {
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
// kBtWaves waves; each keeps kBtSteps k-steps of W1's loads and kBtTiles
// update tiles' loads in flight (addresses clamped inside the arrays,
// unused values zeroed after the load; the products run in the order of a
// load-per-step loop).
constexpr int kBtWaves = 8, kBtSteps = 8, kBtTiles = 2;
__global__ void __launch_bounds__(64 * kBtWaves) k_bt_large(int n, int nref, int k,
                                                             const double* __restrict__ A,
                                                             const double* __restrict__ Tall,
                                                             double* __restrict__ Y,
                                                             const int* __restrict__ active,
                                                             const int* __restrict__ status,
                                                             int m) {
  if (!mol_live(active, status, m)) return;
  __shared__ double part[kBtWaves][256];
  __shared__ double W2[256];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lr = lane & 15, lk = lane >> 4;
  const int j = 16 * blockIdx.x + lr;
  const bool jok = j < k;
  const int jc = jok ? j : k - 1;
  const int ng = (nref + 15) / 16;
  for (int g = ng - 1; g >= 0; --g) {
    const int k0 = 16 * g, b = min(16, nref - k0);
    const int c = k0 + lr, cc = min(c, n - 1);
    v4d w1 = {0, 0, 0, 0};
    for (int i0 = k0 + 1 + 4 * wave; i0 < n; i0 += 4 * kBtWaves * kBtSteps) {
      double a[kBtSteps], y[kBtSteps];
#pragma unroll
      for (int u = 0; u < kBtSteps; ++u) {
        const int i = i0 + 4 * kBtWaves * u + lk, ic = min(i, n - 1);
        a[u] = A[(int64_t)cc * n + ic];
        y[u] = Y[(int64_t)ic * k + jc];
      }
#pragma unroll
      for (int u = 0; u < kBtSteps; ++u) {
        const int i = i0 + 4 * kBtWaves * u + lk;
        if (i0 + 4 * kBtWaves * u < n) {
          const double av = (lr < b && i < n && i > c) ? a[u] : 0.0;
          const double yv = (i < n && jok) ? y[u] : 0.0;
          w1 = mfma(av, yv, w1);
        }
      }
    }
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) part[wave][(4 * rr + lk) * 16 + lr] = w1[rr];
    __syncthreads();
    if (wave == 0) {
      v4d s;
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int idx = (4 * rr + lk) * 16 + lr;
        double t = part[0][idx];
#pragma unroll
        for (int w = 1; w < kBtWaves; ++w) t += part[w][idx];
        s[rr] = t;
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
    double wq[4];
#pragma unroll
    for (int st = 0; st < 4; ++st) wq[st] = W2[(4 * st + lk) * 16 + lr];
    for (int it0 = wave; it0 < IT; it0 += kBtWaves * kBtTiles) {
      v4d cur[kBtTiles];
      double av[kBtTiles][4];
#pragma unroll
      for (int u = 0; u < kBtTiles; ++u) {
        const int I0 = k0 + 1 + 16 * (it0 + kBtWaves * u);
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
          const int row = min(I0 + 4 * rr + lk, n - 1);
          cur[u][rr] = Y[(int64_t)row * k + jc];
        }
#pragma unroll
        for (int st = 0; st < 4; ++st) {
          const int q = 4 * st + lk, i = min(I0 + lr, n - 1);
          av[u][st] = A[(int64_t)min(k0 + q, n - 1) * n + i];
        }
      }
#pragma unroll
      for (int u = 0; u < kBtTiles; ++u) {
        const int it = it0 + kBtWaves * u;
        if (it >= IT) break;
        const int I0 = k0 + 1 + 16 * it;
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
          const int row = I0 + 4 * rr + lk;
          if (!(row < n && jok)) cur[u][rr] = 0.0;
        }
#pragma unroll
        for (int st = 0; st < 4; ++st) {
          const int q = 4 * st + lk, i = I0 + lr, cq = k0 + q;
          const double a = (q < b && i < n && i > cq) ? -av[u][st] : 0.0;
          cur[u] = mfma(a, wq[st], cur[u]);
        }
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) {
          const int row = I0 + 4 * rr + lk;
          if (row < n && jok) Y[(int64_t)row * k + j] = cur[u][rr];
        }
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
  for (int k0 = 0; k0 < nref; k0 += kLargeNb) {
    const int nb = min(kLargeNb, nref - k0);
    k_lt_x0<<<(n - k0 + kColRows - 1) / kColRows, kColRows, 0, s>>>(n, L.A, L.xv, L.part2, k0,
                                                                     act, st, m);
    for (int j = 0; j < nb; ++j) {
      const int c = k0 + j;
      const int rows = n - c - 1;
      const dim3 pg((rows + 4 * kProdWaves - 1) / (4 * kProdWaves),
                    (rows + kLargeChunkCols - 1) / kLargeChunkCols);
      k_lt_prod<<<pg, 64 * kProdWaves, 0, s>>>(n, L.A, L.V, L.W, L.xv, L.part2, L.part, L.pyp,
                                               L.partvp, d, e, tau, c, j, act, st, m);
      k_lt_cols<<<(rows + kColRows - 1) / kColRows, kColRows, 0, s>>>(
          n, L.A, L.V, L.W, L.pyp, L.part, L.partvp, tau, L.xv, L.part2, c, j, j + 1 < nb, act,
          st, m);
    }
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
  k_bt_large<<<(k + 15) / 16, 64 * kBtWaves, 0, s>>>(n, nref, k, L.A, L.T, L.Y, act, b.status, m);
}

}  // namespace sg
