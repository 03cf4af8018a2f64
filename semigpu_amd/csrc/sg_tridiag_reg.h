// sg_tridiag_reg.h -- Householder tridiagonalisation of a trailing block held
// on chip by one workgroup (included by sg_linalg.hip after its cross-lane
// helpers).
//
// The one-workgroup reduction of k_tridiag reads the trailing block from
// global memory once per column (the matrix-vector product), which makes it
// bound by HBM: about N^3/6 doubles per molecule.  Here the last MR x MR block
// (MR = 32 NT) stays in the vector registers and LDS of eight wavefronts, so a
// column costs no memory traffic beyond its reflector.
//
// Layout: 32 x 32 tiles (I, J), J <= I, numbered column by column over an
// NT x NT grid.  A molecule's block of m0 rows takes the last m0 rows and
// columns of the grid (grid index = block index + MR - m0), so small
// molecules use the bottom right tiles.  The workgroup has NW wavefronts; the
// last NW * NR tiles live in registers (tile (first register tile) + w + NW s
// in slot s of wavefront w), the first ones (the first to drop out as the
// reduction advances, used only by molecules that fill the grid) in LDS, tile
// t with wavefront t % NW.  Lane l = 8 g_r + g_c holds the 4 x 4 block at rows
// 4 g_r .. 4 g_r + 3 and columns 4 g_c .. 4 g_c + 3 of a tile.  Diagonal tiles
// are held in full (both triangles), so their row sums are their whole
// contribution to the product.  Each wavefront keeps its own copy of v and w
// in LDS, written at the start of the product and of the update (no barrier:
// only that wavefront reads it).  The launches are bucketed by N, each bucket
// with the smallest grid that holds its molecules and as few wavefronts as
// its tiles need, so that several small molecules share a CU.
//
// Per column (unblocked, Golub & Van Loan 8.3.1, as LAPACK dsytd2):
//   x = A[c:, c] (LDS, written by the lanes that hold column c),
//   alpha = x_1, s2 = |x_2:|^2, beta = -sign(alpha) sqrt(alpha^2 + s2),
//   tau = (beta - alpha) / beta, v = x / (alpha - beta) with v_1 = 1;
//   p = tau A22 v: per tile, row sums (reduced over the 8 lanes of a lane row
//   by DPP) and, below the diagonal, column sums (over the 8 lanes of a lane
//   column by permlane swaps and DPP), added per row in a fixed order;
//   w = p - (tau/2)(p.v) v;  A22 -= v w^T + w v^T on chip.
// Three barriers per column; every sum runs in a fixed order, so a molecule's
// result does not depend on the rest of its launch.
//
// A molecule with N > MR has its first N - MR columns reduced by k_tridiag
// (stopped there, the trailing block updated in the work matrix); this kernel
// then loads the trailing block and reduces the rest.  Reflector c goes to the
// upper part of row c of the work matrix and d, e, tau to the molecule's
// vectors, as k_tridiag leaves them.
#pragma once
#include <type_traits>

namespace sg {

constexpr int kDppRor8 = 0x128;  // row_ror:8, lane i <-> i ^ 8 within 16

// Tile coordinates of tile number t (column by column over the lower triangle
// of an NT x NT grid).
template <int NT>
__device__ __forceinline__ void tile_of(int t, int& I, int& J) {
  I = -1;
  J = 0;
  int base = 0;
  for (int j = 0; j < NT; ++j) {
    const int len = NT - j;
    if (t >= base && t < base + len) {
      I = j + (t - base);
      J = j;
    }
    base += len;
  }
}
template <int NT>
__host__ __device__ constexpr int tile_id(int I, int J) {
  return J * NT - J * (J - 1) / 2 + (I - J);
}

// NT x NT tiles (MR = 32 NT) over NW wavefronts with NR register tiles each;
// the rest of the tiles in LDS.
template <int NT, int NW, int NR>
struct RegTiles {
  static constexpr int kRows = 32 * NT;
  static constexpr int kTiles = NT * (NT + 1) / 2;
  static constexpr int kReg = NW * NR < kTiles ? NW * NR : kTiles;  // register tiles
  static constexpr int kLds = kTiles - kReg;                         // LDS tiles: 0 .. kLds - 1
  static constexpr int kLdsSlots = (kLds + NW - 1) / NW;             // per wavefront
  // waves per SIMD the register budget allows (32 registers per tile slot)
  static constexpr int kWaves = NR >= 4 ? 2 : (NR >= 3 ? 3 : 4);
};

template <int NT, int NW, int NR>
__global__ void __launch_bounds__(64 * NW)
__attribute__((amdgpu_waves_per_eu(RegTiles<NT, NW, NR>::kWaves)))
k_tridiag_reg(const MolDesc* __restrict__ mols, const double* __restrict__ F,
              double* __restrict__ work, double* __restrict__ dA, double* __restrict__ eA,
              double* __restrict__ tauA, const int* __restrict__ active,
              const int* __restrict__ status, const int* __restrict__ order,
              const int* __restrict__ prec) {
  using T = RegTiles<NT, NW, NR>;
  constexpr int MR = T::kRows, NTILE = T::kTiles, NL = T::kLdsSlots;
  constexpr int NTH = 64 * NW;
  extern __shared__ double lds[];
  // LDS tiles: element q of lane l of tile t at (t * 8 + q / 2) * 128 + 2 l + q % 2
  double* ltile = lds;
  double* xbuf0 = ltile + (size_t)T::kLds * 1024;  // the current column (double-buffered)
  double* xbuf1 = xbuf0 + MR;
  double* pbuf = xbuf1 + MR;            // p = tau A22 v
  double* prow = pbuf + MR;             // per tile: row sums (32)
  double* pcol = prow + NTILE * 32;     // per tile below the diagonal: column sums (32)
  double* pvw = pcol + NTILE * 32;      // per-wavefront parts of p.v
  double* vloc = pvw + 16;               // per wavefront: v (MR), then w (MR)
  const int m = order ? order[blockIdx.x] : blockIdx.x;
  if (!mol_live(active, status, m)) return;
  if (prec && prec[m]) return;  // reduced in FP32 by k_tridiag<., float>
  const MolDesc M = mols[m];
  const int n = M.n_ao;
  const int m0 = n > MR ? MR : n;  // rows of the block here
  const int off = MR - m0;         // grid index of the block's first row
  const int gbase = n - MR;        // global index = gbase + grid index
  // the wavefront index as a scalar, so that the tile bookkeeping and the
  // tile branches stay in scalar registers
  const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int gr = lane >> 3, gc = lane & 7;
  double* A = work + 5 * M.mat0;
  const double* src = n > MR ? A : F + M.mat0;
  double* d = dA + M.vec0;
  double* e = eA + M.vec0;
  double* tau = tauA + M.vec0;

  // tiles of this wavefront (I = -1: none)
  int rI[NR], rJ[NR], lI[NL > 0 ? NL : 1], lJ[NL > 0 ? NL : 1];
#pragma unroll
  for (int s = 0; s < NR; ++s) {
    const int t = T::kLds + wave + NW * s;
    if (t < NTILE) tile_of<NT>(t, rI[s], rJ[s]);
    else rI[s] = -1;
  }
#pragma unroll
  for (int s = 0; s < NL; ++s) {
    const int t = wave + NW * s;
    if (t < T::kLds) tile_of<NT>(t, lI[s], lJ[s]);
    else lI[s] = -1;
  }
  const auto fetch = [&](int I, int J, int q) __attribute__((always_inline)) -> double {
    const int row = 32 * I + 4 * gr + (q >> 2), col = 32 * J + 4 * gc + (q & 3);
    if (I < 0 || row < off || col < off) return 0.0;
    const int64_t gi = gbase + row, gj = gbase + col;
    return col <= row ? src[gi * n + gj] : src[gj * n + gi];
  };
  double a[NR][16];
#pragma unroll
  for (int s = 0; s < NR; ++s)
#pragma unroll
    for (int q = 0; q < 16; ++q) a[s][q] = fetch(rI[s], rJ[s], q);
#pragma unroll
  for (int s = 0; s < NL; ++s)
    if (lI[s] >= 0 && 32 * lI[s] + 31 >= off) {
      double* tp = ltile + (size_t)(wave + NW * s) * 1024;
#pragma unroll
      for (int q = 0; q < 16; ++q) tp[(q >> 1) * 128 + 2 * lane + (q & 1)] = fetch(lI[s], lJ[s], q);
    }
  for (int i = tid; i < MR; i += NTH) {
    xbuf0[i] = 0.0;
    xbuf1[i] = 0.0;
    pbuf[i] = 0.0;
  }
  for (int i = tid; i < 2 * NW * MR; i += NTH) vloc[i] = 0.0;
  __syncthreads();
  // the block's first column: x[i] = A(i, off) for i >= off
  for (int i = off + tid; i < MR; i += NTH) {
    const int64_t gi = gbase + i, gj = gbase + off;
    xbuf0[i] = src[gi * n + gj];
  }
  __syncthreads();

  // tile numbers of this wavefront's slots
  int rid[NR], lid[NL > 0 ? NL : 1];
#pragma unroll
  for (int s = 0; s < NR; ++s) rid[s] = T::kLds + wave + NW * s;
#pragma unroll
  for (int s = 0; s < NL; ++s) lid[s] = wave + NW * s;
  double* vw = vloc + (size_t)wave * 2 * MR;  // this wavefront's v; w at vw + MR
  // a tile takes part while one of its columns is above k
  const auto live = [](int I, int J, int k) __attribute__((always_inline)) {
    return I >= 0 && 32 * J + 31 > k;
  };
  const auto ld4 = [](const double* p, double (&v)[4]) __attribute__((always_inline)) {
    const double2 u = *(const double2*)p, w = *(const double2*)(p + 2);
    v[0] = u.x;
    v[1] = u.y;
    v[2] = w.x;
    v[3] = w.y;
  };
  // row and column sums of the product over one tile, rows R0 .. R0 + NRW - 1
  // of the lanes' 4 x 4 blocks from t (an LDS tile goes by halves, so that
  // only 8 of its values are in registers at a time)
  const auto acc_rows = [&](auto r0c, auto nrc, const double* t, const double (&vr)[4],
                            const double (&vc)[4], double (&rp)[4], double (&cp)[4])
      __attribute__((always_inline)) {
    constexpr int R0 = decltype(r0c)::value, NRW = decltype(nrc)::value;
#pragma unroll
    for (int ri = R0; ri < R0 + NRW; ++ri)
#pragma unroll
      for (int ci = 0; ci < 4; ++ci) {
        const double u = t[4 * (ri - R0) + ci];
        rp[ri] = ci == 0 ? u * vc[0] : fma(u, vc[ci], rp[ri]);
        cp[ci] = ri == 0 ? u * vr[0] : fma(u, vr[ri], cp[ci]);
      }
  };
  const auto reduce_write = [&](int I, int J, int id, const double (&rp)[4], const double (&cp)[4])
      __attribute__((always_inline)) {
    // the tile's slot offset made opaque here, so that the compiler forms the
    // LDS addresses below at each use instead of keeping one per tile in
    // registers across the column loop (which spills them)
    asm volatile("" : "+s"(id));
    // rows: over the 8 lanes of a lane row (bits 2, 1, 0 of the lane)
    {
      const bool b2 = lane & 4, b1 = lane & 2;
      double h[2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const double keep = b2 ? rp[q + 2] : rp[q], send = b2 ? rp[q] : rp[q + 2];
        h[q] = keep + dpp_f64<kDppMirror8>(send);
      }
      const double keep = b1 ? h[1] : h[0], send = b1 ? h[0] : h[1];
      double r1 = keep + dpp_f64<kDppXor2>(send);
      r1 += dpp_f64<kDppXor1>(r1);
      if ((lane & 1) == 0) prow[id * 32 + 4 * gr + 2 * (int)b2 + (int)b1] = r1;
    }
    // columns (below the diagonal): over the 8 lanes of a lane column
    // (bits 5, 4, 3)
    if (I != J) {
      double x0 = cp[0], x2 = cp[2], x1 = cp[1], x3 = cp[3];
      swap32(x0, x2);
      swap32(x1, x3);
      double y0 = x0 + x2, y1 = x1 + x3;
      swap16(y0, y1);
      double z = y0 + y1;
      z += dpp_f64<kDppRor8>(z);
      if ((lane & 8) == 0) pcol[id * 32 + 4 * gc + 2 * ((lane >> 5) & 1) + ((lane >> 4) & 1)] = z;
    }
  };
  // A22 -= v w^T + w v^T over rows R0 .. R0 + NRW - 1 of a tile, and column
  // k + 1 for the next step
  const auto upd_rows = [&](auto r0c, auto nrc, double* t, int I, int J, const double (&vr)[4],
                            const double (&vc)[4], const double (&wr)[4], const double (&wc)[4],
                            int kn, double* xn) __attribute__((always_inline)) {
    constexpr int R0 = decltype(r0c)::value, NRW = decltype(nrc)::value;
#pragma unroll
    for (int ri = R0; ri < R0 + NRW; ++ri)
#pragma unroll
      for (int ci = 0; ci < 4; ++ci) {
        double u = t[4 * (ri - R0) + ci];
        u = fma(-vr[ri], wc[ci], u);
        u = fma(-wr[ri], vc[ci], u);
        t[4 * (ri - R0) + ci] = u;
      }
    if (J == (kn >> 5) && gc == ((kn & 31) >> 2)) {
      const int cn = kn & 3;
#pragma unroll
      for (int ri = R0; ri < R0 + NRW; ++ri) {
        const int row = 32 * I + 4 * gr + ri;
        const double a0 = t[4 * (ri - R0)], a1 = t[4 * (ri - R0) + 1], a2 = t[4 * (ri - R0) + 2],
                     a3 = t[4 * (ri - R0) + 3];
        const double val = cn == 0 ? a0 : (cn == 1 ? a1 : (cn == 2 ? a2 : a3));
        if (row >= kn) xn[row] = val;
      }
    }
  };
  using C0 = std::integral_constant<int, 0>;
  using C2 = std::integral_constant<int, 2>;
  using C4 = std::integral_constant<int, 4>;

  int buf = 0;
  for (int k = off; k + 2 < MR; ++k) {
    const int c = gbase + k, kn = k + 1;
    const double* x = buf ? xbuf1 : xbuf0;
    double* xn = buf ? xbuf0 : xbuf1;
    // |x_2:|^2: the same sum in every wavefront
    constexpr int NJ = (MR + 63) / 64;  // lane strides over a column
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const int i = k + 2 + lane + 64 * j;
      const double xi = i < MR ? x[i] : 0.0;
      s = fma(xi, xi, s);
    }
    const double s2 = wsum(s);
    const double alpha = x[kn];
    // s2 = 0: no reflection (tau = 0, v = e_1, so p = w = 0)
    const double beta = s2 == 0.0 ? alpha : -copysign(sqrt(alpha * alpha + s2), alpha);
    const double tk = s2 == 0.0 ? 0.0 : (beta - alpha) / beta;
    const double sc = s2 == 0.0 ? 0.0 : 1.0 / (alpha - beta);
    if (tid == 0) {
      d[c] = x[k];
      e[c] = beta;
      tau[c] = tk;
    }
    // v (0 at or before k, 1 at k + 1): this wavefront's copy, and the
    // reflector into row c of the work matrix
    // (entries before k were cleared at their own column)
    {
      // all loads first: the stores could alias them for the compiler
      double xv[NJ];
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        const int i = k + lane + 64 * j;
        xv[j] = i < MR ? x[i] : 0.0;
      }
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        const int i = k + lane + 64 * j;
        if (i < MR) {
          const double v = i > kn ? xv[j] * sc : (i == kn ? 1.0 : 0.0);
          vw[i] = v;
          // the reflector's 64-entry chunk j is stored by wavefront j % NW
          if (i > k && j % NW == wave) A[(int64_t)c * n + gbase + i] = v;
        }
      }
    }
    // p: row and column sums per tile
#pragma unroll
    for (int sl = 0; sl < NR; ++sl)
      if (live(rI[sl], rJ[sl], k)) {
        const int I = rI[sl], J = rJ[sl];
        double vr[4], vc[4], rp[4], cp[4];
        ld4(vw + 32 * I + 4 * gr, vr);
        ld4(vw + 32 * J + 4 * gc, vc);
        acc_rows(C0(), C4(), a[sl], vr, vc, rp, cp);
        reduce_write(I, J, rid[sl], rp, cp);
      }
#pragma unroll
    for (int sl = 0; sl < NL; ++sl)
      if (live(lI[sl], lJ[sl], k)) {
        const int I = lI[sl], J = lJ[sl];
        const double* tp = ltile + (size_t)lid[sl] * 1024 + 2 * lane;
        double vr[4], vc[4], rp[4], cp[4], t[8];
        ld4(vw + 32 * I + 4 * gr, vr);
        ld4(vw + 32 * J + 4 * gc, vc);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const double2 u = *(const double2*)(tp + (4 * h + q) * 128);
            t[2 * q] = u.x;
            t[2 * q + 1] = u.y;
          }
          if (h == 0) acc_rows(C0(), C2(), t, vr, vc, rp, cp);
          else acc_rows(C2(), C2(), t, vr, vc, rp, cp);
        }
        reduce_write(I, J, lid[sl], rp, cp);
      }
    __syncthreads();
    double pvp = 0.0;
    for (int i = kn + tid; i < MR; i += NTH) {
      const int I = i >> 5, r = i & 31, J0 = kn >> 5;
      // row sums of tiles (I, J0 .. I), then column sums of tiles (I + 1 .. NT - 1, I),
      // in that order (the unrolled loads can issue together)
      double sum = 0.0;
#pragma unroll
      for (int q = 0; q < NT; ++q) {
        const int jr = q <= I ? q : I;  // clamped: the skipped loads stay inside the array
        const double u = prow[tile_id<NT>(I, jr) * 32 + r];
        sum += (q >= J0 && q <= I) ? u : 0.0;
      }
#pragma unroll
      for (int q = 1; q < NT; ++q) {
        const int ic = q > I ? q : NT - 1;
        const double u = pcol[tile_id<NT>(ic, I < NT - 1 ? I : 0) * 32 + r];
        sum += q > I ? u : 0.0;
      }
      const double p = tk * sum;
      pbuf[i] = p;
      pvp += p * vw[i];
    }
    pvp = wsum(pvp);
    if (lane == 0) pvw[wave] = pvp;
    __syncthreads();
    double pv = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) pv += pvw[w];
    const double al = -0.5 * tk * pv;
    // w = p + al v (0 at or before k): this wavefront's copy
    {
      double pv_[NJ], vv_[NJ];
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        const int i = k + lane + 64 * j;
        pv_[j] = i < MR ? pbuf[i] : 0.0;
        vv_[j] = i < MR ? vw[i] : 0.0;
      }
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        const int i = k + lane + 64 * j;
        if (i < MR) vw[MR + i] = i > k ? pv_[j] + al * vv_[j] : 0.0;
      }
    }
    // A22 -= v w^T + w v^T, and column k + 1 for the next step
#pragma unroll
    for (int sl = 0; sl < NR; ++sl)
      if (live(rI[sl], rJ[sl], k)) {
        const int I = rI[sl], J = rJ[sl];
        double vr[4], vc[4], wr[4], wc[4];
        ld4(vw + 32 * I + 4 * gr, vr);
        ld4(vw + 32 * J + 4 * gc, vc);
        ld4(vw + MR + 32 * I + 4 * gr, wr);
        ld4(vw + MR + 32 * J + 4 * gc, wc);
        upd_rows(C0(), C4(), a[sl], I, J, vr, vc, wr, wc, kn, xn);
      }
#pragma unroll
    for (int sl = 0; sl < NL; ++sl)
      if (live(lI[sl], lJ[sl], k)) {
        const int I = lI[sl], J = lJ[sl];
        double* tp = ltile + (size_t)lid[sl] * 1024 + 2 * lane;
        double vr[4], vc[4], wr[4], wc[4], t[8];
        ld4(vw + 32 * I + 4 * gr, vr);
        ld4(vw + 32 * J + 4 * gc, vc);
        ld4(vw + MR + 32 * I + 4 * gr, wr);
        ld4(vw + MR + 32 * J + 4 * gc, wc);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const double2 u = *(const double2*)(tp + (4 * h + q) * 128);
            t[2 * q] = u.x;
            t[2 * q + 1] = u.y;
          }
          if (h == 0) upd_rows(C0(), C2(), t, I, J, vr, vc, wr, wc, kn, xn);
          else upd_rows(C2(), C2(), t, I, J, vr, vc, wr, wc, kn, xn);
#pragma unroll
          for (int q = 0; q < 4; ++q)
            *(double2*)(tp + (4 * h + q) * 128) = make_double2(t[2 * q], t[2 * q + 1]);
        }
      }
    __syncthreads();
    buf ^= 1;
  }
  // the last 2 x 2 block: column MR - 2 is the last published column, and
  // entry (MR - 1, MR - 1) is entry 15 of lane 63 in the last register tile
  if (tid == 0 && m0 >= 2) {
    const double* x = buf ? xbuf1 : xbuf0;
    d[n - 2] = x[MR - 2];
    e[n - 2] = x[MR - 1];
  }
#pragma unroll
  for (int sl = 0; sl < NR; ++sl)
    if (lane == 63 && rI[sl] == NT - 1 && rJ[sl] == NT - 1) {
      d[n - 1] = a[sl][15];
      e[n - 1] = 0.0;
    }
}

// Dynamic LDS of k_tridiag_reg<NT, NW, NR>: the LDS tiles, two column
// buffers, p, the per-tile row and column sums, the per-wavefront dots and
// the per-wavefront copies of v and w.
template <int NT, int NW, int NR>
constexpr size_t tridiag_reg_lds() {
  using T = RegTiles<NT, NW, NR>;
  return sizeof(double) * ((size_t)T::kLds * 1024 + 3 * T::kRows + 64 * T::kTiles + 16 +
                           2 * NW * (size_t)T::kRows);
}

}  // namespace sg
