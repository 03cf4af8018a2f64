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

namespace sg {

constexpr int kCoreChunk = 2048;  // core-core pair terms per round of k_core_core

// ---------------------------------------------------------------------------
// Multipole interactions in Hartree; (R, element of first pair, element of
// second pair).  Same formulas and evaluation order as multipole.cpp.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double pp_(int z, int l) {
  return l == 0 ? c_par.pp0[z] : (l == 1 ? c_par.pp1[z] : c_par.pp2[z]);
}
__device__ __forceinline__ double addt(int a, int l, int b, int m) {
  double t = pp_(a, l) + pp_(b, m);
  return t * t;
}
__device__ __forceinline__ double rcp(double x) { return 1.0 / x; }

__device__ double mp_qq(double R, int a, int b) { return 1.0 / sqrt(R * R + addt(a, 0, b, 0)); }
__device__ double mp_qUz(double R, int a, int b) {
  double t = addt(a, 0, b, 1), f = 0.0, x;
  x = R + c_par.dd0[b]; f += rcp(2.0 * sqrt(x * x + t));
  x = R - c_par.dd0[b]; f -= rcp(2.0 * sqrt(x * x + t));
  return f;
}
__device__ double mp_qQpipi(double R, int a, int b) {
  double t = addt(a, 0, b, 2), f = 0.0, x;
  x = (2.0 * c_par.dd1[b]) * (2.0 * c_par.dd1[b]);
  f += rcp(2.0 * sqrt(R * R + x + t));
  f -= rcp(2.0 * sqrt(R * R + t));
  return f;
}
__device__ double mp_qQzz(double R, int a, int b) {
  double t = addt(a, 0, b, 2), f = 0.0, x;
  x = R + 2.0 * c_par.dd1[b]; f += rcp(4.0 * sqrt(x * x + t));
  f -= rcp(2.0 * sqrt(R * R + t));
  x = R - 2.0 * c_par.dd1[b]; f += rcp(4.0 * sqrt(x * x + t));
  return f;
}
__device__ double mp_UpiUpi(double R, int a, int b) {
  double t = addt(a, 1, b, 1), f = 0.0, x;
  x = c_par.dd0[a] - c_par.dd0[b]; f += rcp(2.0 * sqrt(R * R + x * x + t));
  x = c_par.dd0[a] + c_par.dd0[b]; f -= rcp(2.0 * sqrt(R * R + x * x + t));
  return f;
}
__device__ double mp_UzUz(double R, int a, int b) {
  double t = addt(a, 1, b, 1), f = 0.0, x;
  const double da = c_par.dd0[a], db = c_par.dd0[b];
  x = da + db + R;    f += -1.0 * rcp(4.0 * sqrt(x * x + t));
  x = da + -db + R;   f += 1.0 * rcp(4.0 * sqrt(x * x + t));
  x = -da + db + R;   f += 1.0 * rcp(4.0 * sqrt(x * x + t));
  x = -da + -db + R;  f += -1.0 * rcp(4.0 * sqrt(x * x + t));
  return f;
}
__device__ double mp_UpiQpiz(double R, int a, int b) {
  double t = addt(a, 1, b, 2), f = 0.0, x, y;
  const double da = c_par.dd0[a], db = c_par.dd1[b];
  const double sg[2] = {1.0, -1.0}, sn[3] = {-1.0, 1.0, -1.0};
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      x = R + sg[i] * db;
      y = da + sg[j] * db;
      f += sn[i + j] * rcp(4.0 * sqrt(x * x + y * y + t));
    }
  return f;
}
__device__ double mp_UzQpipi(double R, int a, int b) {
  double t = addt(a, 1, b, 2), f = 0.0, x, y;
  const double da = c_par.dd0[a], db = c_par.dd1[b];
  const double sg[2] = {1.0, -1.0}, sn[2] = {-1.0, 1.0};
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    x = R + sg[i] * da;
    y = 2.0 * db;
    f += sn[i] * rcp(4.0 * sqrt(x * x + y * y + t));
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    x = R + sg[i] * da;
    f += sg[i] * rcp(4.0 * sqrt(x * x + t));
  }
  return f;
}
__device__ double mp_UzQz(double R, int a, int b) {
  double t = addt(a, 1, b, 2), f = 0.0, x;
  const double da = c_par.dd0[a], db = c_par.dd1[b];
  const double sg[2] = {1.0, -1.0}, sn[4] = {-1.0, -1.0, 1.0, 1.0};
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      x = R + sg[i] * da + sg[j] * 2.0 * db;
      f += sn[2 * i + j] * rcp(8.0 * sqrt(x * x + t));
    }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    x = R + sg[i] * da;
    f += sg[i] * rcp(4.0 * sqrt(x * x + t));
  }
  return f;
}
__device__ double mp_QpipiQpipi(double R, int a, int b) {
  double t = addt(a, 2, b, 2), f = 0.0, x, y;
  const double da = c_par.dd1[a], db = c_par.dd1[b];
  y = da + db;  x = 4.0 * (y * y); f += rcp(8.0 * sqrt(R * R + x + t));
  y = da + -db; x = 4.0 * (y * y); f += rcp(8.0 * sqrt(R * R + x + t));
  y = 2.0 * da; x = y * y; f -= rcp(4.0 * sqrt(R * R + x + t));
  y = 2.0 * db; x = y * y; f -= rcp(4.0 * sqrt(R * R + x + t));
  f += rcp(4.0 * sqrt(R * R + t));
  return f;
}
__device__ double mp_QxxQyy(double R, int a, int b) {
  double t = addt(a, 2, b, 2), f = 0.0, x;
  const double ya = 2.0 * c_par.dd1[a], yb = 2.0 * c_par.dd1[b];
  x = ya * ya + yb * yb; f += rcp(4.0 * sqrt(R * R + x + t));
  x = ya * ya;           f -= rcp(4.0 * sqrt(R * R + x + t));
  x = yb * yb;           f -= rcp(4.0 * sqrt(R * R + x + t));
  f += rcp(4.0 * sqrt(R * R + t));
  return f;
}
__device__ double mp_QpipiQzz(double R, int a, int b) {
  double t = addt(a, 2, b, 2), f = 0.0, aux, auxx, x;
  const double ya = 2.0 * c_par.dd1[a], db = c_par.dd1[b];
  auxx = ya * ya + t;
  const double sg[2] = {1.0, -1.0};
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    x = R + 2.0 * sg[i] * db;
    aux = x * x;
    f += rcp(8.0 * sqrt(aux + auxx));
    f -= rcp(8.0 * sqrt(aux + t));
    f -= rcp(8.0 * sqrt(R * R + auxx));
  }
  f += rcp(4.0 * sqrt(R * R + t));
  return f;
}
__device__ double mp_QzzQzz(double R, int a, int b) {
  double t = addt(a, 2, b, 2), f = 0.0, x;
  const double da = c_par.dd1[a], db = c_par.dd1[b];
  const double sg[2] = {1.0, -1.0};
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      x = R + sg[i] * 2.0 * da + sg[j] * 2.0 * db;
      f += rcp(16.0 * sqrt(x * x + t));
    }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    x = R + 2.0 * sg[i] * da; f -= rcp(8.0 * sqrt(x * x + t));
    x = R + 2.0 * sg[i] * db; f -= rcp(8.0 * sqrt(x * x + t));
  }
  f += rcp(4.0 * sqrt(R * R + t));
  return f;
}
__device__ double mp_QpizQpiz(double R, int a, int b) {
  double t = addt(a, 2, b, 2), f = 0.0, x, y;
  const double da = c_par.dd1[a], db = c_par.dd1[b];
  const double sg[2] = {1.0, -1.0};
  const double sn[8] = {1.0, -1.0, -1.0, 1.0, -1.0, 1.0, 1.0, -1.0};
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        x = R + sg[i] * da + sg[j] * db;
        y = da + sg[k] * db;
        f += sn[4 * i + 2 * j + k] * rcp(8.0 * sqrt(x * x + y * y + t));
      }
  return f;
}

// Local frame: row 2 = unit vector A->C, rows 0/1 complete it
// (ComputeRotationMatrix, twocenterintegral.cpp:468-490, including its fixed
// frame when A->C is within 1e-10 of the z axis).
__device__ __forceinline__ double dist3(double ax, double ay, double az, double bx, double by,
                                        double bz) {
  double r = 0.0;
  r += (ax - bx) * (ax - bx);
  r += (ay - by) * (ay - by);
  r += (az - bz) * (az - bz);
  return sqrt(r);
}

struct Frame {
  double m[3][3];
};

__device__ __forceinline__ Frame local_frame(const AtomDesc& A, const AtomDesc& C, double R) {
  Frame f;
  f.m[2][0] = (C.x - A.x) / R;
  f.m[2][1] = (C.y - A.y) / R;
  f.m[2][2] = (C.z - A.z) / R;
  if (fabs(f.m[2][0] - 0.0) < 1e-10 && fabs(f.m[2][1] - 0.0) < 1e-10) {
    f.m[1][0] = 0.0; f.m[1][1] = 1.0; f.m[1][2] = 0.0;
    f.m[0][0] = 1.0; f.m[0][1] = 0.0; f.m[0][2] = 0.0;
  } else {
    double s = 1.0 / sqrt(f.m[2][0] * f.m[2][0] + f.m[2][1] * f.m[2][1]);
    f.m[1][0] = s * -f.m[2][1];
    f.m[1][1] = s * f.m[2][0];
    f.m[1][2] = 0.0;
    f.m[0][0] = f.m[2][1] * f.m[1][2] - f.m[2][2] * f.m[1][1];
    f.m[0][1] = f.m[2][2] * f.m[1][0] - f.m[2][0] * f.m[1][2];
    f.m[0][2] = f.m[2][0] * f.m[1][1] - f.m[2][1] * f.m[1][0];
  }
  return f;
}

// The local-frame combinations an X-X pair needs, each formed once with the
// same additions as the reference's IntegralType* functions.
struct LocalXX {
  double qq, qUz, Uzq;
  double sspp_pi, sspp_z, ppss_pi, ppss_z;   // SS|PP and PP|SS (inverted) sums
  double sppi, spzz;                          // SP|SP: UpiUpi, UzUz
  double sppp1, sppp2, sppp3;                 // SP|PP
  double ppsp1, ppsp2, ppsp3;                 // PP|SP (inverted)
  double zzzz, pppp, zzpp, ppzz, pzpz, xxyy, xyxy;
};

__device__ LocalXX local_xx(double R, int a, int c) {
  LocalXX v;
  v.qq = mp_qq(R, a, c);
  v.qUz = mp_qUz(R, a, c);
  v.Uzq = mp_qUz(R, c, a) * -1.0;
  const double qQpp = mp_qQpipi(R, a, c), Qppq = mp_qQpipi(R, c, a);
  const double qQzz = mp_qQzz(R, a, c), Qzzq = mp_qQzz(R, c, a);
  v.sspp_pi = v.qq + qQpp;
  v.sspp_z = v.qq + qQzz;
  v.ppss_pi = v.qq + Qppq;
  v.ppss_z = v.qq + Qzzq;
  v.sppi = mp_UpiUpi(R, a, c);
  v.spzz = mp_UzUz(R, a, c);
  v.sppp1 = v.Uzq + mp_UzQpipi(R, a, c);
  v.sppp2 = v.Uzq + mp_UzQz(R, a, c);
  v.sppp3 = mp_UpiQpiz(R, a, c);
  v.ppsp1 = v.qUz + mp_UzQpipi(R, c, a) * -1.0;
  v.ppsp2 = v.qUz + mp_UzQz(R, c, a) * -1.0;
  v.ppsp3 = mp_UpiQpiz(R, c, a) * -1.0;
  v.zzzz = v.qq + qQzz + Qzzq + mp_QzzQzz(R, a, c);
  v.pppp = v.qq + qQpp + Qppq + mp_QpipiQpipi(R, a, c);
  v.zzpp = v.qq + qQpp + Qzzq + mp_QpipiQzz(R, c, a);
  v.ppzz = v.qq + Qppq + qQzz + mp_QpipiQzz(R, a, c);
  v.pzpz = mp_QpizQpiz(R, a, c);
  v.xxyy = v.qq + qQpp + Qppq + mp_QxxQyy(R, a, c);
  v.xyxy = (v.pppp - v.xxyy) * 0.5;
  return v;
}

// Rotated entry (pair K of A | pair L of C) in eV; K, L in MOPAC order.
template <int K, int L>
__device__ __forceinline__ double xx_entry(const LocalXX& v, const Frame& F) {
  constexpr int s0 = mop_a(K), s1 = mop_b(K), s2 = mop_a(L), s3 = mop_b(L);
  constexpr int pa0 = (s0 > 0) + (s1 > 0), pb0 = (s2 > 0) + (s3 > 0);
  constexpr bool inv = pa0 > pb0;  // CorrectOrderIntegral (:785-796)
  constexpr int pa = inv ? pb0 : pa0, pb = inv ? pa0 : pb0;
  constexpr int t0 = inv ? s2 : s0, t1 = inv ? s3 : s1, t2 = inv ? s0 : s2, t3 = inv ? s1 : s3;
  const double(&m)[3][3] = F.m;
  double r;
  if constexpr (pa == 0 && pb == 0) {
    r = v.qq;
  } else if constexpr (pa == 0 && pb == 1) {
    r = (inv ? v.Uzq : v.qUz) * m[2][t3 - 1];
  } else if constexpr (pa == 0 && pb == 2) {
    double vp = inv ? v.ppss_pi : v.sspp_pi, vz = inv ? v.ppss_z : v.sspp_z;
    vp = vp * m[0][t2 - 1] * m[0][t3 - 1] + vp * m[1][t2 - 1] * m[1][t3 - 1];
    vz = vz * m[2][t2 - 1] * m[2][t3 - 1];
    r = vp + vz;
  } else if constexpr (pa == 1 && pb == 1) {
    double v1 = v.sppi * m[0][t1 - 1] * m[0][t3 - 1] + v.sppi * m[1][t1 - 1] * m[1][t3 - 1];
    double v2 = v.spzz * m[2][t1 - 1] * m[2][t3 - 1];
    r = v1 + v2;
  } else if constexpr (pa == 1 && pb == 2) {
    const double w1 = inv ? v.ppsp1 : v.sppp1, w2 = inv ? v.ppsp2 : v.sppp2,
                 w3 = inv ? v.ppsp3 : v.sppp3;
    double v1 = w1 * m[2][t1 - 1] * m[0][t2 - 1] * m[0][t3 - 1] +
                w1 * m[2][t1 - 1] * m[1][t2 - 1] * m[1][t3 - 1];
    double v2 = w2 * m[2][t1 - 1] * m[2][t2 - 1] * m[2][t3 - 1];
    double v3 = w3 * m[0][t1 - 1] * m[2][t2 - 1] * m[0][t3 - 1] +
                w3 * m[1][t1 - 1] * m[2][t2 - 1] * m[1][t3 - 1] +
                w3 * m[0][t1 - 1] * m[2][t3 - 1] * m[0][t2 - 1] +
                w3 * m[1][t1 - 1] * m[2][t3 - 1] * m[1][t2 - 1];
    r = v1 + v2 + v3;
  } else {
    const int a = t0 - 1, b = t1 - 1, c = t2 - 1, d = t3 - 1;
    double zz = v.zzzz * m[2][a] * m[2][b] * m[2][c] * m[2][d];
    double pp = v.pppp * m[0][a] * m[0][b] * m[0][c] * m[0][d] +
                v.pppp * m[1][a] * m[1][b] * m[1][c] * m[1][d];
    double zp = v.zzpp * m[2][a] * m[2][b] * m[0][c] * m[0][d] +
                v.zzpp * m[2][a] * m[2][b] * m[1][c] * m[1][d];
    double pz = v.ppzz * m[0][a] * m[0][b] * m[2][c] * m[2][d] +
                v.ppzz * m[1][a] * m[1][b] * m[2][c] * m[2][d];
    double q = v.pzpz;
    double pzpz = q * m[0][a] * m[2][b] * m[0][c] * m[2][d] + q * m[0][a] * m[2][b] * m[0][d] * m[2][c] +
                  q * m[0][b] * m[2][a] * m[0][c] * m[2][d] + q * m[0][b] * m[2][a] * m[0][d] * m[2][c] +
                  q * m[1][a] * m[2][b] * m[1][c] * m[2][d] + q * m[1][a] * m[2][b] * m[1][d] * m[2][c] +
                  q * m[1][b] * m[2][a] * m[1][c] * m[2][d] + q * m[1][b] * m[2][a] * m[1][d] * m[2][c];
    double xxyy = v.xxyy * m[0][a] * m[0][b] * m[1][c] * m[1][d] +
                  v.xxyy * m[1][a] * m[1][b] * m[0][c] * m[0][d];
    double w = v.xyxy;
    double xyxy = w * m[0][a] * m[1][b] * m[0][c] * m[1][d] + w * m[0][a] * m[1][b] * m[0][d] * m[1][c] +
                  w * m[0][b] * m[1][a] * m[0][c] * m[1][d] + w * m[0][b] * m[1][a] * m[0][d] * m[1][c];
    r = zz + pp + zp + pz + pzpz + xxyy + xyxy;
  }
  return r * kHartreeToEv;
}

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

// ---------------------------------------------------------------------------
// Off-diagonal Hcore blocks beta * S / 2 for one atom pair per lane, launched
// per pair type so the lanes of a wave do the same work.  The STO-6G overlap
// is Overlap::ComputeOverlap (overlap.cpp:68-134): a sum over the 6 x 6
// primitive pairs of exp(-mu R^2) times one Gaussian moment per axis.  The AO
// pairs of one shell pair (s-s, s-p, p-s, p-p) share their primitive
// exponents, so exp(-mu R^2), the Gaussian centre and the axis moments are
// formed once per primitive pair and shell pair.
template <int NA, int NB>
__device__ __forceinline__ void hcore_shell(const AtomDesc& A, int sa, const AtomDesc& B, int sb,
                                            double ab, double* tot) {
  const double ac[3] = {A.x, A.y, A.z}, bc[3] = {B.x, B.y, B.z};
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      const double al = sa == 0 ? c_sto.expS[A.zn][i] : c_sto.expP[A.zn][i];
      const double be = sb == 0 ? c_sto.expS[B.zn][j] : c_sto.expP[B.zn][j];
      const double p = al + be;
      const double mu = al * be / p;
      double o = exp(-mu * ab);
      o *= sa == 0 ? c_sto.cS[A.zn][i] : c_sto.cP[A.zn][i];
      o *= sb == 0 ? c_sto.cS[B.zn][j] : c_sto.cP[B.zn][j];
      // Per axis, the reference's 12-point Gauss-Hermite sums of 1, (x+pa),
      // (x+pb) and (x+pa)(x+pb) against exp(-p x^2) (overlap.cpp:114-124,
      // 177-192) are exact for these degrees, so their closed forms are used:
      // sqrt(pi/p) {1, pa, pb, pa pb + 1/(2p)}
      const double i00 = sqrt(M_PI / p);
      const double h2 = 0.5 / p;
      double f10[3], f01[3], f11[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        double P = al * ac[k];
        P += be * bc[k];
        P /= p;
        const double pa = P - ac[k], pb = P - bc[k];
        f10[k] = pa * i00;
        f01[k] = pb * i00;
        f11[k] = (pa * pb + h2) * i00;
      }
      // AO u of shell sa (0 = s; 1..3 = px, py, pz when sa = 1), likewise v
#pragma unroll
      for (int u = 0; u < NA; ++u)
#pragma unroll
        for (int v = 0; v < NB; ++v) {
          const int la = sa ? u + 1 : 0, lb = sb ? v + 1 : 0;
          double t = o;
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            const bool pa = la == k + 1, pb = lb == k + 1;
            t *= pa ? (pb ? f11[k] : f10[k]) : (pb ? f01[k] : i00);
          }
          tot[u * NB + v] += t;
        }
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
}

}  // namespace sg
