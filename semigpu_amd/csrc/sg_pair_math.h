// sg_pair_math.h -- the two-center pair terms as functions of the scalar type.
//
// Multipole interactions, the local frame, the rotated (mu nu | lambda sigma)
// entries and the STO-6G overlaps of one atom pair, written once for T =
// double (the integral and Hcore kernels, whose arithmetic is the reference's
// operation order) and for T = ad::Dual, a forward-mode dual number carrying
// one directional derivative, from which the analytic nuclear gradient takes
// the derivatives of every pair term at fixed density.
#pragma once
#include <type_traits>

#include "sg_common.h"

namespace sg {

namespace ad {
// Value and one directional derivative; arithmetic by the chain rule.  The
// functions below are found by argument-dependent lookup, so code written
// for a scalar T calls sqrt / exp of either type unchanged.
struct Dual {
  double v, d;
  __host__ __device__ Dual(double x = 0.0) : v(x), d(0.0) {}
  __host__ __device__ Dual(double x, double dx) : v(x), d(dx) {}
};
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
__device__ __forceinline__ Dual operator+(Dual a, double b) { return Dual(a.v + b, a.d); }
__device__ __forceinline__ Dual operator+(double a, Dual b) { return Dual(a + b.v, b.d); }
__device__ __forceinline__ Dual operator-(Dual a, double b) { return Dual(a.v - b, a.d); }
__device__ __forceinline__ Dual operator-(double a, Dual b) { return Dual(a - b.v, -b.d); }
__device__ __forceinline__ Dual operator*(Dual a, double b) { return Dual(a.v * b, a.d * b); }
__device__ __forceinline__ Dual operator*(double a, Dual b) { return Dual(a * b.v, a * b.d); }
__device__ __forceinline__ Dual operator/(Dual a, double b) { return Dual(a.v / b, a.d / b); }
__device__ __forceinline__ Dual operator/(double a, Dual b) {
  const double q = a / b.v;
  return Dual(q, -q * b.d / b.v);
}
__device__ __forceinline__ Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__device__ __forceinline__ Dual& operator+=(Dual& a, Dual b) { return a = a + b; }
__device__ __forceinline__ Dual& operator-=(Dual& a, Dual b) { return a = a - b; }
__device__ __forceinline__ Dual& operator*=(Dual& a, Dual b) { return a = a * b; }
__device__ __forceinline__ Dual& operator/=(Dual& a, Dual b) { return a = a / b; }
__device__ __forceinline__ Dual sqrt(Dual a) {
  const double r = ::sqrt(a.v);
  return Dual(r, a.d / (2.0 * r));
}
__device__ __forceinline__ Dual exp(Dual a) {
  const double e = ::exp(a.v);
  return Dual(e, a.d * e);
}
}  // namespace ad

template <typename T>
__host__ __device__ constexpr bool is_plain() {
  return std::is_same<T, double>::value;
}
// |x| of the value, for the frame's axial test
__device__ __forceinline__ double vabs(double x) { return fabs(x); }
__device__ __forceinline__ double vabs(ad::Dual x) { return fabs(x.v); }

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
template <typename T>
__device__ __forceinline__ T rcp(T x) { return 1.0 / x; }

template <typename T>
__device__ T mp_qq(T R, int a, int b) { return 1.0 / sqrt(R * R + addt(a, 0, b, 0)); }
template <typename T>
__device__ T mp_qUz(T R, int a, int b) {
  const double t = addt(a, 0, b, 1);
  T f = 0.0, x;
  x = R + c_par.dd0[b]; f += rcp(2.0 * sqrt(x * x + t));
  x = R - c_par.dd0[b]; f -= rcp(2.0 * sqrt(x * x + t));
  return f;
}
template <typename T>
__device__ T mp_qQpipi(T R, int a, int b) {
  const double t = addt(a, 0, b, 2);
  T f = 0.0, x;
  x = (2.0 * c_par.dd1[b]) * (2.0 * c_par.dd1[b]);
  f += rcp(2.0 * sqrt(R * R + x + t));
  f -= rcp(2.0 * sqrt(R * R + t));
  return f;
}
template <typename T>
__device__ T mp_qQzz(T R, int a, int b) {
  const double t = addt(a, 0, b, 2);
  T f = 0.0, x;
  x = R + 2.0 * c_par.dd1[b]; f += rcp(4.0 * sqrt(x * x + t));
  f -= rcp(2.0 * sqrt(R * R + t));
  x = R - 2.0 * c_par.dd1[b]; f += rcp(4.0 * sqrt(x * x + t));
  return f;
}
template <typename T>
__device__ T mp_UpiUpi(T R, int a, int b) {
  const double t = addt(a, 1, b, 1);
  T f = 0.0, x;
  x = c_par.dd0[a] - c_par.dd0[b]; f += rcp(2.0 * sqrt(R * R + x * x + t));
  x = c_par.dd0[a] + c_par.dd0[b]; f -= rcp(2.0 * sqrt(R * R + x * x + t));
  return f;
}
template <typename T>
__device__ T mp_UzUz(T R, int a, int b) {
  const double t = addt(a, 1, b, 1);
  T f = 0.0, x;
  const double da = c_par.dd0[a], db = c_par.dd0[b];
  x = da + db + R;    f += -1.0 * rcp(4.0 * sqrt(x * x + t));
  x = da + -db + R;   f += 1.0 * rcp(4.0 * sqrt(x * x + t));
  x = -da + db + R;   f += 1.0 * rcp(4.0 * sqrt(x * x + t));
  x = -da + -db + R;  f += -1.0 * rcp(4.0 * sqrt(x * x + t));
  return f;
}
template <typename T>
__device__ T mp_UpiQpiz(T R, int a, int b) {
  const double t = addt(a, 1, b, 2);
  T f = 0.0, x, y;
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
template <typename T>
__device__ T mp_UzQpipi(T R, int a, int b) {
  const double t = addt(a, 1, b, 2);
  T f = 0.0, x, y;
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
template <typename T>
__device__ T mp_UzQz(T R, int a, int b) {
  const double t = addt(a, 1, b, 2);
  T f = 0.0, x;
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
template <typename T>
__device__ T mp_QpipiQpipi(T R, int a, int b) {
  const double t = addt(a, 2, b, 2);
  T f = 0.0, x, y;
  const double da = c_par.dd1[a], db = c_par.dd1[b];
  y = da + db;  x = 4.0 * (y * y); f += rcp(8.0 * sqrt(R * R + x + t));
  y = da + -db; x = 4.0 * (y * y); f += rcp(8.0 * sqrt(R * R + x + t));
  y = 2.0 * da; x = y * y; f -= rcp(4.0 * sqrt(R * R + x + t));
  y = 2.0 * db; x = y * y; f -= rcp(4.0 * sqrt(R * R + x + t));
  f += rcp(4.0 * sqrt(R * R + t));
  return f;
}
template <typename T>
__device__ T mp_QxxQyy(T R, int a, int b) {
  const double t = addt(a, 2, b, 2);
  T f = 0.0, x;
  const double ya = 2.0 * c_par.dd1[a], yb = 2.0 * c_par.dd1[b];
  x = ya * ya + yb * yb; f += rcp(4.0 * sqrt(R * R + x + t));
  x = ya * ya;           f -= rcp(4.0 * sqrt(R * R + x + t));
  x = yb * yb;           f -= rcp(4.0 * sqrt(R * R + x + t));
  f += rcp(4.0 * sqrt(R * R + t));
  return f;
}
template <typename T>
__device__ T mp_QpipiQzz(T R, int a, int b) {
  const double t = addt(a, 2, b, 2);
  T f = 0.0, aux, auxx, x;
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
template <typename T>
__device__ T mp_QzzQzz(T R, int a, int b) {
  const double t = addt(a, 2, b, 2);
  T f = 0.0, x;
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
template <typename T>
__device__ T mp_QpizQpiz(T R, int a, int b) {
  const double t = addt(a, 2, b, 2);
  T f = 0.0, x, y;
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
template <typename T>
__device__ __forceinline__ T dist3(T ax, T ay, T az, T bx, T by, T bz) {
  T r = 0.0;
  r += (ax - bx) * (ax - bx);
  r += (ay - by) * (ay - by);
  r += (az - bz) * (az - bz);
  return sqrt(r);
}

template <typename T>
struct FrameT {
  T m[3][3];
};
using Frame = FrameT<double>;

// For the forward-mode derivatives (T a dual number) the fixed frame of an
// axial pair is replaced by the frame Gram-Schmidt builds from the y axis,
// which is the same frame at the axis (up to the sign of row 0, which every
// rotated integral carries an even number of times) but follows the pair off
// it, so its derivatives are those of an orthonormal frame.
template <typename T>
__device__ __forceinline__ FrameT<T> local_frame_xyz(T ax, T ay, T az, T cx, T cy, T cz, T R) {
  FrameT<T> f;
  f.m[2][0] = (cx - ax) / R;
  f.m[2][1] = (cy - ay) / R;
  f.m[2][2] = (cz - az) / R;
  if (vabs(f.m[2][0] - 0.0) < 1e-10 && vabs(f.m[2][1] - 0.0) < 1e-10) {
    if constexpr (is_plain<T>()) {
      f.m[1][0] = 0.0; f.m[1][1] = 1.0; f.m[1][2] = 0.0;
      f.m[0][0] = 1.0; f.m[0][1] = 0.0; f.m[0][2] = 0.0;
    } else {
      T y0 = 0.0 - f.m[2][1] * f.m[2][0], y1 = 1.0 - f.m[2][1] * f.m[2][1],
        y2 = 0.0 - f.m[2][1] * f.m[2][2];
      const T s = 1.0 / sqrt(y0 * y0 + y1 * y1 + y2 * y2);
      f.m[1][0] = y0 * s; f.m[1][1] = y1 * s; f.m[1][2] = y2 * s;
      f.m[0][0] = f.m[1][1] * f.m[2][2] - f.m[1][2] * f.m[2][1];
      f.m[0][1] = f.m[1][2] * f.m[2][0] - f.m[1][0] * f.m[2][2];
      f.m[0][2] = f.m[1][0] * f.m[2][1] - f.m[1][1] * f.m[2][0];
    }
  } else {
    T s = 1.0 / sqrt(f.m[2][0] * f.m[2][0] + f.m[2][1] * f.m[2][1]);
    f.m[1][0] = s * -f.m[2][1];
    f.m[1][1] = s * f.m[2][0];
    f.m[1][2] = 0.0;
    f.m[0][0] = f.m[2][1] * f.m[1][2] - f.m[2][2] * f.m[1][1];
    f.m[0][1] = f.m[2][2] * f.m[1][0] - f.m[2][0] * f.m[1][2];
    f.m[0][2] = f.m[2][0] * f.m[1][1] - f.m[2][1] * f.m[1][0];
  }
  return f;
}

__device__ __forceinline__ Frame local_frame(const AtomDesc& A, const AtomDesc& C, double R) {
  return local_frame_xyz<double>(A.x, A.y, A.z, C.x, C.y, C.z, R);
}

// The local-frame combinations an X-X pair needs, each formed once with the
// same additions as the reference's IntegralType* functions.
template <typename T>
struct LocalXXT {
  T qq, qUz, Uzq;
  T sspp_pi, sspp_z, ppss_pi, ppss_z;   // SS|PP and PP|SS (inverted) sums
  T sppi, spzz;                          // SP|SP: UpiUpi, UzUz
  T sppp1, sppp2, sppp3;                 // SP|PP
  T ppsp1, ppsp2, ppsp3;                 // PP|SP (inverted)
  T zzzz, pppp, zzpp, ppzz, pzpz, xxyy, xyxy;
};
using LocalXX = LocalXXT<double>;

template <typename T>
__device__ LocalXXT<T> local_xx(T R, int a, int c) {
  LocalXXT<T> v;
  v.qq = mp_qq(R, a, c);
  v.qUz = mp_qUz(R, a, c);
  v.Uzq = mp_qUz(R, c, a) * -1.0;
  const T qQpp = mp_qQpipi(R, a, c), Qppq = mp_qQpipi(R, c, a);
  const T qQzz = mp_qQzz(R, a, c), Qzzq = mp_qQzz(R, c, a);
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
// The same entry for slot indices known only at run time; m(i, j) gives the
// frame (in registers when K and L are constants, as in xx_entry; from LDS
// for the gradient's lanes, one entry per lane).
template <typename T, typename M>
__device__ __forceinline__ T xx_entry_rt(const LocalXXT<T>& v, const M& m, int K, int L) {
  const int s0 = mop_a(K), s1 = mop_b(K), s2 = mop_a(L), s3 = mop_b(L);
  const int pa0 = (s0 > 0) + (s1 > 0), pb0 = (s2 > 0) + (s3 > 0);
  const bool inv = pa0 > pb0;  // CorrectOrderIntegral (:785-796)
  const int pa = inv ? pb0 : pa0, pb = inv ? pa0 : pb0;
  const int t0 = inv ? s2 : s0, t1 = inv ? s3 : s1, t2 = inv ? s0 : s2, t3 = inv ? s1 : s3;
  T r;
  if (pa == 0 && pb == 0) {
    r = v.qq;
  } else if (pa == 0 && pb == 1) {
    r = (inv ? v.Uzq : v.qUz) * m(2, t3 - 1);
  } else if (pa == 0 && pb == 2) {
    T vp = inv ? v.ppss_pi : v.sspp_pi, vz = inv ? v.ppss_z : v.sspp_z;
    vp = vp * m(0, t2 - 1) * m(0, t3 - 1) + vp * m(1, t2 - 1) * m(1, t3 - 1);
    vz = vz * m(2, t2 - 1) * m(2, t3 - 1);
    r = vp + vz;
  } else if (pa == 1 && pb == 1) {
    T v1 = v.sppi * m(0, t1 - 1) * m(0, t3 - 1) + v.sppi * m(1, t1 - 1) * m(1, t3 - 1);
    T v2 = v.spzz * m(2, t1 - 1) * m(2, t3 - 1);
    r = v1 + v2;
  } else if (pa == 1 && pb == 2) {
    const T w1 = inv ? v.ppsp1 : v.sppp1, w2 = inv ? v.ppsp2 : v.sppp2,
                 w3 = inv ? v.ppsp3 : v.sppp3;
    T v1 = w1 * m(2, t1 - 1) * m(0, t2 - 1) * m(0, t3 - 1) +
                w1 * m(2, t1 - 1) * m(1, t2 - 1) * m(1, t3 - 1);
    T v2 = w2 * m(2, t1 - 1) * m(2, t2 - 1) * m(2, t3 - 1);
    T v3 = w3 * m(0, t1 - 1) * m(2, t2 - 1) * m(0, t3 - 1) +
                w3 * m(1, t1 - 1) * m(2, t2 - 1) * m(1, t3 - 1) +
                w3 * m(0, t1 - 1) * m(2, t3 - 1) * m(0, t2 - 1) +
                w3 * m(1, t1 - 1) * m(2, t3 - 1) * m(1, t2 - 1);
    r = v1 + v2 + v3;
  } else {
    const int a = t0 - 1, b = t1 - 1, c = t2 - 1, d = t3 - 1;
    T zz = v.zzzz * m(2, a) * m(2, b) * m(2, c) * m(2, d);
    T pp = v.pppp * m(0, a) * m(0, b) * m(0, c) * m(0, d) +
                v.pppp * m(1, a) * m(1, b) * m(1, c) * m(1, d);
    T zp = v.zzpp * m(2, a) * m(2, b) * m(0, c) * m(0, d) +
                v.zzpp * m(2, a) * m(2, b) * m(1, c) * m(1, d);
    T pz = v.ppzz * m(0, a) * m(0, b) * m(2, c) * m(2, d) +
                v.ppzz * m(1, a) * m(1, b) * m(2, c) * m(2, d);
    T q = v.pzpz;
    T pzpz = q * m(0, a) * m(2, b) * m(0, c) * m(2, d) + q * m(0, a) * m(2, b) * m(0, d) * m(2, c) +
                  q * m(0, b) * m(2, a) * m(0, c) * m(2, d) + q * m(0, b) * m(2, a) * m(0, d) * m(2, c) +
                  q * m(1, a) * m(2, b) * m(1, c) * m(2, d) + q * m(1, a) * m(2, b) * m(1, d) * m(2, c) +
                  q * m(1, b) * m(2, a) * m(1, c) * m(2, d) + q * m(1, b) * m(2, a) * m(1, d) * m(2, c);
    T xxyy = v.xxyy * m(0, a) * m(0, b) * m(1, c) * m(1, d) +
                  v.xxyy * m(1, a) * m(1, b) * m(0, c) * m(0, d);
    T w = v.xyxy;
    T xyxy = w * m(0, a) * m(1, b) * m(0, c) * m(1, d) + w * m(0, a) * m(1, b) * m(0, d) * m(1, c) +
                  w * m(0, b) * m(1, a) * m(0, c) * m(1, d) + w * m(0, b) * m(1, a) * m(0, d) * m(1, c);
    r = zz + pp + zp + pz + pzpz + xxyy + xyxy;
  }
  return r * kHartreeToEv;
}

template <int K, int L, typename T>
__device__ __forceinline__ T xx_entry(const LocalXXT<T>& v, const FrameT<T>& F) {
  const auto m = [&F](int i, int j) -> T { return F.m[i][j]; };
  return xx_entry_rt<T>(v, m, K, L);
}


// ---------------------------------------------------------------------------
// Off-diagonal Hcore blocks beta * S / 2 for one atom pair per lane, launched
// per pair type so the lanes of a wave do the same work.  The STO-6G overlap
// is Overlap::ComputeOverlap (overlap.cpp:68-134): a sum over the 6 x 6
// primitive pairs of exp(-mu R^2) times one Gaussian moment per axis.  The AO
// pairs of one shell pair (s-s, s-p, p-s, p-p) share their primitive
// exponents, so exp(-mu R^2), the Gaussian centre and the axis moments are
// formed once per primitive pair and shell pair.
template <int NA, int NB, typename T>
__device__ __forceinline__ void hcore_shell_xyz(int za, const T (&ac)[3], int sa, int zb,
                                                const T (&bc)[3], int sb, T ab, T* tot) {
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      const double al = sa == 0 ? c_sto.expS[za][i] : c_sto.expP[za][i];
      const double be = sb == 0 ? c_sto.expS[zb][j] : c_sto.expP[zb][j];
      const double p = al + be;
      const double mu = al * be / p;
      T o = exp(-mu * ab);
      o *= sa == 0 ? c_sto.cS[za][i] : c_sto.cP[za][i];
      o *= sb == 0 ? c_sto.cS[zb][j] : c_sto.cP[zb][j];
      // Per axis, the reference's 12-point Gauss-Hermite sums of 1, (x+pa),
      // (x+pb) and (x+pa)(x+pb) against exp(-p x^2) (overlap.cpp:114-124,
      // 177-192) are exact for these degrees, so their closed forms are used:
      // sqrt(pi/p) {1, pa, pb, pa pb + 1/(2p)}
      const double i00 = sqrt(M_PI / p);
      const double h2 = 0.5 / p;
      T f10[3], f01[3], f11[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        T P = al * ac[k];
        P += be * bc[k];
        P /= p;
        const T pa = P - ac[k], pb = P - bc[k];
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
          T t = o;
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            const bool pa = la == k + 1, pb = lb == k + 1;
            t *= pa ? (pb ? f11[k] : f10[k]) : (pb ? f01[k] : i00);
          }
          tot[u * NB + v] += t;
        }
    }
}

template <int NA, int NB>
__device__ __forceinline__ void hcore_shell(const AtomDesc& A, int sa, const AtomDesc& B, int sb,
                                            double ab, double* tot) {
  const double ac[3] = {A.x, A.y, A.z}, bc[3] = {B.x, B.y, B.z};
  hcore_shell_xyz<NA, NB, double>(A.zn, ac, sa, B.zn, bc, sb, ab, tot);
}

}  // namespace sg
