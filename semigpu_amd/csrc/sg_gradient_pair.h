// sg_gradient_pair.h -- one atom pair's geometry-dependent MNDO energy at a
// fixed density, as a function of the scalar type (the analytic gradient runs
// it on dual numbers; see sg_gradient.hip for the decomposition).
#pragma once
#include "sg_common.h"
#include "sg_pair_math.h"

namespace sg {
namespace grad {

using ad::Dual;

// Dense density entry (row-major N x N, symmetric).
struct Dens {
  const double* p;
  int n;
  __device__ double operator()(int i, int j) const { return p[(int64_t)i * n + j]; }
};

// Coulomb weight of slot k (AO pair (a, b), a <= b) of the atom whose first
// AO is o.
__device__ __forceinline__ double weight(const Dens& P, int o, int k) {
  const int a = mop_a(k), b = mop_b(k);
  return a == b ? P(o + a, o + a) : 2.0 * P(o + a, o + b);
}

// Exchange coefficient of slots k (on A, first AO oa) and l (on C, first AO
// oc): sum over the orderings (mu, lambda) of k and (nu, sigma) of l of
// P(mu, nu) P(lambda, sigma).
__device__ __forceinline__ double exch(const Dens& P, int oa, int k, int oc, int l) {
  const int a = mop_a(k), b = mop_b(k), c = mop_a(l), d = mop_b(l);
  double x = P(oa + a, oc + c) * P(oa + b, oc + d);
  if (a != b && c != d) {
    x += P(oa + a, oc + d) * P(oa + b, oc + c);
    x *= 2.0;
  } else if (a != b || c != d) {
    x *= 2.0;
  }
  return x;
}

__device__ __forceinline__ double coef(const Dens& P, const AtomDesc& A, int k, const AtomDesc& C,
                                       int l) {
  const double wa = weight(P, A.ao0, k), wc = weight(P, C.ao0, l);
  double d = wa * wc - 0.5 * exch(P, A.ao0, k, C.ao0, l);
  if (l == 0) d -= (double)core_charge(C.zn) * wa;
  if (k == 0) d -= (double)core_charge(A.zn) * wc;
  return d;
}

template <int E>
struct XXSum {
  template <typename T>
  __device__ static __forceinline__ T run(const LocalXXT<T>& v, const FrameT<T>& F, const Dens& P,
                                          const AtomDesc& A, const AtomDesc& C) {
    const T s = XXSum<E - 1>::run(v, F, P, A, C);
    constexpr int k = (E - 1) / 10, l = (E - 1) % 10;
    return s + coef(P, A, k, C, l) * xx_entry<k, l>(v, F);
  }
};
template <>
struct XXSum<0> {
  template <typename T>
  __device__ static __forceinline__ T run(const LocalXXT<T>&, const FrameT<T>&, const Dens&,
                                          const AtomDesc&, const AtomDesc&) {
    return T(0.0);
  }
};

// The derivative parts of a heavy pair's 100 rotated entries (eV), in block
// order, into out[0..99].
template <int E>
struct XXDeriv {
  __device__ static __forceinline__ void run(const LocalXXT<Dual>& v, const FrameT<Dual>& F,
                                             double* out) {
    XXDeriv<E - 1>::run(v, F, out);
    out[E - 1] = xx_entry<(E - 1) / 10, (E - 1) % 10>(v, F).d;
  }
};
template <>
struct XXDeriv<0> {
  __device__ static __forceinline__ void run(const LocalXXT<Dual>&, const FrameT<Dual>&, double*) {}
};

// Resonance term: sum over AO pairs of P (beta_mu + beta_nu) S, the overlaps
// of the shell pairs in the Hcore kernel's form (S(AO of A, AO of C)); one
// shell pair (NA x NB AOs from A's first AO oa, C's oc) at a time, with
// compile-time shapes so that the overlaps stay in registers.
template <int NA, int NB, typename T>
__device__ __forceinline__ T resonance_shells(const Dens& P, const AtomDesc& A, int oa,
                                              const T (&ac)[3], const AtomDesc& C, int oc,
                                              const T (&cc)[3], T ab) {
  T t[NA * NB];
#pragma unroll
  for (int q = 0; q < NA * NB; ++q) t[q] = 0.0;
  hcore_shell_xyz<NA, NB, T>(A.zn, ac, NA == 1 ? 0 : 1, C.zn, cc, NB == 1 ? 0 : 1, ab, t);
  const double bu = NA == 1 ? c_par.betaS[A.zn] : c_par.betaP[A.zn];
  const double bv = NB == 1 ? c_par.betaS[C.zn] : c_par.betaP[C.zn];
  T g = 0.0;
#pragma unroll
  for (int u = 0; u < NA; ++u)
#pragma unroll
    for (int w = 0; w < NB; ++w) g += P(A.ao0 + oa + u, C.ao0 + oc + w) * (bu + bv) * t[u * NB + w];
  return g;
}

template <typename T>
__device__ T resonance(const Dens& P, const AtomDesc& A, const T (&ac)[3], const AtomDesc& C,
                       const T (&cc)[3], T ab) {
  T g = resonance_shells<1, 1, T>(P, A, 0, ac, C, 0, cc, ab);
  if (C.zn > 1) g += resonance_shells<1, 3, T>(P, A, 0, ac, C, 1, cc, ab);
  if (A.zn > 1) g += resonance_shells<3, 1, T>(P, A, 1, ac, C, 0, cc, ab);
  if (A.zn > 1 && C.zn > 1) g += resonance_shells<3, 3, T>(P, A, 1, ac, C, 1, cc, ab);
  return g;
}

// Core-core term Z_A Z_C (ss|ss) f(R) of CoreCoreRepulsion (corecorerepulsion.cpp:12-65).
template <typename T>
__device__ T core_term(const AtomDesc& A, const AtomDesc& C, T R) {
  T t = mp_qq(R, A.zn, C.zn) * kHartreeToEv * (double)(core_charge(A.zn) * core_charge(C.zn));
  const T r = R * kBohrToAng;
  T f;
  if (C.zn == 1 && (A.zn == 7 || A.zn == 8))
    f = 1.0 + r * exp(-c_par.alpha[A.zn] * r) + exp(-c_par.alpha[C.zn] * r);
  else if (A.zn == 1 && (C.zn == 7 || C.zn == 8))
    f = 1.0 + r * exp(-c_par.alpha[C.zn] * r) + exp(-c_par.alpha[A.zn] * r);
  else
    f = 1.0 + exp(-c_par.alpha[A.zn] * r) + exp(-c_par.alpha[C.zn] * r);
  return t * f;
}

// The geometry-dependent energy E_AC(P) of the pair (A = row atom, C = column
// atom) at atom positions ac, cc (bohr), of the given kind: 0 both heavy, 1
// one hydrogen, 2 both H.  With T a dual number seeded on one coordinate of C
// its derivative part is dE_AC / dR_C along that axis.
template <int KIND, typename T>
__device__ T pair_energy(const AtomDesc& A, const T (&ac)[3], const AtomDesc& C, const T (&cc)[3],
                         const Dens& P) {
  const T R = dist3(ac[0], ac[1], ac[2], cc[0], cc[1], cc[2]);
  T e = 0.0;
  if constexpr (KIND == 0) {
    const FrameT<T> F = local_frame_xyz(ac[0], ac[1], ac[2], cc[0], cc[1], cc[2], R);
    const LocalXXT<T> v = local_xx(R, A.zn, C.zn);
    e += XXSum<100>::run(v, F, P, A, C);
  } else if constexpr (KIND == 1) {
    // the frame runs from the hydrogen to the heavy atom, and entry k is
    // (ss of H | slot k of the heavy atom) (k_pair_hx)
    const bool a_is_h = A.zn == 1;
    const FrameT<T> F = a_is_h ? local_frame_xyz(ac[0], ac[1], ac[2], cc[0], cc[1], cc[2], R)
                               : local_frame_xyz(cc[0], cc[1], cc[2], ac[0], ac[1], ac[2], R);
    const int h = a_is_h ? A.zn : C.zn, x = a_is_h ? C.zn : A.zn;
    const T qq = mp_qq(R, h, x);
    const T qUz = mp_qUz(R, h, x);
    const T vpi = qq + mp_qQpipi(R, h, x);
    const T vz = qq + mp_qQzz(R, h, x);
    const auto& m = F.m;
    for (int k = 0; k < 10; ++k) {
      const int t2 = mop_a(k), t3 = mop_b(k);
      T r;
      if (k == 0) {
        r = qq;
      } else if (t2 == 0) {
        r = qUz * m[2][t3 - 1];
      } else {
        T a = vpi * m[0][t2 - 1] * m[0][t3 - 1] + vpi * m[1][t2 - 1] * m[1][t3 - 1];
        T b = vz * m[2][t2 - 1] * m[2][t3 - 1];
        r = a + b;
      }
      const double d = a_is_h ? coef(P, A, 0, C, k) : coef(P, A, k, C, 0);
      e += d * (r * kHartreeToEv);
    }
  } else {
    e += coef(P, A, 0, C, 0) * (mp_qq(R, A.zn, C.zn) * kHartreeToEv);
  }
  T ab = 0.0;
  {
    const T d0 = ac[0] - cc[0], d1 = ac[1] - cc[1], d2 = ac[2] - cc[2];
    ab += d0 * d0;
    ab += d1 * d1;
    ab += d2 * d2;
  }
  e += resonance(P, A, ac, C, cc, ab);
  e += core_term(A, C, R);
  return e;
}

// dE_AC / dR_C (eV / bohr) along axis dir.
template <int KIND>
__device__ double pair_grad(const AtomDesc& A, const AtomDesc& C, const Dens& P, int dir) {
  const Dual ac[3] = {Dual(A.x), Dual(A.y), Dual(A.z)};
  const Dual cc[3] = {Dual(C.x, dir == 0 ? 1.0 : 0.0), Dual(C.y, dir == 1 ? 1.0 : 0.0),
                      Dual(C.z, dir == 2 ? 1.0 : 0.0)};
  return pair_energy<KIND, Dual>(A, ac, C, cc, P).d;
}

// Pair index of (row a, column j < a) inside its molecule.
__device__ __forceinline__ int64_t pair_slot(int a, int j) { return (int64_t)a * (a - 1) / 2 + j; }


}  // namespace grad
}  // namespace sg
