/*
 * pac_arith.h -- device arithmetic for the precision ladder.
 *
 *   FE<float>, FE<double>  value + certified absolute error bound.  Every
 *                          + - * adds the propagated input errors and the
 *                          op's own rounding (u|result|) plus an underflow
 *                          guard; a sign is certified when |v| > e*slack.
 *                          Requires -ffp-contract=off (each op rounds once
 *                          as modelled; explicit fma only where intended).
 *   FEDD                   double-double value (Joldes-Muller-Popescu 2017
 *                          AccurateDWPlusDW / DWTimesDW3, relative error
 *                          <= 4u^2) with a double error bound charged at
 *                          2^-100 per op.
 *   WI                     storage of a WL-limb integer (alpha scaled by
 *                          2^(2S)); the exact stage itself computes with
 *                          the wavefront-cooperative integers of
 *                          pac_wide.h on inputs scaled by 2^S.
 *
 * The predicate polynomials are written once over any of these rings; they
 * restate the reference filters (AlphaComplexLargeData.cu:167-190, :298-348,
 * :531-591) and emptiness tests (:597-735, :807-1015) multiplied through by
 * their positive denominators so that exact evaluation needs no division.
 */
#ifndef PAC_ARITH_H
#define PAC_ARITH_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pac_variant.h"

#define PAC_DEV __device__ __forceinline__

namespace pac {
namespace PAC_V {

/* ------------------------------------------------------------------ */
template <class T>
struct V3 {
  T x, y, z;
};
template <class T>
PAC_DEV V3<T> vsub(const V3<T>& a, const V3<T>& b) {
  return V3<T>{a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class T>
PAC_DEV V3<T> vadd(const V3<T>& a, const V3<T>& b) {
  return V3<T>{a.x + b.x, a.y + b.y, a.z + b.z};
}
template <class T>
PAC_DEV T vdot(const V3<T>& a, const V3<T>& b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class T>
PAC_DEV V3<T> vcross(const V3<T>& a, const V3<T>& b) {
  return V3<T>{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z,
               a.x * b.y - a.y * b.x};
}
template <class T>
PAC_DEV V3<T> vscale(const T& s, const V3<T>& a) {
  return V3<T>{s * a.x, s * a.y, s * a.z};
}

/* ------------------------------------------------------------------ */
/* FE: float with certified absolute error                             */
/* ------------------------------------------------------------------ */
template <class R>
struct FEc;
template <>
struct FEc<float> {
  static PAC_DEV float u() { return 0x1p-24f; }
  static PAC_DEV float tiny() { return 0x1p-100f; }
  static PAC_DEV float slack() { return 1.0f + 0x1p-10f; }
};
template <>
struct FEc<double> {
  static PAC_DEV double u() { return 0x1p-53; }
  static PAC_DEV double tiny() { return 0x1p-1000; }
  static PAC_DEV double slack() { return 1.0 + 0x1p-20; }
};

template <class R>
struct FE {
  R v, e;
};
template <class R>
PAC_DEV FE<R> operator+(const FE<R>& a, const FE<R>& b) {
  R s = a.v + b.v;
  return FE<R>{s, a.e + b.e + FEc<R>::u() * fabs(s) + FEc<R>::tiny()};
}
template <class R>
PAC_DEV FE<R> operator-(const FE<R>& a, const FE<R>& b) {
  R s = a.v - b.v;
  return FE<R>{s, a.e + b.e + FEc<R>::u() * fabs(s) + FEc<R>::tiny()};
}
template <class R>
PAC_DEV FE<R> operator*(const FE<R>& a, const FE<R>& b) {
  R p = a.v * b.v;
  return FE<R>{p, fabs(a.v) * b.e + fabs(b.v) * a.e + a.e * b.e +
                      FEc<R>::u() * fabs(p) + FEc<R>::tiny()};
}
/* 1: certainly > 0, -1: certainly < 0, 0: undecided */
template <class R>
PAC_DEV int fe_sign(const FE<R>& a) {
  R b = a.e * FEc<R>::slack();
  return a.v > b ? 1 : (a.v < -b ? -1 : 0);
}

/* ------------------------------------------------------------------ */
/* double-double                                                       */
/* ------------------------------------------------------------------ */
struct dd {
  double hi, lo;
};
PAC_DEV dd two_sum(double a, double b) {
  double s = a + b;
  double t = s - a;
  double e = (a - (s - t)) + (b - t);
  return dd{s, e};
}
PAC_DEV dd fast_two_sum(double a, double b) {
  double s = a + b;
  double e = b - (s - a);
  return dd{s, e};
}
PAC_DEV dd two_prod(double a, double b) {
  double p = a * b;
  double e = __builtin_fma(a, b, -p);
  return dd{p, e};
}
PAC_DEV dd dd_add(const dd& x, const dd& y) {
  dd s = two_sum(x.hi, y.hi);
  dd t = two_sum(x.lo, y.lo);
  double c = s.lo + t.hi;
  dd v = fast_two_sum(s.hi, c);
  double w = t.lo + v.lo;
  return fast_two_sum(v.hi, w);
}
PAC_DEV dd dd_neg(const dd& x) { return dd{-x.hi, -x.lo}; }
PAC_DEV dd dd_mul(const dd& x, const dd& y) {
  dd c = two_prod(x.hi, y.hi);
  double tl0 = x.lo * y.lo;
  double tl1 = __builtin_fma(x.hi, y.lo, tl0);
  double cl2 = __builtin_fma(x.lo, y.hi, tl1);
  double cl3 = c.lo + cl2;
  return fast_two_sum(c.hi, cl3);
}

struct FEDD {
  dd v;
  double e;
};
PAC_DEV double ddu() { return 0x1p-100; }
PAC_DEV FEDD operator+(const FEDD& a, const FEDD& b) {
  dd s = dd_add(a.v, b.v);
  return FEDD{s, a.e + b.e + ddu() * fabs(s.hi) + 0x1p-1000};
}
PAC_DEV FEDD operator-(const FEDD& a, const FEDD& b) {
  dd s = dd_add(a.v, dd_neg(b.v));
  return FEDD{s, a.e + b.e + ddu() * fabs(s.hi) + 0x1p-1000};
}
PAC_DEV FEDD operator*(const FEDD& a, const FEDD& b) {
  dd p = dd_mul(a.v, b.v);
  return FEDD{p, fabs(a.v.hi) * b.e + fabs(b.v.hi) * a.e + a.e * b.e +
                     ddu() * fabs(p.hi) + 0x1p-1000};
}
PAC_DEV int fe_sign(const FEDD& a) {
  double b = a.e * (1.0 + 0x1p-20);
  return a.v.hi > b ? 1 : (a.v.hi < -b ? -1 : 0);
}

/* ------------------------------------------------------------------ */
/* WI: storage of a WL-limb two's-complement integer (alpha * 2^(2S),   */
/* built on the host); the exact stage computes with the wavefront-     */
/* cooperative integers of pac_wide.h                                   */
/* ------------------------------------------------------------------ */
struct WI {
  uint32_t l[WL];
};

/* ------------------------------------------------------------------ */
/* Ring constants                                                      */
/* ------------------------------------------------------------------ */
template <class T>
struct K;
template <class R>
struct K<FE<R>> {
  static PAC_DEV FE<R> c(int v) { return FE<R>{(R)v, (R)0}; }
};
template <>
struct K<FEDD> {
  static PAC_DEV FEDD c(int v) { return FEDD{dd{(double)v, 0.0}, 0.0}; }
};

/* ------------------------------------------------------------------ */
/* Ball loaders: position and weight of a ball in each ring.           */
/* Balls are (x, y, z, r) with r = r_raw + solvent as the pipeline     */
/* computed it; inputs are exact in FP64 and in double-double.          */
/* ------------------------------------------------------------------ */
struct ExactScale {
  int S;          /* coordinates scaled by 2^S are integers */
  WI alpha;       /* alpha * 2^(2S) */
};

template <class T>
struct Ld;
template <class R>
struct Ld<FE<R>> {
  static PAC_DEV V3<FE<R>> p(const Ball& b) {
    return V3<FE<R>>{FE<R>{(R)b.x, (R)0}, FE<R>{(R)b.y, (R)0},
                     FE<R>{(R)b.z, (R)0}};
  }
  static PAC_DEV FE<R> w(const Ball& b) {
    FE<R> r{(R)b.w, (R)0};
    return r * r;
  }
};
template <>
struct Ld<FEDD> {
  static PAC_DEV V3<FEDD> p(const Ball& b) {
    return V3<FEDD>{FEDD{dd{(double)b.x, 0.0}, 0.0},
                    FEDD{dd{(double)b.y, 0.0}, 0.0},
                    FEDD{dd{(double)b.z, 0.0}, 0.0}};
  }
  static PAC_DEV FEDD w(const Ball& b) {
    FEDD r{dd{(double)b.w, 0.0}, 0.0};
    return r * r;
  }
};
/* ------------------------------------------------------------------ */
/* Predicate polynomials (ring-generic).                               */
/* ------------------------------------------------------------------ */

/* Edge (LD:167-190): u = p_j - p_i, A = |u|^2 + w_i - w_j;
 * rho <= alpha  <=>  4|u|^2 (w_i + alpha) - A^2 >= 0. */
template <class T>
PAC_DEV T edge_val(const V3<T>& pi, const T& wi, const V3<T>& pj, const T& wj,
                   const T& al) {
  V3<T> u = vsub(pj, pi);
  T uu = vdot(u, u);
  T A = uu + wi - wj;
  return K<T>::c(4) * uu * (wi + al) - A * A;
}

/* Triangle (LD:298-348): x = c - p_i = (A (v x n) - B (u x n)) / (2|n|^2);
 * rho <= alpha <=> 4|n|^2 (w_i+alpha) - (A^2|v|^2 - 2AB u.v + B^2|u|^2) >= 0.
 */
template <class T>
PAC_DEV T tri_val(const V3<T>& pi, const T& wi, const V3<T>& pj, const T& wj,
                  const V3<T>& pk, const T& wk, const T& al, T* nn_out) {
  V3<T> u = vsub(pj, pi), v = vsub(pk, pi);
  V3<T> n = vcross(u, v);
  T nn = vdot(n, n);
  T uu = vdot(u, u), vv = vdot(v, v), uv = vdot(u, v);
  T A = uu + wi - wj, B = vv + wi - wk;
  *nn_out = nn;
  T lhs = A * A * vv - K<T>::c(2) * A * B * uv + B * B * uu;
  return K<T>::c(4) * nn * (wi + al) - lhs;
}

/* Tet (LD:531-591): x = N/(2D), N = A(v x t) + B(t x u) + C(u x v),
 * D = u.(v x t); rho <= alpha <=> 4 D^2 (w_i + alpha) - |N|^2 >= 0. */
template <class T>
PAC_DEV void tet_DN(const V3<T>& pi, const T& wi, const V3<T>& pj, const T& wj,
                    const V3<T>& pk, const T& wk, const V3<T>& pl, const T& wl,
                    T* D, V3<T>* N) {
  V3<T> u = vsub(pj, pi), v = vsub(pk, pi), t = vsub(pl, pi);
  V3<T> vt = vcross(v, t), tu = vcross(t, u), uv = vcross(u, v);
  *D = vdot(u, vt);
  T A = vdot(u, u) + wi - wj, B = vdot(v, v) + wi - wk,
    C = vdot(t, t) + wi - wl;
  *N = vadd(vadd(vscale(A, vt), vscale(B, tu)), vscale(C, uv));
}
template <class T>
PAC_DEV T tet_val(const T& D, const V3<T>& N, const T& wi, const T& al) {
  return K<T>::c(4) * D * D * (wi + al) - vdot(N, N);
}

/* Triangle orthocentre numerator X = A (v x n) - B (u x n) and |n|^2. */
template <class T>
PAC_DEV void tri_X(const V3<T>& pi, const T& wi, const V3<T>& pj, const T& wj,
                   const V3<T>& pk, const T& wk, T* nn, V3<T>* X) {
  V3<T> u = vsub(pj, pi), v = vsub(pk, pi);
  V3<T> n = vcross(u, v);
  *nn = vdot(n, n);
  T A = vdot(u, u) + wi - wj, B = vdot(v, v) + wi - wk;
  *X = vsub(vscale(A, vcross(v, n)), vscale(B, vcross(u, n)));
}

/* power of ball e minus rho, times the positive denominator:
 *   triangle: (L_e |n|^2 - y.X),  edge: (L_e |u|^2 - (y.u) A),
 *   tet:      (L_e D - y.N)  (sign(pi_e - rho) = sign(.) * sign(D)),
 * with y = p_e - p_i, L_e = |y|^2 - w_e + w_i.   Tests <= 0 reject
 * (AlphaComplexLargeData.cu:838-842, :971-976), tet < 0 rejects. */
template <class T>
PAC_DEV T power_gap(const V3<T>& pi, const T& wi, const V3<T>& pe,
                    const T& we, const T& den, const V3<T>& num) {
  V3<T> y = vsub(pe, pi);
  T Le = vdot(y, y) - we + wi;
  return Le * den - vdot(y, num);
}

}  // namespace PAC_V
}  // namespace pac
#endif
