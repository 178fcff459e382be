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
 *   WI                     two's-complement integer in WL u32 limbs (512
 *                          bits for FP32 balls, 768 for FP64; exact stage
 *                          on inputs scaled by 2^S).
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
/* WI: WL-limb two's-complement integer                                */
/* ------------------------------------------------------------------ */
struct WI {
  uint32_t l[WL];
};
PAC_DEV WI wi_from_i64(int64_t v) {
  WI r;
  r.l[0] = (uint32_t)(uint64_t)v;
  r.l[1] = (uint32_t)((uint64_t)v >> 32);
  uint32_t ext = v < 0 ? 0xffffffffu : 0u;
#pragma unroll
  for (int i = 2; i < WL; ++i) r.l[i] = ext;
  return r;
}
PAC_DEV WI operator+(const WI& a, const WI& b) {
  WI r;
  uint64_t c = 0;
#pragma unroll
  for (int i = 0; i < WL; ++i) {
    c += (uint64_t)a.l[i] + b.l[i];
    r.l[i] = (uint32_t)c;
    c >>= 32;
  }
  return r;
}
PAC_DEV WI operator-(const WI& a, const WI& b) {
  WI r;
  uint64_t borrow = 0;
#pragma unroll
  for (int i = 0; i < WL; ++i) {
    uint64_t t = (uint64_t)a.l[i] - (uint64_t)b.l[i] - borrow;
    r.l[i] = (uint32_t)t;
    borrow = (t >> 63) & 1u;
  }
  return r;
}
PAC_DEV WI wi_neg(const WI& a) {
  WI z;
#pragma unroll
  for (int i = 0; i < WL; ++i) z.l[i] = 0;
  return z - a;
}
PAC_DEV int wi_sign(const WI& a) {
  if (a.l[WL - 1] >> 31) return -1;
  uint32_t o = 0;
#pragma unroll
  for (int i = 0; i < WL; ++i) o |= a.l[i];
  return o ? 1 : 0;
}
PAC_DEV WI operator*(const WI& a, const WI& b) {
  bool na = (a.l[WL - 1] >> 31) != 0, nb = (b.l[WL - 1] >> 31) != 0;
  WI ma = na ? wi_neg(a) : a;
  WI mb = nb ? wi_neg(b) : b;
  WI r;
#pragma unroll
  for (int i = 0; i < WL; ++i) r.l[i] = 0;
  int la = WL, lb = WL;
  while (la > 0 && ma.l[la - 1] == 0) --la;
  while (lb > 0 && mb.l[lb - 1] == 0) --lb;
  for (int i = 0; i < la; ++i) {
    uint64_t carry = 0;
    uint64_t ai = ma.l[i];
    int j = 0;
    for (; j < lb && i + j < WL; ++j) {
      uint64_t t = ai * mb.l[j] + r.l[i + j] + carry;
      r.l[i + j] = (uint32_t)t;
      carry = t >> 32;
    }
    if (i + j < WL) r.l[i + j] = (uint32_t)carry;
  }
  return (na != nb) ? wi_neg(r) : r;
}
PAC_DEV int fe_sign(const WI& a) { return wi_sign(a); }

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
template <>
struct K<WI> {
  static PAC_DEV WI c(int v) { return wi_from_i64(v); }
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
/* v * 2^S as an integer; S is at least minus the exponent of v's lowest
 * set bit, so the product is exact.  FP32: |v| 2^S < 2^63 (checked on the
 * host), one int64 conversion. */
PAC_DEV WI wi_scaled(float v, int S) {
  return wi_from_i64((int64_t)__builtin_ldexp((double)v, S));
}
/* FP64: the 53-bit significand shifted into place limb by limb. */
PAC_DEV WI wi_scaled(double v, int S) {
  WI r;
  uint64_t bits = (uint64_t)__double_as_longlong(v);
  int E = (int)((bits >> 52) & 0x7ff);
  uint64_t m = bits & ((1ull << 52) - 1);
  int e;
  if (E == 0) {
    e = -1074;
  } else {
    m |= 1ull << 52;
    e = E - 1075;
  }
  int sh = e + S;
  if (sh < 0) { /* the shifted-out bits are zero by the choice of S */
    m >>= -sh;
    sh = 0;
  }
  const int q = sh >> 5, b = sh & 31;
  const uint64_t lo = m << b;
  const uint32_t hi = b ? (uint32_t)(m >> (64 - b)) : 0u;
#pragma unroll
  for (int i = 0; i < WL; ++i)
    r.l[i] = i == q ? (uint32_t)lo
                    : (i == q + 1 ? (uint32_t)(lo >> 32) : (i == q + 2 ? hi : 0u));
  if (bits >> 63) {
    WI z;
#pragma unroll
    for (int i = 0; i < WL; ++i) z.l[i] = 0;
    return z - r;
  }
  return r;
}
struct LdWI {
  int S;
  PAC_DEV V3<WI> p(const Ball& b) const {
    return V3<WI>{wi_scaled(b.x, S), wi_scaled(b.y, S), wi_scaled(b.z, S)};
  }
  PAC_DEV WI w(const Ball& b) const {
    WI r = wi_scaled(b.w, S);
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
