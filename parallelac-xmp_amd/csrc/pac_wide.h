/*
 * pac_wide.h -- wavefront-cooperative wide integers for the exact stage of
 * the precision ladder.
 *
 * A value of 32*TPI bits is held by a group of TPI consecutive lanes of a
 * 64-lane wavefront, one u32 limb per lane (limb i in lane i of the group),
 * in two's complement.  Arithmetic is modulo 2^(32 TPI): every predicate is
 * a polynomial with integer coefficients in the scaled inputs, so its value
 * modulo 2^(32 TPI) is exact whatever the intermediates overflow, and the
 * value itself is read off as a signed number because the host bounds its
 * magnitude below 2^(32 TPI - 1) (ExactScale).
 *
 * Every lane of a group executes the same operations (group-uniform
 * control flow); groups of one wave may diverge.  Cross-lane traffic:
 *   add / sub  each lane adds its limbs; two 64-bit ballots collect the
 *              group's generate (carry out) and propagate (sum all ones)
 *              bits G, P; the carry into every limb is ((G << 1) + P) ^ P
 *              over the group's bits -- a carry-lookahead resolved by one
 *              integer addition;
 *   mul        magnitudes trimmed to their highest non-zero limb (a group
 *              ballot); lane k forms the column sum_{i+j=k} a_i b_j with
 *              v_mad_u64_u32, a_i and b_j read from lanes i and j
 *              (ds_bpermute), splits it into three 32-bit words and two
 *              carry-resolved additions place them at limbs k, k+1, k+2;
 *   sign       the top limb's sign bit and a ballot of non-zero limbs.
 *
 * Replaces the reference's CGBN core that MP_float builds on
 * (cgbn/core/core_add_sub.cu, cgbn/core/core_mul_xmad.cu; carry resolution
 * by a warp ballot of generate / propagate bits, sum = 2g + p,
 * cgbn/core/warp_resolver.cu:68-80): limbs spread over cooperating lanes
 * and a ballot carry resolution are the same idea, here on 64-lane
 * wavefronts with a column-sum product and modular evaluation.
 */
#ifndef PAC_WIDE_H
#define PAC_WIDE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pac_arith.h"

namespace pac {
namespace PAC_V {

/* lanes per wide integer: 512 bits for FP32 balls, 1024 for FP64 balls */
constexpr int kTPI = kBallF64 ? 32 : 16;

template <int TPI>
struct Grp {
  static_assert(TPI == 8 || TPI == 16 || TPI == 32 || TPI == 64, "TPI");
  static __device__ __forceinline__ int lane() { return (int)(threadIdx.x & 63); }
  static __device__ __forceinline__ int idx() { return lane() & (TPI - 1); }
  static __device__ __forceinline__ int base() { return lane() & ~(TPI - 1); }
  static __device__ __forceinline__ uint64_t bits(uint64_t ballot) {
    return TPI == 64 ? ballot
                     : (ballot >> base()) & ((1ull << (TPI & 63)) - 1ull);
  }
  /* value of v in lane i of this group */
  static __device__ __forceinline__ uint32_t from(uint32_t v, int i) {
    return (uint32_t)__shfl((int)v, base() + i, 64);
  }
};

/* one lane's limb of a wide integer of the group */
template <int TPI>
struct CW {
  uint32_t v;
};

/* carry-resolved sum of the lanes' 33-bit partial sums: s = limb sum plus
 * the lane's own carry-in (at most 2^33 - 1) */
template <int TPI>
__device__ __forceinline__ CW<TPI> cw_resolve(uint64_t s) {
  using G = Grp<TPI>;
  const uint32_t lo = (uint32_t)s;
  const uint64_t g = G::bits(__ballot((s >> 32) != 0));
  const uint64_t p = G::bits(__ballot(lo == 0xffffffffu));
  const uint64_t c = (g << 1) + p ^ p;
  return CW<TPI>{lo + (uint32_t)((c >> G::idx()) & 1ull)};
}
template <int TPI>
__device__ __forceinline__ CW<TPI> operator+(const CW<TPI>& a, const CW<TPI>& b) {
  return cw_resolve<TPI>((uint64_t)a.v + b.v);
}
template <int TPI>
__device__ __forceinline__ CW<TPI> operator-(const CW<TPI>& a, const CW<TPI>& b) {
  /* a + ~b + 1: the +1 enters at limb 0 */
  return cw_resolve<TPI>((uint64_t)a.v + (uint32_t)~b.v +
                         (Grp<TPI>::idx() == 0 ? 1u : 0u));
}
template <int TPI>
__device__ __forceinline__ CW<TPI> cw_neg(const CW<TPI>& a) {
  return CW<TPI>{0u} - a;
}
template <int TPI>
__device__ __forceinline__ bool cw_is_neg(const CW<TPI>& a) {
  return (Grp<TPI>::from(a.v, TPI - 1) >> 31) != 0;
}
template <int TPI>
__device__ __forceinline__ int cw_sign(const CW<TPI>& a) {
  if (cw_is_neg(a)) return -1;
  return Grp<TPI>::bits(__ballot(a.v != 0)) ? 1 : 0;
}
/* number of limbs up to the highest non-zero one (0 for zero) */
template <int TPI>
__device__ __forceinline__ int cw_len(const CW<TPI>& a) {
  const uint64_t nz = Grp<TPI>::bits(__ballot(a.v != 0));
  return nz ? 64 - __clzll((unsigned long long)nz) : 0;
}
/* a * b mod 2^(32 TPI) */
template <int TPI>
__device__ __forceinline__ CW<TPI> operator*(const CW<TPI>& a, const CW<TPI>& b) {
  using G = Grp<TPI>;
  const bool na = cw_is_neg(a), nb = cw_is_neg(b);
  const CW<TPI> ma = na ? cw_neg(a) : a, mb = nb ? cw_neg(b) : b;
  const int la = cw_len(ma), lb = cw_len(mb);
  const int k = G::idx();
  uint64_t lo = 0, hi = 0; /* sums of the low and high words of the column */
  for (int i = 0; i < la; ++i) {
    const uint32_t ai = G::from(ma.v, i);
    const int j = k - i;
    const uint32_t bj = G::from(mb.v, j < 0 ? 0 : j);
    if (j >= 0 && j < lb) {
      const uint64_t p = (uint64_t)ai * bj;
      lo += (uint32_t)p;
      hi += p >> 32;
    }
  }
  /* column k = lo + hi 2^32 = w0 + w1 2^32 + w2 2^64 */
  const uint64_t t = (lo >> 32) + (uint32_t)hi;
  const uint32_t w0 = (uint32_t)lo, w1 = (uint32_t)t,
                 w2 = (uint32_t)((t >> 32) + (hi >> 32));
  const uint32_t up1 = G::from(w1, k == 0 ? 0 : k - 1);
  const uint32_t up2 = G::from(w2, k < 2 ? 0 : k - 2);
  CW<TPI> r = CW<TPI>{w0} + CW<TPI>{k >= 1 ? up1 : 0u};
  r = r + CW<TPI>{k >= 2 ? up2 : 0u};
  return na != nb ? cw_neg(r) : r;
}
template <int TPI>
__device__ __forceinline__ CW<TPI> cw_from_i64(int64_t v) {
  const int k = Grp<TPI>::idx();
  const uint32_t ext = v < 0 ? 0xffffffffu : 0u;
  return CW<TPI>{k == 0 ? (uint32_t)(uint64_t)v
                        : (k == 1 ? (uint32_t)((uint64_t)v >> 32) : ext)};
}
template <int TPI>
__device__ __forceinline__ int fe_sign(const CW<TPI>& a) {
  return cw_sign(a);
}
template <int TPI>
struct K<CW<TPI>> {
  static __device__ __forceinline__ CW<TPI> c(int v) { return cw_from_i64<TPI>(v); }
};

/* v * 2^S as a wide integer (S at least minus the exponent of v's lowest
 * set bit, so the product is an integer): the 53-bit significand shifted
 * into place, each lane taking its 32 bits of it. */
template <int TPI>
__device__ __forceinline__ CW<TPI> cw_scaled(double v, int S) {
  const uint64_t bits = (uint64_t)__double_as_longlong(v);
  const int E = (int)((bits >> 52) & 0x7ff);
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
    m >>= (-sh < 64 ? -sh : 63);
    sh = 0;
  }
  /* bits [32 k, 32 k + 32) of m << sh */
  const int lo_bit = 32 * Grp<TPI>::idx() - sh;
  uint32_t limb;
  if (lo_bit >= 64 || lo_bit <= -32)
    limb = 0;
  else if (lo_bit >= 0)
    limb = (uint32_t)(m >> lo_bit);
  else
    limb = (uint32_t)(m << -lo_bit);
  const CW<TPI> r{m ? limb : 0u};
  return (bits >> 63) ? cw_neg(r) : r;
}
template <int TPI>
__device__ __forceinline__ CW<TPI> cw_scaled(float v, int S) {
  return cw_scaled<TPI>((double)v, S);
}
template <int TPI>
struct LdCW {
  int S;
  __device__ __forceinline__ V3<CW<TPI>> p(const Ball& b) const {
    return V3<CW<TPI>>{cw_scaled<TPI>(b.x, S), cw_scaled<TPI>(b.y, S),
                       cw_scaled<TPI>(b.z, S)};
  }
  __device__ __forceinline__ CW<TPI> w(const Ball& b) const {
    const CW<TPI> r = cw_scaled<TPI>(b.w, S);
    return r * r;
  }
};

/* wide integer -> double (relative error < 2^-50), scaled by 2^-S */
template <int TPI>
__device__ __forceinline__ double cw_to_double(const CW<TPI>& a, int S) {
  const bool neg = cw_is_neg(a);
  const CW<TPI> m = neg ? cw_neg(a) : a;
  const int top = cw_len(m) - 1;
  double v = 0;
  for (int q = top; q >= 0 && q >= top - 2; --q)
    v += __builtin_ldexp((double)Grp<TPI>::from(m.v, q), 32 * q);
  v = __builtin_ldexp(v, -S);
  return neg ? -v : v;
}

/* alpha * 2^(2S) from the host's limbs (ExactScale.alpha, WL limbs) */
template <int TPI>
__device__ __forceinline__ CW<TPI> cw_of(const WI& a) {
  const int k = Grp<TPI>::idx();
  const uint32_t ext = (a.l[WL - 1] >> 31) ? 0xffffffffu : 0u;
  uint32_t v = ext;
#pragma unroll
  for (int i = 0; i < WL; ++i)
    if (i == k) v = a.l[i];
  return CW<TPI>{v};
}

/* Queue entries for groups of TPI lanes: group g of the grid takes entry
 * g, then g + groups, ...  Every lane of a group calls body(t) with the
 * same t. */
template <int TPI, class F>
__device__ __forceinline__ void for_queue_grp(const QRef& q, F&& body) {
  const uint32_t n = q.size();
  const uint32_t groups = (gridDim.x * blockDim.x) / TPI;
  const uint32_t g0 = (uint32_t)(gtid() / TPI);
  if (g0 < n) body(g0);
  for (uint32_t t = g0 + groups; t < n; t += groups) body(t); /* as for_queue */
}

}  // namespace PAC_V
}  // namespace pac
#endif
