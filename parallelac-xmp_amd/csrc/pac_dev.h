/*
 * pac_dev.h -- device helpers shared by the kernel translation units:
 * wave-aggregated counters, queue appends, CSR lookups, the emptiness
 * search box, and the semi-static power-gap test used by the hot
 * emptiness kernels.
 */
#ifndef PAC_DEV_H
#define PAC_DEV_H

#include "pac_internal.h"

namespace pac {
namespace PAC_V {

__device__ __forceinline__ unsigned wave_sum(unsigned v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}
/* Counters live in CTR_SLOTS copies; each wave adds into the copy picked
 * by its block index, so that one-item-per-thread kernels do not serialize
 * hundreds of thousands of atomics on a single address.  The host sums
 * the copies. */
__device__ __forceinline__ void flush_ctr(unsigned long long* ctr, int idx,
                                          unsigned v) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0 && v)
    atomicAdd(&ctr[idx * CTR_SLOTS + (blockIdx.x & (CTR_SLOTS - 1))],
              (unsigned long long)v);
}
/* Block-level versions (every thread of the block must call them): one
 * global atomic per block instead of one per wave.  block_flush takes any
 * block of up to kBlock threads in whole waves. */
__device__ __forceinline__ void block_flush(unsigned long long* ctr, int idx,
                                            unsigned v) {
  __shared__ unsigned s_part[kBlock / 64];
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned t = 0;
    const int waves = (int)(blockDim.x >> 6);
    for (int w = 0; w < waves; ++w) t += s_part[w];
    if (t)
      atomicAdd(&ctr[idx * CTR_SLOTS + (blockIdx.x & (CTR_SLOTS - 1))],
                (unsigned long long)t);
  }
  __syncthreads();
}
__device__ __forceinline__ unsigned block_append(unsigned* counter, bool want) {
  __shared__ unsigned s_cnt[kBlock / 64];
  __shared__ unsigned s_base;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint64_t m = __ballot(want);
  uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  if (lane == 0) s_cnt[w] = (unsigned)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned t = 0;
    for (int q = 0; q < kBlock / 64; ++q) t += s_cnt[q];
    s_base = t ? atomicAdd(counter, t) : 0u;
  }
  __syncthreads();
  unsigned off = s_base;
  for (int q = 0; q < w; ++q) off += s_cnt[q];
  unsigned r = off + (unsigned)__popcll(m & lt);
  __syncthreads();
  return r;
}
/* Append v (where want) to the sub-list of this block (LSLOTS sub-lists
 * of lcap entries, counters cnt[LSLOTS]); every thread of the block must
 * call it. */
__device__ __forceinline__ void list_append(uint32_t* data, unsigned* cnt,
                                            uint32_t lcap, bool want,
                                            uint32_t v) {
  const unsigned k = blockIdx.x & (LSLOTS - 1);
  const unsigned at = block_append(cnt + k, want);
  if (want && at < lcap) data[(uint64_t)k * lcap + at] = v;
}
/* A consumer block's table of the sub-lists' starts in the concatenation:
 * s_pre[k] = entries of sub-lists before k, s_pre[LSLOTS] = the total.
 * Every thread of the block must call it. */
__device__ __forceinline__ void list_table(const unsigned* cnt, uint32_t lcap,
                                           uint32_t* s_pre) {
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    uint32_t c = lane < LSLOTS ? min(cnt[lane], lcap) : 0u;
    uint32_t incl = c;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = (uint32_t)__shfl_up((int)incl, d, 64);
      if (lane >= d) incl += y;
    }
    if (lane < LSLOTS) s_pre[lane + 1] = incl;
    if (lane == 0) s_pre[0] = 0;
  }
  __syncthreads();
}
/* entry w of the concatenation (w < s_pre[LSLOTS]) */
__device__ __forceinline__ uint32_t list_at(const uint32_t* data, uint32_t lcap,
                                            const uint32_t* s_pre, uint32_t w) {
  int lo = 0, hi = LSLOTS - 1; /* last k with s_pre[k] <= w */
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (s_pre[mid] <= w) lo = mid;
    else hi = mid - 1;
  }
  return data[(uint64_t)lo * lcap + (w - s_pre[lo])];
}
/* Loads of a u32 array through a buffer resource: one 32-bit byte offset
 * per load instead of a 64-bit address, for arrays below 4 GiB (offsets
 * at or past the array's end read 0).  kBufWord3 is the descriptor's
 * fourth word for raw 32-bit access on gfx9-family parts (data format
 * 32). */
constexpr int kBufWord3 = 0x00020000;
struct BufU32 {
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ BufU32(const uint32_t* p, uint32_t bytes)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t*>(p), (short)0,
                                            (int)bytes, kBufWord3)) {}
  __device__ __forceinline__ uint32_t at(uint32_t byte_off) const {
    return __builtin_amdgcn_raw_buffer_load_b32(r, (int)byte_off, 0, 0);
  }
};

/* Block-staged output of a one-thread-per-segment enumeration.  The
 * segments of a block own the contiguous output range [base, base+total)
 * (their exclusive-scan offsets); each thread's entries go to LDS at their
 * offset within that range and the block writes the range out coalesced,
 * instead of 64 lanes each streaming into its own region.  Entries past
 * the LDS capacity CAP are stored directly. */
template <int NARR, int CAP>
struct BlockOut {
  uint32_t* s; /* LDS, NARR x CAP */
  uint64_t base, total;
  uint64_t lim; /* capacity of the output arrays: entries at or past it are dropped */
  __device__ __forceinline__ void put(uint32_t* const g[NARR], uint64_t o,
                                      const uint32_t v[NARR]) {
    if (o >= lim) return;
    uint64_t lo = o - base;
    if (lo < (uint64_t)CAP) {
#pragma unroll
      for (int a = 0; a < NARR; ++a) s[a * CAP + lo] = v[a];
    } else {
#pragma unroll
      for (int a = 0; a < NARR; ++a) g[a][o] = v[a];
    }
  }
  /* every thread of the block calls this once, after its last put */
  __device__ __forceinline__ void flush(uint32_t* const g[NARR]) {
    __syncthreads();
    uint64_t m64 = total < (uint64_t)CAP ? total : (uint64_t)CAP;
    if (base >= lim) m64 = 0;
    else if (m64 > lim - base) m64 = lim - base;
    const uint32_t m = (uint32_t)m64;
    for (uint32_t t = threadIdx.x; t < m; t += blockDim.x)
#pragma unroll
      for (int a = 0; a < NARR; ++a) g[a][base + t] = s[a * CAP + t];
  }
};

/* the same for one array of uint4 entries */
template <int CAP>
struct BlockOut4 {
  uint4* s; /* LDS, CAP */
  uint64_t base, total, lim;
  __device__ __forceinline__ void put(uint4* g, uint64_t o, const uint4& v) {
    if (o >= lim) return;
    uint64_t lo = o - base;
    if (lo < (uint64_t)CAP) s[lo] = v;
    else g[o] = v;
  }
  __device__ __forceinline__ void flush(uint4* g) {
    __syncthreads();
    uint64_t m64 = total < (uint64_t)CAP ? total : (uint64_t)CAP;
    if (base >= lim) m64 = 0;
    else if (m64 > lim - base) m64 = lim - base;
    const uint32_t m = (uint32_t)m64;
    for (uint32_t t = threadIdx.x; t < m; t += blockDim.x) g[base + t] = s[t];
  }
};

struct StageCnt {
  unsigned c[4];
  __device__ StageCnt() { c[0] = c[1] = c[2] = c[3] = 0; }
  __device__ void flush(unsigned long long* ctr, int fam) {
#pragma unroll
    for (int s = 0; s < 4; ++s) flush_ctr(ctr, CTR_STAGE + fam * 4 + s, c[s]);
  }
};

__device__ __forceinline__ int64_t gtid() {
  return (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
}
/* Persistent launch over [0, total): the grid's waves take `chunk`
 * consecutive entries (a multiple of 64) at a time, wave w chunks w,
 * w + W, w + 2W, ... (W waves in the grid), each lane every 64th entry of
 * its chunk.  The chunk loop is wave-uniform.  The assignment is static: a
 * shared work counter cost one device-scope atomic per chunk, and with
 * ~10^4 chunks those serialized (second phase 1.57 -> 1.17 ms on the 100k
 * globule at budget 4 without it). */
template <class F>
__device__ __forceinline__ void for_work(unsigned chunk, int64_t total,
                                         F&& body) {
  const int lane = threadIdx.x & 63;
  const int64_t nw = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t base = (gtid() >> 6) * (int64_t)chunk; base < total;
       base += nw * (int64_t)chunk) {
    const int64_t end = base + chunk < total ? base + chunk : total;
    for (int64_t w = base + lane; w < end; w += 64) body(w);
  }
}
__device__ __forceinline__ bool same_pos(const Ball& a, const Ball& b) {
  return a.x == b.x && a.y == b.y && a.z == b.z;
}
__device__ __forceinline__ void push_q(const Queue& q, uint32_t item,
                                       uint32_t aux) {
  const uint32_t k = blockIdx.x & (QSLOTS - 1);
  unsigned p = atomicAdd(q.n + k, 1u);
  if (p < q.cap) q.q[(uint64_t)k * q.cap + p] = QEntry{item, aux};
}

/* A queued (simplex, ball) pair whose simplex another pair has already
 * rejected needs no work (flags only ever go from undecided to 0, so a
 * stale read merely does the work).  Entries with aux == ~0 are full
 * scans and always live. */
__device__ __forceinline__ bool live_entry(const QEntry& e,
                                           const uint8_t* flag) {
  return e.aux == 0xffffffffu || flag[e.item] != 0;
}

/* Entries of a queue for its consumer kernel: thread t takes entry t,
 * then t + grid, ... (a host view's grid covers its entries; a device
 * view's is at most a few hundred blocks). */
template <class F>
__device__ __forceinline__ void for_queue(const QRef& q, F&& body) {
  const uint32_t n = q.size();
  const uint32_t grid = gridDim.x * blockDim.x;
  const uint32_t t0 = (uint32_t)gtid();
  if (t0 < n) body(t0);
  /* a static stride: a shared work counter took one device-scope atomic
   * per 64 entries, and with tens of millions of escalated entries
   * (degenerate inputs) those serialized */
  for (uint32_t t = t0 + grid; t < n; t += grid) body(t);
}

template <class T>
struct Al;
template <class R>
struct Al<FE<R>> {
  static __device__ FE<R> c(Real a) { return FE<R>{(R)a, (R)0}; }
};
template <>
struct Al<FEDD> {
  static __device__ FEDD c(Real a) { return FEDD{dd{(double)a, 0.0}, 0.0}; }
};

enum { REJECT = 0, KEEP = 1, UNDEC = 2 };
__device__ __forceinline__ int code_of(int s) {
  return s > 0 ? KEEP : (s < 0 ? REJECT : UNDEC);
}
/* stage index of the precision type */
template <class T>
struct StageOf;
template <>
struct StageOf<FE<float>> {
  static constexpr int v = 0;
};
template <>
struct StageOf<FE<double>> {
  static constexpr int v = 1;
};
template <>
struct StageOf<FEDD> {
  static constexpr int v = 2;
};

/* ---- CSR lookups ---- */
/* Position of v in the sorted row a[lo, hi) or -1, reading the row eight
 * entries at a time: the eight loads issue together, one round trip per
 * eight entries where a binary search waits on every probe (the rows of
 * the alpha edge and triangle CSRs hold a few to a few tens of entries) */
__device__ __forceinline__ int64_t find_row(const uint32_t* a, uint64_t lo,
                                            uint64_t hi, uint32_t v) {
  for (uint64_t p = lo; p < hi; p += 8) {
    uint32_t x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = p + k < hi ? a[p + k] : 0xffffffffu;
    int hit = 8;
#pragma unroll
    for (int k = 7; k >= 0; --k)
      if (x[k] == v && p + k < hi) hit = k;
    if (hit < 8) return (int64_t)(p + hit);
    if (x[7] > v) return -1; /* sorted: v would have been in this batch */
  }
  return -1;
}

/* The same for two values at once (one pass over the row): the
 * positions of v and w (v < w) in a[lo, hi), each -1 when absent */
__device__ __forceinline__ void find_row2(const uint32_t* a, uint64_t lo,
                                          uint64_t hi, uint32_t v, uint32_t w,
                                          int64_t& pv, int64_t& pw) {
  pv = -1;
  pw = -1;
  for (uint64_t p = lo; p < hi; p += 8) {
    uint32_t x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = p + k < hi ? a[p + k] : 0xffffffffu;
#pragma unroll
    for (int k = 7; k >= 0; --k) {
      if (x[k] == v && p + k < hi) pv = (int64_t)(p + k);
      if (x[k] == w && p + k < hi) pw = (int64_t)(p + k);
    }
    if (x[7] > w) return; /* sorted: both would have been seen */
  }
}

/* ---- emptiness search box ----
 * Every ball that can violate a simplex lies within `reach` of its centre
 * c.  The centre is known to within err per axis and the FP32 cell keys of
 * the balls carry at most `margin` of rounding (length units), so the box
 * [c - reach - margin - err, c + reach + margin + err] in cells covers all
 * of them. */
struct CellBox {
  int lo[3], hi[3];
  int cc[3]; /* cell of the centre, clamped into the box */
};
/* grid of the molecule that ball `b` (sorted id) belongs to */
__device__ __forceinline__ GridP grid_of(const Stage& S, uint32_t b) {
  return S.mol ? S.grids[S.mol[b]] : S.g;
}
/* Rounding allowances of the box computations in the computing type C:
 * FP64 for the plain path and the escalation / exact kernels, FP32 for the
 * staged path's hot kernels.  `cell`: cells added on either side of a
 * box edge (relative to |index| + 1); `quot`: relative rounding of a
 * computed quotient; `add`: relative rounding of the centre's final
 * addition; `rel`: relative allowance on a certified upper bound. */
template <class C>
struct BoxEps;
template <>
struct BoxEps<double> {
  static constexpr double cell = 1e-9, quot = 1e-14, add = 0.0, rel = 1e-12;
};
template <>
struct BoxEps<float> {
  static constexpr float cell = 1e-4f, quot = 1e-6f, add = 2.4e-7f, rel = 1e-5f;
};
template <class C>
__device__ __forceinline__ bool make_box_r(const GridP& g, C reach, C cx, C cy,
                                           C cz, C ex, C ey, C ez, CellBox& b);
__device__ __forceinline__ bool make_box(const GridP& g, double cx, double cy,
                                         double cz, double ex, double ey,
                                         double ez, CellBox& b) {
  return make_box_r<double>(g, g.reach, cx, cy, cz, ex, ey, ez, b);
}
/* A simplex with power-centre c and alpha value rho is violated only by
 * balls with |p_e - c|^2 < rho + w_e <= rho + max w, so its own reach
 * sqrt(rho + max w) (rho certified from above) can replace the global
 * one (exact cell policy; the reference's stencil keeps its 27 cells). */
template <class C>
__device__ __forceinline__ C simplex_reach(const GridP& g, C rho_up) {
  const C full = (C)g.reach * (1 + BoxEps<C>::rel);
  if (!(g.maxw >= 0)) return full;
  const C e = BoxEps<C>::rel;
  const C r2 = rho_up + (C)g.maxw * (1 + e);
  const C r = r2 > 0 ? sqrt(r2) * (1 + e) + e : e;
  return r < full ? r : full;
}
template <class C>
__device__ __forceinline__ bool make_box_r(const GridP& g, C reach, C cx, C cy,
                                           C cz, C ex, C ey, C ez, CellBox& b) {
  const C margin = (C)g.margin, eps = BoxEps<C>::cell;
  C c[3] = {cx, cy, cz}, e[3] = {ex, ey, ez};
  C mn[3] = {(C)g.minx, (C)g.miny, (C)g.minz};
  int ext[3] = {g.ex, g.ey, g.ez};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const C inv = (C)(a == 0 ? g.inv_step : g.inv_yz);
    C r = reach + margin + e[a];
    C ql = (c[a] - r - mn[a]) * inv, qh = (c[a] + r - mn[a]) * inv;
    C qc = (c[a] - mn[a]) * inv;
    if (!(ql == ql) || !(qh == qh) || !(r < (C)1e30)) return false;
    C lo = floor(ql - eps * (fabs(ql) + 1));
    C hi = floor(qh + eps * (fabs(qh) + 1));
    if (lo < 0) lo = 0;
    if (hi > (C)(ext[a] - 1)) hi = (C)(ext[a] - 1);
    if (lo > hi) {
      b.lo[a] = 1;
      b.hi[a] = 0;
      b.cc[a] = 0;
      continue;
    }
    b.lo[a] = (int)lo;
    b.hi[a] = (int)hi;
    C f = floor(qc);
    b.cc[a] = f < lo ? (int)lo : (f > hi ? (int)hi : (int)f);
  }
  return true;
}
/* Visit every ball of the box, the centre cell first (the most likely
 * violators, as in AlphaComplexParallel.cu:632-643), then the rest in
 * row order.  visit(p) returns true to stop. */
template <class F>
__device__ __forceinline__ void scan_box(const GridP& g,
                                         const uint32_t* cstart,
                                         const CellBox& b, F&& visit) {
  if (b.lo[0] > b.hi[0] || b.lo[1] > b.hi[1] || b.lo[2] > b.hi[2]) return;
  {
    int64_t c = (int64_t)g.cbase + b.cc[0] +
                (int64_t)g.ex * (b.cc[1] + (int64_t)g.ey * b.cc[2]);
    for (uint32_t p = cstart[c], e = cstart[c + 1]; p < e; ++p)
      if (visit(p)) return;
  }
  for (int z = b.lo[2]; z <= b.hi[2]; ++z)
    for (int y = b.lo[1]; y <= b.hi[1]; ++y) {
      int64_t row = (int64_t)g.cbase + (int64_t)g.ex * (y + (int64_t)g.ey * z);
      bool crow = (z == b.cc[2] && y == b.cc[1]);
      if (!crow) {
        for (uint32_t p = cstart[row + b.lo[0]], e = cstart[row + b.hi[0] + 1];
             p < e; ++p)
          if (visit(p)) return;
      } else {
        for (uint32_t p = cstart[row + b.lo[0]], e = cstart[row + b.cc[0]];
             p < e; ++p)
          if (visit(p)) return;
        for (uint32_t p = cstart[row + b.cc[0] + 1],
                      e = cstart[row + b.hi[0] + 1];
             p < e; ++p)
          if (visit(p)) return;
      }
    }
}

/* Same traversal, balls fetched four at a time so that each row costs one
 * round trip to memory instead of one per ball; visit(p, ball) returns
 * true to stop.  Indices past the end of a range are clamped to its last
 * element (always a valid address) and not visited. */
template <class F>
__device__ __forceinline__ bool scan_range4(const Ball* ball, uint32_t b,
                                            uint32_t e, F& visit) {
  for (uint32_t p = b; p < e; p += 4) {
    uint32_t l = e - 1;
    Ball E0 = ball[p];
    Ball E1 = ball[min(p + 1, l)];
    Ball E2 = ball[min(p + 2, l)];
    Ball E3 = ball[min(p + 3, l)];
    if (visit(p, E0)) return true;
    if (p + 1 < e && visit(p + 1, E1)) return true;
    if (p + 2 < e && visit(p + 2, E2)) return true;
    if (p + 3 < e && visit(p + 3, E3)) return true;
  }
  return false;
}
/* The ranges of scan_box's traversal one at a time (centre cell, then the
 * rows with the centre cell cut out), so that a scan has a single loop
 * over ranges and its per-ball code is emitted once. */
struct BoxIter {
  int z, y, part; /* part: -1 centre cell, 0 whole row / left of centre, 1 right */
};
__device__ __forceinline__ bool box_next(const GridP& g, const uint32_t* cstart,
                                         const CellBox& b, BoxIter& it,
                                         uint32_t& lo, uint32_t& hi) {
  for (;;) {
    if (it.part == -1) {
      int64_t c = (int64_t)g.cbase + b.cc[0] +
                  (int64_t)g.ex * (b.cc[1] + (int64_t)g.ey * b.cc[2]);
      lo = cstart[c];
      hi = cstart[c + 1];
      it.part = 0;
      return true;
    }
    if (it.z > b.hi[2]) return false;
    const int64_t row =
        (int64_t)g.cbase + (int64_t)g.ex * (it.y + (int64_t)g.ey * it.z);
    const bool crow = it.z == b.cc[2] && it.y == b.cc[1];
    if (!crow) {
      lo = cstart[row + b.lo[0]];
      hi = cstart[row + b.hi[0] + 1];
    } else if (it.part == 0) {
      lo = cstart[row + b.lo[0]];
      hi = cstart[row + b.cc[0]];
    } else {
      lo = cstart[row + b.cc[0] + 1];
      hi = cstart[row + b.hi[0] + 1];
    }
    if (crow && it.part == 0) {
      it.part = 1;
    } else {
      it.part = 0;
      if (++it.y > b.hi[1]) {
        it.y = b.lo[1];
        ++it.z;
      }
    }
    if (hi > lo) return true;
  }
}
template <class F>
__device__ __forceinline__ void scan_box4(const GridP& g,
                                          const uint32_t* cstart,
                                          const Ball* ball, const CellBox& b,
                                          F&& visit) {
  if (b.lo[0] > b.hi[0] || b.lo[1] > b.hi[1] || b.lo[2] > b.hi[2]) return;
  BoxIter it{b.lo[2], b.lo[1], -1};
  uint32_t lo, hi;
  while (box_next(g, cstart, b, it, lo, hi))
    if (scan_range4(ball, lo, hi, visit)) return;
}

/* ---- emptiness scans over the fine grid ----
 * The cells visited are those of the certified box that meet the sphere
 * of radius R = reach + margin + |centre error| about the computed centre
 * (a violator lies within reach of the exact centre, which lies within the
 * error of the computed one): the box rows (y, z) whose cell square comes
 * within R, each cut to the sphere's x-extent over the row's nearest
 * point, the centre's row first.  Worked in cells relative to the
 * box's low corner, in FP32, with R inflated far past the rounding (every
 * ball lies inside its own fine cell: keys are computed in double). */
struct SphBox {
  CellBox b;
  float fx, fy, fz; /* centre minus the box's low corner, in cells */
  float R2;         /* squared radius in y-z cells, rounded up */
  float kx;         /* x cells per y-z cell */
  int cy, cz;       /* the box row nearest the centre */
};
__device__ __forceinline__ int clamp_cell(float f, int lo, int hi) {
  const float q = floorf(fminf(fmaxf(f, -1.0e6f), 1.0e6f));
  const int c = lo + (int)q;
  return c < lo ? lo : (c > hi ? hi : c);
}
template <class C>
__device__ __forceinline__ bool make_sph(const GridP& g, C reach, const C c[3],
                                         const C e[3], SphBox& s) {
  if (!make_box_r<C>(g, reach, c[0], c[1], c[2], e[0], e[1], e[2], s.b))
    return false;
  const C iyz = (C)g.inv_yz;
  const C el = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
  const C R = (reach + (C)g.margin + el) * iyz * (C)(1 + 1e-5) + (C)1e-3;
  s.fx = (float)((c[0] - (C)g.minx) * (C)g.inv_step - (C)s.b.lo[0]);
  s.fy = (float)((c[1] - (C)g.miny) * iyz - (C)s.b.lo[1]);
  s.fz = (float)((c[2] - (C)g.minz) * iyz - (C)s.b.lo[2]);
  s.R2 = (float)(R * R) * 1.0001f;
  s.kx = g.xsplit * 1.0001f;
  s.cy = clamp_cell(s.fy, s.b.lo[1], s.b.hi[1]);
  s.cz = clamp_cell(s.fz, s.b.lo[2], s.b.hi[2]);
  return true;
}
/* squared distance from f to the cell [q, q+1) */
__device__ __forceinline__ float axis_d2(float f, int q) {
  const float d = fmaxf(0.f, fmaxf((float)q - f, f - (float)(q + 1)));
  return d * d;
}
/* The fine-grid cells of a sphere box (above) in one flat loop: each
 * iteration fetches the next row's range when the current one is used up
 * and then hands the next (up to) four balls of the range to group(p, n)
 * (fine-grid positions p .. p+n-1; true stops the scan).  Rows: the
 * centre row first, then the box's other rows in (z, y) order.  Nested
 * row and ball loops make a wave step through every lane's rows in lock
 * step, each row as long as its longest lane's; here lanes at different
 * rows share iterations, and the per-ball code is emitted once. */
template <class F>
__device__ __forceinline__ void scan_sph_flat(const GridP& g,
                                              const uint32_t* fstart,
                                              const SphBox& s, F&& group) {
  const CellBox& b = s.b;
  if (b.lo[0] > b.hi[0] || b.lo[1] > b.hi[1] || b.lo[2] > b.hi[2]) return;
  int ry = s.cy, rz = s.cz;        /* the row fetched next */
  int ty = b.lo[1], tz = b.lo[2];  /* (z, y) order cursor, at the first row */
  bool more = true;
  uint32_t p = 0, hi = 0;
  for (;;) {
    if (p >= hi) {
      if (!more) break;
      const float d2 = axis_d2(s.fz, rz - b.lo[2]) + axis_d2(s.fy, ry - b.lo[1]);
      if (d2 <= s.R2) {
        /* the hardware square root (v_sqrt_f32, within 1 ulp): the row's
         * half-width in cells is padded by 1e-3 cells, far above an ulp of
         * the few cells it measures, so the cell range stays conservative;
         * the correctly rounded sqrtf costs ~15 more instructions per row */
        const float rx = __builtin_amdgcn_sqrtf(s.R2 - d2) * s.kx + 1e-3f;
        const int xl = max(0, (int)floorf(fmaxf(s.fx - rx, -1.0e6f)));
        const int xh =
            min(b.hi[0] - b.lo[0], (int)floorf(fminf(s.fx + rx, 1.0e6f)));
        if (xl <= xh) {
          /* fine-grid cell ids stay below 2^31 (the host uses the fine
           * grid only then): 32-bit unsigned index arithmetic */
          const uint32_t row =
              (uint32_t)((int)g.cbase + b.lo[0] + g.ex * (ry + g.ey * rz));
          p = fstart[row + (uint32_t)xl];
          hi = fstart[row + (uint32_t)xh + 1u];
        }
      }
      /* next in (z, y) order; the centre row's turn goes to the first row,
       * whose own turn (the cursor's start) is skipped */
      if (++ty > b.hi[1]) {
        ty = b.lo[1];
        ++tz;
      }
      more = tz <= b.hi[2];
      ry = ty;
      rz = tz;
      if (ry == s.cy && rz == s.cz) {
        ry = b.lo[1];
        rz = b.lo[2];
      }
    }
    if (p < hi) {
      const uint32_t n = min(hi - p, 4u);
      if (group(p, n)) break;
      p += n;
    }
  }
}
template <class C>
__device__ __forceinline__ bool place_box(const GridP& g, C reach, const C c[3],
                                          const C e[3], CellBox& b) {
  return make_box_r<C>(g, reach, c[0], c[1], c[2], e[0], e[1], e[2], b);
}
template <class C>
__device__ __forceinline__ bool place_box(const GridP& g, C reach, const C c[3],
                                          const C e[3], SphBox& b) {
  return make_sph<C>(g, reach, c, e, b);
}
__device__ __forceinline__ GridP sgrid_of(const Stage& S, uint32_t b) {
  return S.sgrids ? S.sgrids[S.mol[b]] : S.sg;
}
/* The same with the choice made at compile time: MOL (a batch) reads the
 * molecule's grid from its table; otherwise the single grid is read from
 * the kernel's arguments (scalar loads, once per wave).  The run-time
 * choice above becomes one generic (flat) pointer whose 88-byte grid is
 * reloaded per item through the vector memory path. */
template <bool MOL>
__device__ __forceinline__ GridP grid_at(const Stage& S, uint32_t b) {
  if constexpr (MOL) return S.grids[S.mol[b]];
  else return S.g;
}
template <bool MOL>
__device__ __forceinline__ GridP sgrid_at(const Stage& S, uint32_t b) {
  if constexpr (MOL) return S.sgrids[S.mol[b]];
  else return S.sg;
}

/* Centre c = p0 + num / (2 den) and per-axis error bounds from certified
 * num, den, with two divisions.  With D = |den|, eD its bound (D > eD):
 * |num/den - num'/den'| <= (eN + |N| eD / (D - eD)) / (D - eD), which is
 * at least quot_err's bound; 1e-14 |N| / (D - eD) covers the rounding of
 * the computed quotient.  Returns false when den is not certified. */
template <class C, class R>
__device__ __forceinline__ bool centre_err(const FE<R>& den,
                                           const V3<FE<R>>& num, const Ball& p0,
                                           const FE<R>& wa, C c[3], C e[3],
                                           C* rho_up) {
  const C rel = BoxEps<C>::rel;
  const C D = fabs((C)den.v), eD = (C)den.e;
  if (!(D > eD * (C)1.01)) return false;
  const C inv2 = (C)0.5 / (C)den.v, im = (C)0.5 / (D - eD) * (1 + rel);
  const C n[3] = {(C)num.x.v, (C)num.y.v, (C)num.z.v};
  const C en[3] = {(C)num.x.e, (C)num.y.e, (C)num.z.e};
  const C p[3] = {(C)p0.x, (C)p0.y, (C)p0.z};
  C nn = 0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    c[a] = p[a] + n[a] * inv2;
    e[a] = (en[a] + fabs(n[a]) * eD * (2 * im)) * im * (C)1.001 +
           fabs(n[a]) * im * BoxEps<C>::quot + fabs(c[a]) * BoxEps<C>::add;
    const C up = (fabs(n[a]) + en[a]) * im; /* >= |x_a| */
    nn += up * up;
  }
  /* rho = |x|^2 - w_a with x = num / (2 den) */
  *rho_up = nn * (1 + rel) - ((C)wa.v - (C)wa.e) * (1 - rel);
  return true;
}

/* per-axis error of x = num/(k den) from certified errors of num, den */
template <class R>
__device__ __forceinline__ double quot_err(const FE<R>& num, const FE<R>& den,
                                           double k) {
  double D = fabs((double)den.v), eD = (double)den.e;
  double N = fabs((double)num.v), eN = (double)num.e;
  if (!(D > eD * 1.01)) return INFINITY;
  return ((eN * D + N * eD) / ((D - eD) * D)) / k * 1.001 + (N / D) / k * 1e-14;
}
__device__ __forceinline__ double quot_err(const FEDD& num, const FEDD& den,
                                           double k) {
  double D = fabs(den.v.hi), eD = den.e;
  double N = fabs(num.v.hi), eN = num.e;
  if (!(D > eD * 1.01)) return INFINITY;
  return ((eN * D + N * eD) / ((D - eD) * D)) / k * 1.001 + (N / D) / k * 1e-14;
}

/* ---- semi-static power-gap test ----
 * For a simplex with vertex a, certified den and num (value + error bound)
 * and a ball e, evaluates g = L*den - y.num with y = p_e - p_a,
 * L = |y|^2 - w_e + w_a in precision R with the a-priori bound
 *   |g - g_exact| <= Labs*(c u |den| + e_den) + sum|y_i|(c u |num_i| + e_num_i)
 *                    + e_wa |den| + tiny,     Labs = |y|^2 + w_e + w_a,
 * c = 12 covering the <= 8 roundings on any path of the expression (y, the
 * squares and their sums, w_e = r_e^2, the products with den and num, the
 * final difference) plus second-order terms.  Returns +1 / -1 when the sign
 * is certified, 0 otherwise. */
/* fused multiply-add in the precision of R */
__device__ __forceinline__ float fmaR(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}
__device__ __forceinline__ double fmaR(double a, double b, double c) {
  return __builtin_fma(a, b, c);
}

template <class R>
struct PairSS {
  R ax, ay, az, wa;
  R den, nx, ny, nz; /* times sign(den) (exact): the gap below comes out
                      * as sign(den) g, so that -1 is a violator */
  R K1, K2, Mx, My, Mz;
  int sden;
};
template <class R>
__device__ __forceinline__ bool pairss_init(PairSS<R>& P, const V3<FE<R>>& pa,
                                            const FE<R>& wa, const FE<R>& den,
                                            const V3<FE<R>>& num) {
  P.sden = fe_sign(den);
  if (!P.sden) return false;
  const R cu = (R)12 * FEc<R>::u();
  P.ax = pa.x.v;
  P.ay = pa.y.v;
  P.az = pa.z.v;
  P.wa = wa.v;
  const bool neg = P.sden < 0;
  P.den = neg ? -den.v : den.v;
  P.nx = neg ? -num.x.v : num.x.v;
  P.ny = neg ? -num.y.v : num.y.v;
  P.nz = neg ? -num.z.v : num.z.v;
  R sl = FEc<R>::slack();
  P.K1 = (cu * fabs(den.v) + den.e) * sl;
  P.K2 = (wa.e * fabs(den.v)) * sl + (R)64 * FEc<R>::tiny();
  P.Mx = (cu * fabs(num.x.v) + num.x.e) * sl;
  P.My = (cu * fabs(num.y.v) + num.y.e) * sl;
  P.Mz = (cu * fabs(num.z.v) + num.z.e) * sl;
  return true;
}
/* Semi-static setup: D, N (tets) or den, num (triangles, edges) evaluated
 * once in precision R together with the same expressions on absolute
 * values (perm); with leaves y = p - p_a and w = r^2 each rounded once, a
 * tree of height h has |computed - exact| <= gamma_{h} * perm, and
 * gamma_{h} <= (h + 1) u for h u << 1.  Heights used: D 6, N 10 (tets);
 * |n|^2 6, X 10 (triangles); |u|^2 4, A u 8 (edges), rounded up. */
template <class R>
struct SS {
  R v, p; /* value, permanent */
};
template <class R>
__device__ __forceinline__ FE<R> ss_fe(const SS<R>& a, int h) {
  R g = (R)(h + 2) * FEc<R>::u();
  return FE<R>{a.v, g * a.p * FEc<R>::slack() + (R)64 * FEc<R>::tiny()};
}
template <class R>
__device__ __forceinline__ void cross_ss(R ax, R ay, R az, R bx, R by, R bz,
                                         SS<R>* c) {
  R aX = fabs(ax), aY = fabs(ay), aZ = fabs(az), bX = fabs(bx), bY = fabs(by),
    bZ = fabs(bz);
  c[0] = SS<R>{ay * bz - az * by, aY * bZ + aZ * bY};
  c[1] = SS<R>{az * bx - ax * bz, aZ * bX + aX * bZ};
  c[2] = SS<R>{ax * by - ay * bx, aX * bY + aY * bX};
}

/* The pair test's gap g = sign(den) (L den - y . num) for ball e and the
 * bound B on its rounding error.  Negating den and num negates every
 * rounded result exactly, so g is the unnormalised gap times sign(den),
 * bit for bit, and B is unchanged. */
template <class R>
__device__ __forceinline__ void pairss_gap(const PairSS<R>& P, const Ball& e,
                                           R& g, R& B) {
  /* fused multiply-adds round once where the bound counts two roundings,
   * so the bound still holds */
  R yx = (R)e.x - P.ax, yy = (R)e.y - P.ay, yz = (R)e.z - P.az;
  R q = fmaR(yx, yx, fmaR(yy, yy, yz * yz));
  R we = (R)e.w * (R)e.w;
  R L = (q - we) + P.wa;
  R yn = fmaR(yx, P.nx, fmaR(yy, P.ny, yz * P.nz));
  g = fmaR(L, P.den, -yn);
  R Labs = (q + we) + P.wa;
  B = fmaR(Labs, P.K1,
           fmaR(fabs(yx), P.Mx, fmaR(fabs(yy), P.My, fmaR(fabs(yz), P.Mz, P.K2))));
}
/* -1: e certainly violates (power below rho), 0: undecided, 1: not */
template <class R>
__device__ __forceinline__ int pairss_sign(const PairSS<R>& P,
                                           const Ball& e) {
  R g, B;
  pairss_gap(P, e, g, B);
  return g > B ? 1 : (g < -B ? -1 : 0);
}
/* The same for four balls as bit masks: bit k of viol when ball k
 * certainly violates, of tie when its test is undecided (|g| <= B, or a
 * NaN: pairss_sign's 0). */
template <class R>
__device__ __forceinline__ void pairss_masks(const PairSS<R>& P,
                                             const Ball E[4], unsigned& viol,
                                             unsigned& tie) {
  viol = 0u;
  tie = 0u;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    R g, B;
    pairss_gap(P, E[k], g, B);
    viol |= g < -B ? 1u << k : 0u;
    tie |= !(fabs(g) > B) ? 1u << k : 0u;
  }
}

/* Orthocentre numerators N and denominator D of a tet at precision R with
 * a-priori (semi-static) error bounds, and w_a = r_a^2 with its rounding
 * bound: the same D, N as tet_DN, each bounded by gamma_h times its
 * permanent instead of a running error per operation. */
template <class R>
struct TetCert {
  FE<R> D;
  V3<FE<R>> N;
  FE<R> wa;
};
template <class R>
__device__ __forceinline__ void tet_dn_ss(const Ball v[4], TetCert<R>& T) {
  const R ax = v[0].x, ay = v[0].y, az = v[0].z;
  R ux = (R)v[1].x - ax, uy = (R)v[1].y - ay, uz = (R)v[1].z - az;
  R vx = (R)v[2].x - ax, vy = (R)v[2].y - ay, vz = (R)v[2].z - az;
  R tx = (R)v[3].x - ax, ty = (R)v[3].y - ay, tz = (R)v[3].z - az;
  R wa = (R)v[0].w * (R)v[0].w;
  R w1 = (R)v[1].w * (R)v[1].w, w2 = (R)v[2].w * (R)v[2].w,
    w3 = (R)v[3].w * (R)v[3].w;
  /* the permanents of the cross products bounded by max norms:
   * |v_y t_z| + |v_z t_y| <= 2 |v|_inf |t|_inf for every component */
  const R mu = fmax(fabs(ux), fmax(fabs(uy), fabs(uz)));
  const R mv = fmax(fabs(vx), fmax(fabs(vy), fabs(vz)));
  const R mt = fmax(fabs(tx), fmax(fabs(ty), fabs(tz)));
  const R pvt = ((R)2 * mv) * mt, ptu = ((R)2 * mt) * mu, puv = ((R)2 * mu) * mv;
  const R vt[3] = {vy * tz - vz * ty, vz * tx - vx * tz, vx * ty - vy * tx};
  const R tu[3] = {ty * uz - tz * uy, tz * ux - tx * uz, tx * uy - ty * ux};
  const R uv[3] = {uy * vz - uz * vy, uz * vx - ux * vz, ux * vy - uy * vx};
  SS<R> D{ux * vt[0] + uy * vt[1] + uz * vt[2],
          ((fabs(ux) + fabs(uy)) + fabs(uz)) * pvt};
  R uu = ux * ux + uy * uy + uz * uz, vv = vx * vx + vy * vy + vz * vz,
    tt = tx * tx + ty * ty + tz * tz;
  SS<R> A{(uu + wa) - w1, (uu + wa) + w1}, B{(vv + wa) - w2, (vv + wa) + w2},
      C{(tt + wa) - w3, (tt + wa) + w3};
  const R np = A.p * pvt + B.p * ptu + C.p * puv;
  SS<R> N[3];
#pragma unroll
  for (int q = 0; q < 3; ++q)
    N[q] = SS<R>{A.v * vt[q] + B.v * tu[q] + C.v * uv[q], np};
  T.D = ss_fe(D, 6);
  T.N = V3<FE<R>>{ss_fe(N[0], 10), ss_fe(N[1], 10), ss_fe(N[2], 10)};
  T.wa = FE<R>{wa, FEc<R>::u() * wa * FEc<R>::slack()};
}
/* Sign of V = 4 D^2 (w_a + alpha) - |N|^2 (alpha exact) from the
 * certified D, N, w_a: 1 certainly > 0, -1 certainly < 0, 0 undecided.
 * One bound instead of a running error per operation:
 *  - the inputs' errors: |4 D*^2 (w*+a) - 4 D^2 (w+a)| <=
 *    4 (2|D| eD + eD^2)(|w+a| + ew) + 4 D^2 ew, and
 *    ||N*|^2 - |N|^2| <= sum_q (2|N_q| + eN_q) eN_q;
 *  - the evaluation's roundings: 4 D is exact, 4 D^2 (w+a) and |N|^2 each
 *    carry at most three roundings (gamma_3 / (1 - gamma_3) < 4u of their
 *    computed magnitudes), the difference one more (u / (1 - u) < 2u of
 *    the computed V);
 *  - 64 tiny for products that underflow; the slack covers the rounding
 *    of the bound's own evaluation. */
template <class R>
__device__ __forceinline__ int tet_val_ss(const TetCert<R>& T, R alpha) {
  const R u = FEc<R>::u();
  const R D = T.D.v, eD = T.D.e;
  const R s = T.wa.v + alpha, es = T.wa.e;
  const R t2 = ((R)4 * D) * D;
  const R t3 = t2 * s;
  const R nx = T.N.x.v, ny = T.N.y.v, nz = T.N.z.v;
  const R n2 = (nx * nx + ny * ny) + nz * nz;
  const R V = t3 - n2;
  const R eDD = ((R)2 * fabs(D) + eD) * eD;
  const R eN = ((R)2 * fabs(nx) + T.N.x.e) * T.N.x.e +
               ((R)2 * fabs(ny) + T.N.y.e) * T.N.y.e +
               ((R)2 * fabs(nz) + T.N.z.e) * T.N.z.e;
  const R ein = (R)4 * eDD * (fabs(s) + es) + t2 * es + eN;
  const R B = (ein + (R)4 * u * (fabs(t3) + n2) + (R)2 * u * fabs(V) +
               (R)64 * FEc<R>::tiny()) *
              FEc<R>::slack();
  return V > B ? 1 : (V < -B ? -1 : 0);
}

}  // namespace PAC_V
}  // namespace pac
#endif
