/*
 * k_pipeline.hip -- gfx950 kernels of the weighted alpha complex pipeline.
 *
 * One thread per ball / edge / triangle / tet, 256-thread blocks (four
 * wave64s).  Every sign decision goes through the precision ladder of
 * pac_arith.h: FP32 (staged path only) -> FP64 -> double-double (staged)
 * with certified error bounds; undecided cases are appended to a queue that
 * a wide-integer kernel resolves exactly.  Stage evaluations are counted per
 * predicate family with wave-aggregated atomics.
 *
 * Reference kernels restated (AlphaComplexLargeData.cu unless noted):
 *   k_cell_key           computeCellIndex :65-78
 *   k_edge_count/fill    computePossibleEdgeCount / fillPotentialEdges
 *                        :80-165 + markRealEdges :167-190 (fused)
 *   k_tri_count/fill     computePossibleTriangleCount / fillPossibleTriangles
 *                        :192-296 + markRealTriangles :298-348 (fused)
 *   k_tet_count/fill     computePossibleTetCount / fillPossibleTets :350-529
 *                        + markRealTets :531-591 (fused)
 *   k_tet_ac2            checkTetsOrthosphere :597-735
 *   k_mark_tet_faces     generateTrisInTets + markTrisForDeletion :737-805
 *                        (membership lookup instead of sort/unique)
 *   k_tri_ac2            checkDanglingTris :807-881
 *   k_mark_tri_edges     generateEdgesInTris + markEdgesForDeletion :883-926
 *   k_edge_ac2           checkDanglingEdges :928-1015
 */
#include "pac_internal.h"
#include "pac_sos_table.h"

namespace pac {

/* ------------------------------------------------------------------ */
/* small helpers                                                       */
/* ------------------------------------------------------------------ */
__device__ __forceinline__ unsigned wave_sum(unsigned v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}
__device__ __forceinline__ void flush_ctr(unsigned long long* ctr, int idx,
                                          unsigned v) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(&ctr[idx], (unsigned long long)v);
}
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
__device__ __forceinline__ bool same_pos(const float4& a, const float4& b) {
  return a.x == b.x && a.y == b.y && a.z == b.z;
}
__device__ __forceinline__ void push_q(QEntry* q, unsigned* qn, uint32_t cap,
                                       uint32_t item, uint32_t aux) {
  unsigned p = atomicAdd(qn, 1u);
  if (p < cap) q[p] = QEntry{item, aux};
}

template <class T>
struct Al;
template <class R>
struct Al<FE<R>> {
  static __device__ FE<R> c(float a) { return FE<R>{(R)a, (R)0}; }
};
template <>
struct Al<FEDD> {
  static __device__ FEDD c(float a) { return FEDD{dd{(double)a, 0.0}, 0.0}; }
};

/* decision codes */
enum { REJECT = 0, KEEP = 1, UNDEC = 2 };
__device__ __forceinline__ int code_of(int s) {
  return s > 0 ? KEEP : (s < 0 ? REJECT : UNDEC);
}

/* ------------------------------------------------------------------ */
/* bounds / keys / gather                                              */
/* ------------------------------------------------------------------ */
__device__ __forceinline__ int neg_lsb_exp(float v) {
  uint32_t b = __float_as_uint(v) & 0x7fffffffu;
  if (b == 0) return 0;
  uint32_t E = b >> 23, m = b & 0x7fffffu;
  int e;
  if (E == 0) {
    e = -149;
  } else {
    m |= 0x800000u;
    e = (int)E - 150;
  }
  int lsb = e + __builtin_ctz(m);
  return lsb < 0 ? -lsb : 0;
}

struct BRed {
  float v[9];  /* minx miny minz -maxx -maxy -maxz -maxrawr -maxr -maxabs */
  int shift, nonfinite;
};
__device__ void bred_init(BRed& b) {
#pragma unroll
  for (int q = 0; q < 9; ++q) b.v[q] = INFINITY;
  b.shift = 0;
  b.nonfinite = 0;
}
__device__ void bred_merge(BRed& a, const BRed& b) {
#pragma unroll
  for (int q = 0; q < 9; ++q) a.v[q] = fminf(a.v[q], b.v[q]);
  a.shift = max(a.shift, b.shift);
  a.nonfinite += b.nonfinite;
}
__device__ void bred_block(BRed& mine, BRed* sh) {
  sh[threadIdx.x] = mine;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) bred_merge(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  mine = sh[0];
}

__global__ __launch_bounds__(kBlock) void k_bounds_part(const float4* raw,
                                                        int64_t n,
                                                        float solvent,
                                                        BRed* part) {
  __shared__ BRed sh[kBlock];
  BRed b;
  bred_init(b);
  for (int64_t i = gtid(); i < n; i += (int64_t)gridDim.x * blockDim.x) {
    float4 a = raw[i];
    if (!isfinite(a.x) || !isfinite(a.y) || !isfinite(a.z) || !isfinite(a.w)) {
      b.nonfinite++;
      continue;
    }
    float r = __fadd_rn(a.w, solvent);
    b.v[0] = fminf(b.v[0], a.x);
    b.v[1] = fminf(b.v[1], a.y);
    b.v[2] = fminf(b.v[2], a.z);
    b.v[3] = fminf(b.v[3], -a.x);
    b.v[4] = fminf(b.v[4], -a.y);
    b.v[5] = fminf(b.v[5], -a.z);
    b.v[6] = fminf(b.v[6], -a.w);
    b.v[7] = fminf(b.v[7], -fabsf(r));
    b.v[8] = fminf(b.v[8], -fmaxf(fabsf(a.x), fmaxf(fabsf(a.y), fabsf(a.z))));
    int s = max(max(neg_lsb_exp(a.x), neg_lsb_exp(a.y)),
                max(neg_lsb_exp(a.z), neg_lsb_exp(r)));
    b.shift = max(b.shift, s);
  }
  bred_block(b, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = b;
}
__global__ __launch_bounds__(kBlock) void k_bounds_final(const BRed* part,
                                                         int np, Bounds* out) {
  __shared__ BRed sh[kBlock];
  BRed b;
  bred_init(b);
  for (int i = threadIdx.x; i < np; i += blockDim.x) bred_merge(b, part[i]);
  bred_block(b, sh);
  if (threadIdx.x == 0) {
    Bounds o;
    o.minx = b.v[0];
    o.miny = b.v[1];
    o.minz = b.v[2];
    o.maxx = -b.v[3];
    o.maxy = -b.v[4];
    o.maxz = -b.v[5];
    o.max_raw_r = -b.v[6];
    o.max_r = -b.v[7];
    o.max_abs = -b.v[8];
    o.shift = b.shift;
    o.nonfinite = b.nonfinite;
    o.pad = 0;
    *out = o;
  }
}
void launch_bounds(const float4* raw, int64_t n, float solvent, Bounds* part,
                   Bounds* out, PipeLaunch L) {
  int nb = (int)std::min<int64_t>(1024, (n + kBlock - 1) / kBlock);
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(k_bounds_part, dim3(nb), dim3(kBlock), 0, L.s, raw, n,
                     solvent, (BRed*)part);
  hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(kBlock), 0, L.s,
                     (const BRed*)part, nb, out);
}

/* computeCellIndex (LD:65-78): FP32 subtract, IEEE divide, truncate. */
__global__ __launch_bounds__(kBlock) void k_cell_key(const float4* raw,
                                                     int64_t n, GridP g,
                                                     uint32_t* key,
                                                     uint32_t* idx) {
  int64_t i = gtid();
  if (i >= n) return;
  float4 a = raw[i];
  int cx = (int)__fdiv_rn(__fsub_rn(a.x, g.minx), g.step);
  int cy = (int)__fdiv_rn(__fsub_rn(a.y, g.miny), g.step);
  int cz = (int)__fdiv_rn(__fsub_rn(a.z, g.minz), g.step);
  key[i] = (uint32_t)(cx + g.ex * cy + g.ex * g.ey * cz);
  idx[i] = (uint32_t)i;
}
void launch_cell_key(const float4* raw, int64_t n, GridP g, uint32_t* key,
                     uint32_t* idx, PipeLaunch L) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_cell_key, dim3((unsigned)((n + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, L.s, raw, n, g, key, idx);
}

__global__ __launch_bounds__(kBlock) void k_gather(const float4* raw,
                                                   const uint32_t* sidx,
                                                   const uint32_t* skey,
                                                   int64_t n, float solvent,
                                                   float4* ball, uint32_t* perm,
                                                   uint32_t* cellk) {
  int64_t s = gtid();
  if (s >= n) return;
  uint32_t o = sidx[s];
  float4 a = raw[o];
  ball[s] = make_float4(a.x, a.y, a.z, __fadd_rn(a.w, solvent));
  perm[s] = o;
  cellk[s] = skey[s];
}
void launch_gather(const float4* raw, const uint32_t* sidx,
                   const uint32_t* skey, int64_t n, float solvent,
                   float4* ball, uint32_t* perm, uint32_t* cellk,
                   PipeLaunch L) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_gather, dim3((unsigned)((n + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, L.s, raw, sidx, skey, n, solvent, ball,
                     perm, cellk);
}

/* cstart[c] = first sorted ball with key >= c, for c in [0, G] */
__global__ __launch_bounds__(kBlock) void k_cell_starts(const uint32_t* cellk,
                                                        int64_t n, int64_t G,
                                                        uint32_t* cstart) {
  int64_t s = gtid();
  if (s > n) return;
  int64_t lo = (s == 0) ? 0 : (int64_t)cellk[s - 1] + 1;
  int64_t hi = (s == n) ? G : (int64_t)cellk[s];
  for (int64_t c = lo; c <= hi; ++c) cstart[c] = (uint32_t)s;
}
void launch_cell_starts(const uint32_t* cellk, int64_t n, int64_t G,
                        uint32_t* cstart, PipeLaunch L) {
  hipLaunchKernelGGL(k_cell_starts,
                     dim3((unsigned)((n + 1 + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, L.s, cellk, n, G, cstart);
}

/* ------------------------------------------------------------------ */
/* filter decisions                                                    */
/* ------------------------------------------------------------------ */
template <class T>
__device__ int edge_code(const float4& a, const float4& b, float alpha) {
  T v = edge_val(Ld<T>::p(a), Ld<T>::w(a), Ld<T>::p(b), Ld<T>::w(b),
                 Al<T>::c(alpha));
  return code_of(fe_sign(v));
}
template <bool STAGED>
__device__ int edge_decide(const float4& a, const float4& b, float alpha,
                           StageCnt& sc) {
  if (same_pos(a, b)) return REJECT;
  int r;
  if (STAGED) {
    sc.c[0]++;
    r = edge_code<FE<float>>(a, b, alpha);
    if (r != UNDEC) return r;
  }
  sc.c[1]++;
  r = edge_code<FE<double>>(a, b, alpha);
  if (r != UNDEC) return r;
  if (STAGED) {
    sc.c[2]++;
    r = edge_code<FEDD>(a, b, alpha);
  }
  return r;
}
__device__ int edge_exact(const float4& a, const float4& b,
                          const ExactScale& X) {
  if (same_pos(a, b)) return REJECT;
  LdWI L{X.S};
  WI v = edge_val(L.p(a), L.w(a), L.p(b), L.w(b), X.alpha);
  return wi_sign(v) >= 0 ? KEEP : REJECT;
}

template <class T>
__device__ int tri_code(const float4& a, const float4& b, const float4& c,
                        float alpha) {
  T nn;
  T v = tri_val(Ld<T>::p(a), Ld<T>::w(a), Ld<T>::p(b), Ld<T>::w(b),
                Ld<T>::p(c), Ld<T>::w(c), Al<T>::c(alpha), &nn);
  if (fe_sign(nn) <= 0) return UNDEC;
  return code_of(fe_sign(v));
}
template <bool STAGED>
__device__ int tri_decide(const float4& a, const float4& b, const float4& c,
                          float alpha, StageCnt& sc) {
  int r;
  if (STAGED) {
    sc.c[0]++;
    r = tri_code<FE<float>>(a, b, c, alpha);
    if (r != UNDEC) return r;
  }
  sc.c[1]++;
  r = tri_code<FE<double>>(a, b, c, alpha);
  if (r != UNDEC) return r;
  if (STAGED) {
    sc.c[2]++;
    r = tri_code<FEDD>(a, b, c, alpha);
  }
  return r;
}
__device__ int tri_exact(const float4& a, const float4& b, const float4& c,
                         const ExactScale& X) {
  LdWI L{X.S};
  WI nn;
  WI v = tri_val(L.p(a), L.w(a), L.p(b), L.w(b), L.p(c), L.w(c), X.alpha, &nn);
  if (wi_sign(nn) == 0) return REJECT;
  return wi_sign(v) >= 0 ? KEEP : REJECT;
}

template <class T>
__device__ int tet_code(const float4& a, const float4& b, const float4& c,
                        const float4& d, float alpha) {
  T D;
  V3<T> N;
  T wa = Ld<T>::w(a);
  tet_DN(Ld<T>::p(a), wa, Ld<T>::p(b), Ld<T>::w(b), Ld<T>::p(c), Ld<T>::w(c),
         Ld<T>::p(d), Ld<T>::w(d), &D, &N);
  if (fe_sign(D) == 0) return UNDEC;
  return code_of(fe_sign(tet_val(D, N, wa, Al<T>::c(alpha))));
}
template <bool STAGED>
__device__ int tet_decide(const float4& a, const float4& b, const float4& c,
                          const float4& d, float alpha, StageCnt& sc) {
  int r;
  if (STAGED) {
    sc.c[0]++;
    r = tet_code<FE<float>>(a, b, c, d, alpha);
    if (r != UNDEC) return r;
  }
  sc.c[1]++;
  r = tet_code<FE<double>>(a, b, c, d, alpha);
  if (r != UNDEC) return r;
  if (STAGED) {
    sc.c[2]++;
    r = tet_code<FEDD>(a, b, c, d, alpha);
  }
  return r;
}
__device__ int tet_exact(const float4& a, const float4& b, const float4& c,
                         const float4& d, const ExactScale& X) {
  LdWI L{X.S};
  WI D;
  V3<WI> N;
  WI wa = L.w(a);
  tet_DN(L.p(a), wa, L.p(b), L.w(b), L.p(c), L.w(c), L.p(d), L.w(d), &D, &N);
  if (wi_sign(D) == 0) return REJECT;
  return wi_sign(tet_val(D, N, wa, X.alpha)) >= 0 ? KEEP : REJECT;
}

/* ------------------------------------------------------------------ */
/* edges                                                               */
/* ------------------------------------------------------------------ */
__device__ __forceinline__ void cell_xyz(uint32_t c, const GridP& g, int& x,
                                         int& y, int& z) {
  x = (int)(c % (uint32_t)g.ex);
  uint32_t t = c / (uint32_t)g.ex;
  y = (int)(t % (uint32_t)g.ey);
  z = (int)(t / (uint32_t)g.ey);
}

__global__ __launch_bounds__(kBlock) void k_edge_count(Stage S, uint32_t* cnt) {
  int64_t i = gtid();
  if (i >= S.n) return;
  uint32_t c = S.cellk[i];
  int gx, gy, gz;
  cell_xyz(c, S.g, gx, gy, gz);
  const int k = S.g.k_enum;
  uint32_t total = 0;
  for (int z = max(gz - k, 0); z <= min(gz + k, S.g.ez - 1); ++z)
    for (int y = max(gy - k, 0); y <= min(gy + k, S.g.ey - 1); ++y) {
      int64_t row = (int64_t)S.g.ex * (y + (int64_t)S.g.ey * z);
      int x0 = max(gx - k, 0), x1 = min(gx + k, S.g.ex - 1);
      int64_t c0 = row + x0, c1 = row + x1;
      if (c1 < (int64_t)c) continue;
      uint32_t b = S.cstart[c0], e = S.cstart[c1 + 1];
      if (b < (uint32_t)i + 1) b = (uint32_t)i + 1;
      if (e > b) total += e - b;
    }
  cnt[i] = total;
}
void launch_edge_count(const Stage& S, uint32_t* cnt, PipeLaunch L) {
  if (S.n <= 0) return;
  hipLaunchKernelGGL(k_edge_count, dim3((unsigned)((S.n + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, L.s, S, cnt);
}

template <bool STAGED>
__global__ __launch_bounds__(kBlock) void k_edge_fill(
    Stage S, const uint64_t* off, uint32_t* cand, uint8_t* flag,
    uint32_t* keep, QEntry* q, unsigned* qn, uint32_t qcap) {
  int64_t i = gtid();
  StageCnt sc;
  if (i < S.n) {
    uint32_t c = S.cellk[i];
    int gx, gy, gz;
    cell_xyz(c, S.g, gx, gy, gz);
    const int k = S.g.k_enum;
    float4 a = S.ball[i];
    uint64_t o = off[i];
    uint32_t kept = 0;
    for (int z = max(gz - k, 0); z <= min(gz + k, S.g.ez - 1); ++z)
      for (int y = max(gy - k, 0); y <= min(gy + k, S.g.ey - 1); ++y) {
        int64_t row = (int64_t)S.g.ex * (y + (int64_t)S.g.ey * z);
        int x0 = max(gx - k, 0), x1 = min(gx + k, S.g.ex - 1);
        int64_t c0 = row + x0, c1 = row + x1;
        if (c1 < (int64_t)c) continue;
        uint32_t b = S.cstart[c0], e = S.cstart[c1 + 1];
        if (b < (uint32_t)i + 1) b = (uint32_t)i + 1;
        for (uint32_t j = b; j < e; ++j) {
          int r = edge_decide<STAGED>(a, S.ball[j], S.alpha, sc);
          cand[o] = j;
          flag[o] = (uint8_t)r;
          if (r == KEEP) ++kept;
          if (r == UNDEC) push_q(q, qn, qcap, (uint32_t)o, (uint32_t)i);
          ++o;
        }
      }
    keep[i] = kept;
  }
  sc.flush(S.ctr, FAM_EDGE);
}
void launch_edge_fill(const Stage& S, const uint64_t* off, uint32_t* cand,
                      uint8_t* flag, uint32_t* keep, QEntry* q,
                      unsigned int* qn, uint32_t qcap, PipeLaunch L) {
  if (S.n <= 0) return;
  dim3 gr((unsigned)((S.n + kBlock - 1) / kBlock));
  if (S.staged)
    hipLaunchKernelGGL(k_edge_fill<true>, gr, dim3(kBlock), 0, L.s, S, off,
                       cand, flag, keep, q, qn, qcap);
  else
    hipLaunchKernelGGL(k_edge_fill<false>, gr, dim3(kBlock), 0, L.s, S, off,
                       cand, flag, keep, q, qn, qcap);
}

__global__ __launch_bounds__(kBlock) void k_edge_exact(
    Stage S, const QEntry* q, uint32_t nq, const uint32_t* cand,
    uint8_t* flag, uint32_t* keep) {
  int64_t t = gtid();
  unsigned cnt = 0;
  if (t < nq) {
    QEntry e = q[t];
    int r = edge_exact(S.ball[e.aux], S.ball[cand[e.item]], S.ex);
    flag[e.item] = (uint8_t)r;
    if (r == KEEP) atomicAdd(&keep[e.aux], 1u);
    cnt = 1;
  }
  flush_ctr(S.ctr, CTR_STAGE + FAM_EDGE * 4 + 3, cnt);
}
void launch_edge_exact(const Stage& S, const QEntry* q, uint32_t nq,
                       const uint32_t* cand, uint8_t* flag, uint32_t* keep,
                       PipeLaunch L) {
  if (!nq) return;
  hipLaunchKernelGGL(k_edge_exact, dim3((nq + kBlock - 1) / kBlock),
                     dim3(kBlock), 0, L.s, S, q, nq, cand, flag, keep);
}

__global__ __launch_bounds__(kBlock) void k_edge_compact(
    int64_t n, const uint64_t* off, const uint32_t* cand, const uint8_t* flag,
    const uint64_t* erow, uint32_t* ecol, uint32_t* eleft) {
  int64_t i = gtid();
  if (i >= n) return;
  uint64_t w = erow[i];
  for (uint64_t p = off[i]; p < off[i + 1]; ++p)
    if (flag[p] == KEEP) {
      ecol[w] = cand[p];
      eleft[w] = (uint32_t)i;
      ++w;
    }
}
void launch_edge_compact(int64_t n, const uint64_t* off, const uint32_t* cand,
                         const uint8_t* flag, const uint64_t* erow,
                         uint32_t* ecol, uint32_t* eleft, PipeLaunch L) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_edge_compact, dim3((unsigned)((n + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, L.s, n, off, cand, flag, erow, ecol,
                     eleft);
}

/* ------------------------------------------------------------------ */
/* triangles: k in row(i) after j, intersected with row(j)             */
/* ------------------------------------------------------------------ */
__global__ __launch_bounds__(kBlock) void k_tri_count(
    int64_t ne, const uint64_t* erow, const uint32_t* ecol,
    const uint32_t* eleft, uint32_t* cnt) {
  int64_t q = gtid();
  if (q >= ne) return;
  uint32_t i = eleft[q], j = ecol[q];
  uint64_t a = (uint64_t)q + 1, ae = erow[i + 1];
  uint64_t b = erow[j], be = erow[j + 1];
  uint32_t c = 0;
  while (a < ae && b < be) {
    uint32_t x = ecol[a], y = ecol[b];
    c += (x == y);
    a += (x <= y);
    b += (y <= x);
  }
  cnt[q] = c;
}
void launch_tri_count(const Stage&, int64_t ne, const uint64_t* erow,
                      const uint32_t* ecol, const uint32_t* eleft,
                      uint32_t* cnt, PipeLaunch L) {
  if (ne <= 0) return;
  hipLaunchKernelGGL(k_tri_count, dim3((unsigned)((ne + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, L.s, ne, erow, ecol, eleft, cnt);
}

template <bool STAGED>
__global__ __launch_bounds__(kBlock) void k_tri_fill(
    Stage S, int64_t ne, const uint64_t* erow, const uint32_t* ecol,
    const uint32_t* eleft, const uint64_t* off, uint32_t* cand, uint8_t* flag,
    uint32_t* keep, QEntry* qq, unsigned* qn, uint32_t qcap) {
  int64_t q = gtid();
  StageCnt sc;
  if (q < ne) {
    uint32_t i = eleft[q], j = ecol[q];
    float4 A = S.ball[i], B = S.ball[j];
    uint64_t a = (uint64_t)q + 1, ae = erow[i + 1];
    uint64_t b = erow[j], be = erow[j + 1];
    uint64_t o = off[q];
    uint32_t kept = 0;
    while (a < ae && b < be) {
      uint32_t x = ecol[a], y = ecol[b];
      if (x == y) {
        int r = tri_decide<STAGED>(A, B, S.ball[x], S.alpha, sc);
        cand[o] = x;
        flag[o] = (uint8_t)r;
        if (r == KEEP) ++kept;
        if (r == UNDEC) push_q(qq, qn, qcap, (uint32_t)o, (uint32_t)q);
        ++o;
      }
      a += (x <= y);
      b += (y <= x);
    }
    keep[q] = kept;
  }
  sc.flush(S.ctr, FAM_TRI);
}
void launch_tri_fill(const Stage& S, int64_t ne, const uint64_t* erow,
                     const uint32_t* ecol, const uint32_t* eleft,
                     const uint64_t* off, uint32_t* cand, uint8_t* flag,
                     uint32_t* keep, QEntry* q, unsigned int* qn,
                     uint32_t qcap, PipeLaunch L) {
  if (ne <= 0) return;
  dim3 gr((unsigned)((ne + kBlock - 1) / kBlock));
  if (S.staged)
    hipLaunchKernelGGL(k_tri_fill<true>, gr, dim3(kBlock), 0, L.s, S, ne, erow,
                       ecol, eleft, off, cand, flag, keep, q, qn, qcap);
  else
    hipLaunchKernelGGL(k_tri_fill<false>, gr, dim3(kBlock), 0, L.s, S, ne, erow,
                       ecol, eleft, off, cand, flag, keep, q, qn, qcap);
}

__global__ __launch_bounds__(kBlock) void k_tri_exact(
    Stage S, const QEntry* q, uint32_t nq, const uint32_t* ecol,
    const uint32_t* eleft, const uint32_t* cand, uint8_t* flag,
    uint32_t* keep) {
  int64_t t = gtid();
  unsigned cnt = 0;
  if (t < nq) {
    QEntry e = q[t];
    uint32_t i = eleft[e.aux], j = ecol[e.aux], k = cand[e.item];
    int r = tri_exact(S.ball[i], S.ball[j], S.ball[k], S.ex);
    flag[e.item] = (uint8_t)r;
    if (r == KEEP) atomicAdd(&keep[e.aux], 1u);
    cnt = 1;
  }
  flush_ctr(S.ctr, CTR_STAGE + FAM_TRI * 4 + 3, cnt);
}
void launch_tri_exact(const Stage& S, const QEntry* q, uint32_t nq,
                      const uint32_t* ecol, const uint32_t* eleft,
                      const uint32_t* cand, uint8_t* flag, uint32_t* keep,
                      PipeLaunch L) {
  if (!nq) return;
  hipLaunchKernelGGL(k_tri_exact, dim3((nq + kBlock - 1) / kBlock),
                     dim3(kBlock), 0, L.s, S, q, nq, ecol, eleft, cand, flag,
                     keep);
}

/* generic segmented compaction: segment s owns candidates [off[s],off[s+1]) */
__global__ __launch_bounds__(kBlock) void k_seg_compact(
    int64_t nseg, const uint64_t* off, const uint32_t* cand,
    const uint8_t* flag, const uint64_t* orow, uint32_t* ocol,
    uint32_t* oseg) {
  int64_t s = gtid();
  if (s >= nseg) return;
  uint64_t w = orow[s];
  for (uint64_t p = off[s]; p < off[s + 1]; ++p)
    if (flag[p] == KEEP) {
      ocol[w] = cand[p];
      if (oseg) oseg[w] = (uint32_t)s;
      ++w;
    }
}
void launch_seg_compact(int64_t nseg, const uint64_t* off,
                        const uint32_t* cand, const uint8_t* flag,
                        const uint64_t* orow, uint32_t* ocol, uint32_t* oseg,
                        PipeLaunch L) {
  if (nseg <= 0) return;
  hipLaunchKernelGGL(k_seg_compact,
                     dim3((unsigned)((nseg + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, L.s, nseg, off, cand, flag, orow, ocol,
                     oseg);
}

/* ------------------------------------------------------------------ */
/* tets: l in T(i,j) after k, T(i,k), T(j,k)                           */
/* ------------------------------------------------------------------ */
__device__ __forceinline__ int64_t find_in(const uint32_t* a, uint64_t lo,
                                           uint64_t hi, uint32_t v) {
  while (lo < hi) {
    uint64_t mid = (lo + hi) >> 1;
    uint32_t x = a[mid];
    if (x < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}
__device__ __forceinline__ int64_t find_edge(const uint64_t* erow,
                                             const uint32_t* ecol, uint32_t a,
                                             uint32_t b) {
  uint64_t lo = erow[a], hi = erow[a + 1];
  int64_t p = find_in(ecol, lo, hi, b);
  return ((uint64_t)p < hi && ecol[p] == b) ? p : -1;
}

struct TetWalk {
  uint64_t a, ae, b, be, c, ce;
};
__device__ __forceinline__ bool tet_walk_init(
    int64_t q, const uint64_t* erow, const uint32_t* ecol,
    const uint32_t* eleft, const uint64_t* trow, const uint32_t* tedge,
    const uint32_t* tk, uint32_t& i, uint32_t& j, uint32_t& k, TetWalk& w) {
  uint32_t eij = tedge[q];
  i = eleft[eij];
  j = ecol[eij];
  k = tk[q];
  w.a = (uint64_t)q + 1;
  w.ae = trow[eij + 1];
  if (w.a >= w.ae) return false;
  int64_t eik = find_edge(erow, ecol, i, k), ejk = find_edge(erow, ecol, j, k);
  if (eik < 0 || ejk < 0) return false;
  w.b = trow[eik];
  w.be = trow[eik + 1];
  w.c = trow[ejk];
  w.ce = trow[ejk + 1];
  return true;
}

__global__ __launch_bounds__(kBlock) void k_tet_count(
    int64_t nt, const uint64_t* erow, const uint32_t* ecol,
    const uint32_t* eleft, const uint64_t* trow, const uint32_t* tk,
    const uint32_t* tedge, uint32_t* cnt) {
  int64_t q = gtid();
  if (q >= nt) return;
  uint32_t i, j, k;
  TetWalk w;
  uint32_t c = 0;
  if (tet_walk_init(q, erow, ecol, eleft, trow, tedge, tk, i, j, k, w)) {
    while (w.a < w.ae && w.b < w.be && w.c < w.ce) {
      uint32_t x = tk[w.a], y = tk[w.b], z = tk[w.c];
      uint32_t m = max(x, max(y, z));
      if (x == m && y == m && z == m) {
        ++c;
        ++w.a;
        ++w.b;
        ++w.c;
      } else {
        w.a += (x < m);
        w.b += (y < m);
        w.c += (z < m);
      }
    }
  }
  cnt[q] = c;
}
void launch_tet_count(int64_t nt, const uint64_t* erow, const uint32_t* ecol,
                      const uint32_t* eleft, const uint64_t* trow,
                      const uint32_t* tk, const uint32_t* tedge,
                      uint32_t* cnt, PipeLaunch L) {
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_tet_count, dim3((unsigned)((nt + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, L.s, nt, erow, ecol, eleft, trow, tk,
                     tedge, cnt);
}

template <bool STAGED>
__global__ __launch_bounds__(kBlock) void k_tet_fill(
    Stage S, int64_t nt, const uint64_t* erow, const uint32_t* ecol,
    const uint32_t* eleft, const uint64_t* trow, const uint32_t* tk,
    const uint32_t* tedge, const uint64_t* off, uint32_t* cand, uint8_t* flag,
    uint32_t* keep, QEntry* qq, unsigned* qn, uint32_t qcap) {
  int64_t q = gtid();
  StageCnt sc;
  if (q < nt) {
    uint32_t i, j, k;
    TetWalk w;
    uint64_t o = off[q];
    uint32_t kept = 0;
    if (tet_walk_init(q, erow, ecol, eleft, trow, tedge, tk, i, j, k, w)) {
      float4 A = S.ball[i], B = S.ball[j], C = S.ball[k];
      while (w.a < w.ae && w.b < w.be && w.c < w.ce) {
        uint32_t x = tk[w.a], y = tk[w.b], z = tk[w.c];
        uint32_t m = max(x, max(y, z));
        if (x == m && y == m && z == m) {
          int r = tet_decide<STAGED>(A, B, C, S.ball[m], S.alpha, sc);
          cand[o] = m;
          flag[o] = (uint8_t)r;
          if (r == KEEP) ++kept;
          if (r == UNDEC) push_q(qq, qn, qcap, (uint32_t)o, (uint32_t)q);
          ++o;
          ++w.a;
          ++w.b;
          ++w.c;
        } else {
          w.a += (x < m);
          w.b += (y < m);
          w.c += (z < m);
        }
      }
    }
    keep[q] = kept;
  }
  sc.flush(S.ctr, FAM_TET);
}
void launch_tet_fill(const Stage& S, int64_t nt, const uint64_t* erow,
                     const uint32_t* ecol, const uint32_t* eleft,
                     const uint64_t* trow, const uint32_t* tk,
                     const uint32_t* tedge, const uint64_t* off,
                     uint32_t* cand, uint8_t* flag, uint32_t* keep, QEntry* q,
                     unsigned int* qn, uint32_t qcap, PipeLaunch L) {
  if (nt <= 0) return;
  dim3 gr((unsigned)((nt + kBlock - 1) / kBlock));
  if (S.staged)
    hipLaunchKernelGGL(k_tet_fill<true>, gr, dim3(kBlock), 0, L.s, S, nt, erow,
                       ecol, eleft, trow, tk, tedge, off, cand, flag, keep, q,
                       qn, qcap);
  else
    hipLaunchKernelGGL(k_tet_fill<false>, gr, dim3(kBlock), 0, L.s, S, nt, erow,
                       ecol, eleft, trow, tk, tedge, off, cand, flag, keep, q,
                       qn, qcap);
}

__global__ __launch_bounds__(kBlock) void k_tet_exact(
    Stage S, const QEntry* q, uint32_t nq, const uint32_t* ecol,
    const uint32_t* eleft, const uint32_t* tk, const uint32_t* tedge,
    const uint32_t* cand, uint8_t* flag, uint32_t* keep) {
  int64_t t = gtid();
  unsigned cnt = 0;
  if (t < nq) {
    QEntry e = q[t];
    uint32_t eij = tedge[e.aux];
    uint32_t i = eleft[eij], j = ecol[eij], k = tk[e.aux], l = cand[e.item];
    int r = tet_exact(S.ball[i], S.ball[j], S.ball[k], S.ball[l], S.ex);
    flag[e.item] = (uint8_t)r;
    if (r == KEEP) atomicAdd(&keep[e.aux], 1u);
    cnt = 1;
  }
  flush_ctr(S.ctr, CTR_STAGE + FAM_TET * 4 + 3, cnt);
}
void launch_tet_exact(const Stage& S, const QEntry* q, uint32_t nq,
                      const uint32_t* ecol, const uint32_t* eleft,
                      const uint32_t* tk, const uint32_t* tedge,
                      const uint32_t* cand, uint8_t* flag, uint32_t* keep,
                      PipeLaunch L) {
  if (!nq) return;
  hipLaunchKernelGGL(k_tet_exact, dim3((nq + kBlock - 1) / kBlock),
                     dim3(kBlock), 0, L.s, S, q, nq, ecol, eleft, tk, tedge,
                     cand, flag, keep);
}

__global__ __launch_bounds__(kBlock) void k_tet_compact(
    int64_t nt, const uint64_t* off, const uint32_t* cand, const uint8_t* flag,
    const uint64_t* orow, const uint32_t* ecol, const uint32_t* eleft,
    const uint32_t* tk, const uint32_t* tedge, uint4* tets) {
  int64_t q = gtid();
  if (q >= nt) return;
  uint64_t w = orow[q];
  uint64_t p = off[q], pe = off[q + 1];
  if (p == pe) return;
  uint32_t eij = tedge[q];
  uint32_t i = eleft[eij], j = ecol[eij], k = tk[q];
  for (; p < pe; ++p)
    if (flag[p] == KEEP) tets[w++] = make_uint4(i, j, k, cand[p]);
}
void launch_tet_compact(int64_t nt, const uint64_t* off, const uint32_t* cand,
                        const uint8_t* flag, const uint64_t* orow,
                        const uint32_t* ecol, const uint32_t* eleft,
                        const uint32_t* tk, const uint32_t* tedge,
                        uint4* tets, PipeLaunch L) {
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_tet_compact, dim3((unsigned)((nt + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, L.s, nt, off, cand, flag, orow, ecol,
                     eleft, tk, tedge, tets);
}

/* ------------------------------------------------------------------ */
/* emptiness scans (AC2)                                               */
/* ------------------------------------------------------------------ */
/* Cell box of all balls that can lie within `reach` of a centre known to
 * within err (per axis).  Keys carry FP32 rounding of at most key_margin
 * (length units), folded into GridP.step-relative slack by the host. */
struct CellBox {
  int lo[3], hi[3];
};
__device__ __forceinline__ bool make_box(const GridP& g, double reach,
                                         double margin, double cx, double cy,
                                         double cz, double ex, double ey,
                                         double ez, CellBox& b) {
  double c[3] = {cx, cy, cz}, e[3] = {ex, ey, ez};
  double mn[3] = {(double)g.minx, (double)g.miny, (double)g.minz};
  int ext[3] = {g.ex, g.ey, g.ez};
  double inv = 1.0 / (double)g.step;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double r = reach + margin + e[a];
    double ql = (c[a] - r - mn[a]) * inv, qh = (c[a] + r - mn[a]) * inv;
    double lo = floor(ql - 1e-9 * (fabs(ql) + 1.0));
    double hi = floor(qh + 1e-9 * (fabs(qh) + 1.0));
    if (!(lo == lo) || !(hi == hi)) return false;
    if (lo < 0) lo = 0;
    if (hi > ext[a] - 1) hi = ext[a] - 1;
    if (lo > hi) {
      b.lo[a] = 1;
      b.hi[a] = 0;
      continue;
    }
    b.lo[a] = (int)lo;
    b.hi[a] = (int)hi;
  }
  return true;
}

/* per-axis error of x = num/(k*den) from certified errors of num and den */
template <class R>
__device__ __forceinline__ double quot_err(const FE<R>& num, const FE<R>& den,
                                           double k) {
  double D = fabs((double)den.v), eD = (double)den.e;
  double N = fabs((double)num.v), eN = (double)num.e;
  if (D <= eD * 1.01) return INFINITY;
  return ((eN * D + N * eD) / ((D - eD) * D)) / k * 1.001 +
         (N / D) / k * 4.0 * (double)FEc<R>::u();
}
__device__ __forceinline__ double quot_err(const FEDD& num, const FEDD& den,
                                           double k) {
  double D = fabs(den.v.hi), eD = den.e;
  double N = fabs(num.v.hi), eN = num.e;
  if (D <= eD * 1.01) return INFINITY;
  return ((eN * D + N * eD) / ((D - eD) * D)) / k * 1.001 + (N / D) / k * 1e-15;
}
template <class R>
__device__ __forceinline__ double val_d(const FE<R>& a) {
  return (double)a.v;
}
__device__ __forceinline__ double val_d(const FEDD& a) { return a.v.hi; }

/* exact lifted row for SoS */
__device__ void lifted_row(const float4& b, const LdWI& L, WI out[4]) {
  V3<WI> p = L.p(b);
  out[0] = p.x;
  out[1] = p.y;
  out[2] = p.z;
  out[3] = p.x * p.x + p.y * p.y + p.z * p.z - L.w(b);
}
__device__ __forceinline__ WI ent(const WI M[5][4], int r, int c) {
  return c == 4 ? wi_from_i64(1) : M[r][c];
}
__device__ WI det2(const WI M[5][4], const int* r, const int* c) {
  return ent(M, r[0], c[0]) * ent(M, r[1], c[1]) -
         ent(M, r[0], c[1]) * ent(M, r[1], c[0]);
}
__device__ WI det3(const WI M[5][4], const int* r, const int* c) {
  WI acc = wi_from_i64(0);
  for (int q = 0; q < 3; ++q) {
    int cc[2], m = 0;
    for (int t = 0; t < 3; ++t)
      if (t != q) cc[m++] = c[t];
    WI term = ent(M, r[0], c[q]) * det2(M, r + 1, cc);
    acc = (q & 1) ? acc - term : acc + term;
  }
  return acc;
}
__device__ WI det4(const WI M[5][4], const int* r, const int* c) {
  WI acc = wi_from_i64(0);
  for (int q = 0; q < 4; ++q) {
    int cc[3], m = 0;
    for (int t = 0; t < 4; ++t)
      if (t != q) cc[m++] = c[t];
    WI term = ent(M, r[0], c[q]) * det3(M, r + 1, cc);
    acc = (q & 1) ? acc - term : acc + term;
  }
  return acc;
}

/* Sign of the SoS-perturbed det[x y z l 1] of the five balls, rows in
 * argument order, perturbation order given by ids (original input ids,
 * as the reference's Atom.index, AlphaComplexParallel.cu:1087). */
__device__ int sos_det5(const float4 b[5], const uint32_t ids[5],
                        const LdWI& L) {
  int ord[5] = {0, 1, 2, 3, 4};
  int swaps = 0;
  for (int a = 0; a < 5; ++a)
    for (int q = 0; q + 1 < 5 - a; ++q)
      if (ids[ord[q + 1]] < ids[ord[q]]) {
        int t = ord[q];
        ord[q] = ord[q + 1];
        ord[q + 1] = t;
        ++swaps;
      }
  WI M[5][4];
  for (int r = 0; r < 5; ++r) lifted_row(b[ord[r]], L, M[r]);
  int parity = (swaps & 1) ? -1 : 1;
  for (int t = 0; t < PAC_SOS_TERMS; ++t) {
    int rows[5], cols[5], m = 0, mc = 0;
    int rm = pac_sos_table[t][0], cm = pac_sos_table[t][1];
    for (int r = 0; r < 5; ++r)
      if (rm & (1 << r)) rows[m++] = r;
    for (int c = 0; c < 5; ++c)
      if (cm & (1 << c)) cols[mc++] = c;
    int s;
    if (m == 1)
      s = wi_sign(ent(M, rows[0], cols[0]));
    else if (m == 2)
      s = wi_sign(det2(M, rows, cols));
    else if (m == 3)
      s = wi_sign(det3(M, rows, cols));
    else
      s = wi_sign(det4(M, rows, cols));
    if (s != 0) return parity * pac_sos_table[t][2] * s;
  }
  return parity;
}

/* WI -> double (relative error < 2^-50), scaled back by 2^-S */
__device__ double wi_to_double(const WI& a, int S) {
  bool neg = (a.l[WL - 1] >> 31) != 0;
  WI m = neg ? wi_neg(a) : a;
  int top = WL - 1;
  while (top > 0 && m.l[top] == 0) --top;
  double v = 0;
  for (int q = top; q >= 0 && q >= top - 2; --q)
    v += __builtin_ldexp((double)m.l[q], 32 * q);
  v = __builtin_ldexp(v, -S);
  return neg ? -v : v;
}

/* ---------------- tet emptiness ---------------- */
template <class T>
struct TetQ {
  T D, wa;
  V3<T> pa, N;
};
template <class T>
__device__ __forceinline__ void tetq_init(TetQ<T>& t, const float4 v[4]) {
  t.pa = Ld<T>::p(v[0]);
  t.wa = Ld<T>::w(v[0]);
  tet_DN(t.pa, t.wa, Ld<T>::p(v[1]), Ld<T>::w(v[1]), Ld<T>::p(v[2]),
         Ld<T>::w(v[2]), Ld<T>::p(v[3]), Ld<T>::w(v[3]), &t.D, &t.N);
}
/* 1 violated, 0 not violated, 2 undecided (needs exact) */
template <class T>
__device__ __forceinline__ int tet_pair_code(const TetQ<T>& t, int sD,
                                             const float4& e) {
  T g = power_gap(t.pa, t.wa, Ld<T>::p(e), Ld<T>::w(e), t.D, t.N);
  int s = fe_sign(g);
  if (s == 0) return 2;
  return (s * sD) < 0 ? 1 : 0;
}

template <bool STAGED>
__global__ __launch_bounds__(kBlock) void k_tet_ac2(
    Stage S, int64_t nk, const uint4* tets, const uint32_t* perm,
    uint8_t* flag, QEntry* q, unsigned* qn, uint32_t qcap, double reach,
    double margin) {
  int64_t id = gtid();
  StageCnt sc;
  unsigned pairs = 0;
  if (id < nk) {
    uint4 t = tets[id];
    float4 v[4] = {S.ball[t.x], S.ball[t.y], S.ball[t.z], S.ball[t.w]};
    /* centre and its certified error from the first stage that certifies D */
    TetQ<FE<float>> t32;
    TetQ<FE<double>> t64;
    TetQ<FEDD> tdd;
    int have = 0; /* bit0: f32, bit1: f64, bit2: dd */
    int sD = 0;
    double cx = 0, cy = 0, cz = 0, errx = INFINITY, erry = INFINITY,
           errz = INFINITY;
    if (STAGED) {
      tetq_init(t32, v);
      have |= 1;
      sD = fe_sign(t32.D);
      if (sD) {
        double D2 = 2.0 * (double)t32.D.v;
        cx = (double)t32.pa.x.v + (double)t32.N.x.v / D2;
        cy = (double)t32.pa.y.v + (double)t32.N.y.v / D2;
        cz = (double)t32.pa.z.v + (double)t32.N.z.v / D2;
        errx = quot_err(t32.N.x, t32.D, 2.0);
        erry = quot_err(t32.N.y, t32.D, 2.0);
        errz = quot_err(t32.N.z, t32.D, 2.0);
      }
    }
    if (!sD || !STAGED) {
      tetq_init(t64, v);
      have |= 2;
      sD = fe_sign(t64.D);
      if (sD) {
        double D2 = 2.0 * t64.D.v;
        cx = t64.pa.x.v + t64.N.x.v / D2;
        cy = t64.pa.y.v + t64.N.y.v / D2;
        cz = t64.pa.z.v + t64.N.z.v / D2;
        errx = quot_err(t64.N.x, t64.D, 2.0);
        erry = quot_err(t64.N.y, t64.D, 2.0);
        errz = quot_err(t64.N.z, t64.D, 2.0);
      }
    }
    if (!sD && STAGED) {
      tetq_init(tdd, v);
      have |= 4;
      sD = fe_sign(tdd.D);
      if (sD) {
        double D2 = 2.0 * tdd.D.v.hi;
        cx = (double)v[0].x + tdd.N.x.v.hi / D2;
        cy = (double)v[0].y + tdd.N.y.v.hi / D2;
        cz = (double)v[0].z + tdd.N.z.v.hi / D2;
        errx = quot_err(tdd.N.x, tdd.D, 2.0);
        erry = quot_err(tdd.N.y, tdd.D, 2.0);
        errz = quot_err(tdd.N.z, tdd.D, 2.0);
      }
    }
    CellBox box;
    bool ok = sD != 0 &&
              make_box(S.g, reach, margin, cx, cy, cz, errx, erry, errz, box);
    uint8_t res = 1;
    if (!ok) {
      /* D not certified in floating point: resolve the tet exactly */
      push_q(q, qn, qcap, (uint32_t)id, 0xffffffffu);
      res = 2;
    } else {
      bool pending = false, dead = false;
      for (int z = box.lo[2]; z <= box.hi[2] && !dead; ++z)
        for (int y = box.lo[1]; y <= box.hi[1] && !dead; ++y) {
          int64_t row = (int64_t)S.g.ex * (y + (int64_t)S.g.ey * z);
          uint32_t b = S.cstart[row + box.lo[0]],
                   e = S.cstart[row + box.hi[0] + 1];
          for (uint32_t p = b; p < e; ++p) {
            if (p == t.x || p == t.y || p == t.z || p == t.w) continue;
            ++pairs;
            float4 E = S.ball[p];
            int c = 2;
            if (STAGED && (have & 1) && fe_sign(t32.D) != 0) {
              sc.c[0]++;
              c = tet_pair_code(t32, sD, E);
            }
            if (c == 2) {
              if (!(have & 2)) {
                tetq_init(t64, v);
                have |= 2;
              }
              if (fe_sign(t64.D) != 0) {
                sc.c[1]++;
                c = tet_pair_code(t64, sD, E);
              }
            }
            if (c == 2 && STAGED) {
              if (!(have & 4)) {
                tetq_init(tdd, v);
                have |= 4;
              }
              if (fe_sign(tdd.D) != 0) {
                sc.c[2]++;
                c = tet_pair_code(tdd, sD, E);
              }
            }
            if (c == 1) {
              dead = true;
              break;
            }
            if (c == 2) {
              pending = true;
              push_q(q, qn, qcap, (uint32_t)id, p);
            }
          }
        }
      res = dead ? 0 : (pending ? 2 : 1);
    }
    flag[id] = res;
  }
  sc.flush(S.ctr, FAM_TET_AC2);
  flush_ctr(S.ctr, CTR_PAIRS + 0, pairs);
}

/* exact resolution of one (tet, ball) pair, or a full exact scan when
 * aux == ~0 (tet whose orientation no float stage could certify) */
__device__ int tet_pair_exact(const Stage& S, const float4 v[4],
                              const uint32_t vid[4], const uint32_t* perm,
                              const WI& D, const V3<WI>& N, const V3<WI>& pa,
                              const WI& wa, uint32_t p, unsigned* sos) {
  LdWI L{S.ex.S};
  float4 E = S.ball[p];
  WI g = power_gap(pa, wa, L.p(E), L.w(E), D, N);
  int s = wi_sign(g), sD = wi_sign(D);
  if (s != 0) return (s * sD) < 0;
  /* exact tie: insphere_w = det[x y z l 1](a,b,c,d,e) with SoS; the tet
   * is violated iff SoS * orient3d(a,b,c,d) > 0, orient3d = -sign(D)
   * (AlphaComplexParallel.cu:653-683). */
  ++*sos;
  float4 bb[5] = {v[0], v[1], v[2], v[3], E};
  uint32_t ids[5] = {perm[vid[0]], perm[vid[1]], perm[vid[2]], perm[vid[3]],
                     perm[p]};
  int ss = sos_det5(bb, ids, L);
  return (ss * -sD) > 0;
}

__global__ __launch_bounds__(kBlock) void k_tet_ac2_exact(
    Stage S, const QEntry* q, uint32_t nq, const uint4* tets,
    const uint32_t* perm, uint8_t* flag, double reach, double margin) {
  int64_t t = gtid();
  unsigned cnt = 0, sos = 0;
  if (t < nq) {
    QEntry e = q[t];
    uint4 tt = tets[e.item];
    uint32_t vid[4] = {tt.x, tt.y, tt.z, tt.w};
    float4 v[4] = {S.ball[tt.x], S.ball[tt.y], S.ball[tt.z], S.ball[tt.w]};
    LdWI L{S.ex.S};
    V3<WI> pa = L.p(v[0]);
    WI wa = L.w(v[0]), D;
    V3<WI> N;
    tet_DN(pa, wa, L.p(v[1]), L.w(v[1]), L.p(v[2]), L.w(v[2]), L.p(v[3]),
           L.w(v[3]), &D, &N);
    if (e.aux != 0xffffffffu) {
      cnt = 1;
      if (tet_pair_exact(S, v, vid, perm, D, N, pa, wa, e.aux, &sos))
        flag[e.item] = 0;
    } else if (wi_sign(D) != 0) {
      double D2 = 2.0 * wi_to_double(D, 3 * S.ex.S);
      double cx = (double)v[0].x + wi_to_double(N.x, 4 * S.ex.S) / D2;
      double cy = (double)v[0].y + wi_to_double(N.y, 4 * S.ex.S) / D2;
      double cz = (double)v[0].z + wi_to_double(N.z, 4 * S.ex.S) / D2;
      double err = 1e-9 * (fabs(cx) + fabs(cy) + fabs(cz) + 1.0);
      CellBox box;
      bool viol = false;
      if (make_box(S.g, reach, margin, cx, cy, cz, err, err, err, box)) {
        for (int z = box.lo[2]; z <= box.hi[2] && !viol; ++z)
          for (int y = box.lo[1]; y <= box.hi[1] && !viol; ++y) {
            int64_t row = (int64_t)S.g.ex * (y + (int64_t)S.g.ey * z);
            uint32_t b = S.cstart[row + box.lo[0]],
                     en = S.cstart[row + box.hi[0] + 1];
            for (uint32_t p = b; p < en; ++p) {
              if (p == tt.x || p == tt.y || p == tt.z || p == tt.w) continue;
              ++cnt;
              if (tet_pair_exact(S, v, vid, perm, D, N, pa, wa, p, &sos)) {
                viol = true;
                break;
              }
            }
          }
      }
      flag[e.item] = viol ? 0 : 2;
    }
  }
  flush_ctr(S.ctr, CTR_STAGE + FAM_TET_AC2 * 4 + 3, cnt);
  flush_ctr(S.ctr, CTR_SOS, sos);
}

/* ---------------- triangle / edge emptiness ---------------- */
template <class T>
struct LowQ {   /* triangle or edge: power gap with den, num */
  T den, wa;
  V3<T> pa, num;
};
template <class T>
__device__ __forceinline__ void triq_init(LowQ<T>& t, const float4& a,
                                          const float4& b, const float4& c) {
  t.pa = Ld<T>::p(a);
  t.wa = Ld<T>::w(a);
  tri_X(t.pa, t.wa, Ld<T>::p(b), Ld<T>::w(b), Ld<T>::p(c), Ld<T>::w(c),
        &t.den, &t.num);
}
template <class T>
__device__ __forceinline__ void edgeq_init(LowQ<T>& t, const float4& a,
                                           const float4& b) {
  t.pa = Ld<T>::p(a);
  t.wa = Ld<T>::w(a);
  V3<T> u = vsub(Ld<T>::p(b), t.pa);
  t.den = vdot(u, u);
  T A = t.den + t.wa - Ld<T>::w(b);
  t.num = vscale(A, u);
}
/* 1 violated (gap <= 0), 0 clear, 2 undecided */
template <class T>
__device__ __forceinline__ int low_pair_code(const LowQ<T>& t,
                                             const float4& e) {
  T g = power_gap(t.pa, t.wa, Ld<T>::p(e), Ld<T>::w(e), t.den, t.num);
  int s = fe_sign(g);
  return s == 0 ? 2 : (s < 0 ? 1 : 0);
}

/* kind 0: triangle (3 vertices), kind 1: edge (2 vertices) */
template <bool STAGED, int KIND>
__device__ uint8_t low_ac2_scan(const Stage& S, const float4* v,
                                const uint32_t* vid, uint32_t item,
                                QEntry* q, unsigned* qn, uint32_t qcap,
                                double reach, double margin, StageCnt& sc,
                                unsigned& pairs) {
  LowQ<FE<float>> q32;
  LowQ<FE<double>> q64;
  LowQ<FEDD> qdd;
  int have = 0;
  auto init = [&](int which) {
    if (which == 1) {
      if (KIND == 0) triq_init(q32, v[0], v[1], v[2]); else edgeq_init(q32, v[0], v[1]);
    } else if (which == 2) {
      if (KIND == 0) triq_init(q64, v[0], v[1], v[2]); else edgeq_init(q64, v[0], v[1]);
    } else {
      if (KIND == 0) triq_init(qdd, v[0], v[1], v[2]); else edgeq_init(qdd, v[0], v[1]);
    }
    have |= which;
  };
  bool okc = false;
  double cx = 0, cy = 0, cz = 0, ex = 0, ey = 0, ez = 0;
  const double k = KIND == 0 ? 2.0 : 2.0;
  if (STAGED) {
    init(1);
    if (fe_sign(q32.den) > 0) {
      double d2 = k * (double)q32.den.v;
      cx = (double)q32.pa.x.v + (double)q32.num.x.v / d2;
      cy = (double)q32.pa.y.v + (double)q32.num.y.v / d2;
      cz = (double)q32.pa.z.v + (double)q32.num.z.v / d2;
      ex = quot_err(q32.num.x, q32.den, k);
      ey = quot_err(q32.num.y, q32.den, k);
      ez = quot_err(q32.num.z, q32.den, k);
      okc = true;
    }
  }
  if (!okc) {
    init(2);
    if (fe_sign(q64.den) > 0) {
      double d2 = k * q64.den.v;
      cx = q64.pa.x.v + q64.num.x.v / d2;
      cy = q64.pa.y.v + q64.num.y.v / d2;
      cz = q64.pa.z.v + q64.num.z.v / d2;
      ex = quot_err(q64.num.x, q64.den, k);
      ey = quot_err(q64.num.y, q64.den, k);
      ez = quot_err(q64.num.z, q64.den, k);
      okc = true;
    }
  }
  if (!okc && STAGED) {
    init(4);
    if (fe_sign(qdd.den) > 0) {
      double d2 = k * qdd.den.v.hi;
      cx = (double)v[0].x + qdd.num.x.v.hi / d2;
      cy = (double)v[0].y + qdd.num.y.v.hi / d2;
      cz = (double)v[0].z + qdd.num.z.v.hi / d2;
      ex = quot_err(qdd.num.x, qdd.den, k);
      ey = quot_err(qdd.num.y, qdd.den, k);
      ez = quot_err(qdd.num.z, qdd.den, k);
      okc = true;
    }
  }
  CellBox box;
  if (!okc || !make_box(S.g, reach, margin, cx, cy, cz, ex, ey, ez, box)) {
    push_q(q, qn, qcap, item, 0xffffffffu);
    return 2;
  }
  bool pending = false;
  for (int z = box.lo[2]; z <= box.hi[2]; ++z)
    for (int y = box.lo[1]; y <= box.hi[1]; ++y) {
      int64_t row = (int64_t)S.g.ex * (y + (int64_t)S.g.ey * z);
      uint32_t b = S.cstart[row + box.lo[0]], e = S.cstart[row + box.hi[0] + 1];
      for (uint32_t p = b; p < e; ++p) {
        if (p == vid[0] || p == vid[1] || (KIND == 0 && p == vid[2])) continue;
        ++pairs;
        float4 E = S.ball[p];
        int c = 2;
        if (STAGED && (have & 1) && fe_sign(q32.den) > 0) {
          sc.c[0]++;
          c = low_pair_code(q32, E);
        }
        if (c == 2) {
          if (!(have & 2)) init(2);
          if (fe_sign(q64.den) > 0) {
            sc.c[1]++;
            c = low_pair_code(q64, E);
          }
        }
        if (c == 2 && STAGED) {
          if (!(have & 4)) init(4);
          if (fe_sign(qdd.den) > 0) {
            sc.c[2]++;
            c = low_pair_code(qdd, E);
          }
        }
        if (c == 1) return 0;
        if (c == 2) {
          pending = true;
          push_q(q, qn, qcap, item, p);
        }
      }
    }
  return pending ? 2 : 1;
}

template <bool STAGED>
__global__ __launch_bounds__(kBlock) void k_tri_ac2(
    Stage S, int64_t nt, const uint32_t* ecol, const uint32_t* eleft,
    const uint32_t* tk, const uint32_t* tedge, const uint8_t* tri_in,
    uint8_t* flag, QEntry* q, unsigned* qn, uint32_t qcap, double reach,
    double margin) {
  int64_t id = gtid();
  StageCnt sc;
  unsigned pairs = 0;
  if (id < nt) {
    uint8_t res = 1;
    if (!tri_in[id]) {
      uint32_t e = tedge[id];
      uint32_t vid[3] = {eleft[e], ecol[e], tk[id]};
      float4 v[3] = {S.ball[vid[0]], S.ball[vid[1]], S.ball[vid[2]]};
      res = low_ac2_scan<STAGED, 0>(S, v, vid, (uint32_t)id, q, qn, qcap, reach,
                                    margin, sc, pairs);
    }
    flag[id] = res;
  }
  sc.flush(S.ctr, FAM_TRI_AC2);
  flush_ctr(S.ctr, CTR_PAIRS + 1, pairs);
}

template <bool STAGED>
__global__ __launch_bounds__(kBlock) void k_edge_ac2(
    Stage S, int64_t ne, const uint32_t* ecol, const uint32_t* eleft,
    const uint8_t* edge_in, uint8_t* flag, QEntry* q, unsigned* qn,
    uint32_t qcap, double reach, double margin) {
  int64_t id = gtid();
  StageCnt sc;
  unsigned pairs = 0;
  if (id < ne) {
    uint8_t res = 1;
    if (!edge_in[id]) {
      uint32_t vid[3] = {eleft[id], ecol[id], 0xffffffffu};
      float4 v[3] = {S.ball[vid[0]], S.ball[vid[1]], S.ball[vid[1]]};
      res = low_ac2_scan<STAGED, 1>(S, v, vid, (uint32_t)id, q, qn, qcap, reach,
                                    margin, sc, pairs);
    }
    flag[id] = res;
  }
  sc.flush(S.ctr, FAM_EDGE_AC2);
  flush_ctr(S.ctr, CTR_PAIRS + 2, pairs);
}

/* exact triangle / edge pair (or full scan when aux == ~0) */
template <int KIND>
__global__ __launch_bounds__(kBlock) void k_low_ac2_exact(
    Stage S, const QEntry* q, uint32_t nq, const uint32_t* ecol,
    const uint32_t* eleft, const uint32_t* tk, const uint32_t* tedge,
    uint8_t* flag, double reach, double margin) {
  int64_t t = gtid();
  unsigned cnt = 0;
  if (t < nq) {
    QEntry e = q[t];
    uint32_t vid[3];
    if (KIND == 0) {
      uint32_t ed = tedge[e.item];
      vid[0] = eleft[ed];
      vid[1] = ecol[ed];
      vid[2] = tk[e.item];
    } else {
      vid[0] = eleft[e.item];
      vid[1] = ecol[e.item];
      vid[2] = 0xffffffffu;
    }
    LdWI L{S.ex.S};
    float4 a = S.ball[vid[0]], b = S.ball[vid[1]];
    V3<WI> pa = L.p(a), num;
    WI wa = L.w(a), den;
    if (KIND == 0) {
      tri_X(pa, wa, L.p(b), L.w(b), L.p(S.ball[vid[2]]), L.w(S.ball[vid[2]]),
            &den, &num);
    } else {
      V3<WI> u = vsub(L.p(b), pa);
      den = vdot(u, u);
      WI A = den + wa - L.w(b);
      num = vscale(A, u);
    }
    if (e.aux != 0xffffffffu) {
      cnt = 1;
      float4 E = S.ball[e.aux];
      if (wi_sign(power_gap(pa, wa, L.p(E), L.w(E), den, num)) <= 0)
        flag[e.item] = 0;
    } else if (wi_sign(den) > 0) {
      double d2 = 2.0 * wi_to_double(den, (KIND == 0 ? 4 : 2) * S.ex.S);
      int sh = (KIND == 0 ? 5 : 3) * S.ex.S;
      double cx = (double)a.x + wi_to_double(num.x, sh) / d2;
      double cy = (double)a.y + wi_to_double(num.y, sh) / d2;
      double cz = (double)a.z + wi_to_double(num.z, sh) / d2;
      double err = 1e-9 * (fabs(cx) + fabs(cy) + fabs(cz) + 1.0);
      CellBox box;
      bool viol = false;
      if (make_box(S.g, reach, margin, cx, cy, cz, err, err, err, box)) {
        for (int z = box.lo[2]; z <= box.hi[2] && !viol; ++z)
          for (int y = box.lo[1]; y <= box.hi[1] && !viol; ++y) {
            int64_t row = (int64_t)S.g.ex * (y + (int64_t)S.g.ey * z);
            uint32_t bb = S.cstart[row + box.lo[0]],
                     en = S.cstart[row + box.hi[0] + 1];
            for (uint32_t p = bb; p < en; ++p) {
              if (p == vid[0] || p == vid[1] || (KIND == 0 && p == vid[2]))
                continue;
              ++cnt;
              float4 E = S.ball[p];
              if (wi_sign(power_gap(pa, wa, L.p(E), L.w(E), den, num)) <= 0) {
                viol = true;
                break;
              }
            }
          }
      }
      flag[e.item] = viol ? 0 : 2;
    } else {
      flag[e.item] = 0; /* degenerate simplex: cannot be an alpha simplex */
    }
  }
  flush_ctr(S.ctr, CTR_STAGE + (KIND == 0 ? FAM_TRI_AC2 : FAM_EDGE_AC2) * 4 + 3,
            cnt);
}

/* ---------------- launchers for emptiness ---------------- */
void launch_tet_ac2(const Stage& S, int64_t nk, const uint4* tets,
                    const uint32_t* perm, uint8_t* flag, QEntry* q,
                    unsigned int* qn, uint32_t qcap, PipeLaunch L) {
  if (!nk) return;
  dim3 gr((unsigned)((nk + kBlock - 1) / kBlock));
  if (S.staged)
    hipLaunchKernelGGL(k_tet_ac2<true>, gr, dim3(kBlock), 0, L.s, S, nk, tets,
                       perm, flag, q, qn, qcap, S.reach, S.margin);
  else
    hipLaunchKernelGGL(k_tet_ac2<false>, gr, dim3(kBlock), 0, L.s, S, nk, tets,
                       perm, flag, q, qn, qcap, S.reach, S.margin);
}
void launch_tet_ac2_exact(const Stage& S, const QEntry* q, uint32_t nq,
                          const uint4* tets, const uint32_t* perm,
                          uint8_t* flag, PipeLaunch L) {
  if (!nq) return;
  hipLaunchKernelGGL(k_tet_ac2_exact, dim3((nq + kBlock - 1) / kBlock),
                     dim3(kBlock), 0, L.s, S, q, nq, tets, perm, flag, S.reach,
                     S.margin);
}
void launch_tri_ac2(const Stage& S, int64_t nt, const uint32_t* ecol,
                    const uint32_t* eleft, const uint32_t* tk,
                    const uint32_t* tedge, const uint8_t* tri_in,
                    uint8_t* flag, QEntry* q, unsigned* qn, uint32_t qcap,
                    PipeLaunch L) {
  if (!nt) return;
  dim3 gr((unsigned)((nt + kBlock - 1) / kBlock));
  if (S.staged)
    hipLaunchKernelGGL(k_tri_ac2<true>, gr, dim3(kBlock), 0, L.s, S, nt, ecol,
                       eleft, tk, tedge, tri_in, flag, q, qn, qcap, S.reach,
                       S.margin);
  else
    hipLaunchKernelGGL(k_tri_ac2<false>, gr, dim3(kBlock), 0, L.s, S, nt, ecol,
                       eleft, tk, tedge, tri_in, flag, q, qn, qcap, S.reach,
                       S.margin);
}
void launch_tri_ac2_exact(const Stage& S, const QEntry* q, uint32_t nq,
                          const uint32_t* ecol, const uint32_t* eleft,
                          const uint32_t* tk, const uint32_t* tedge,
                          uint8_t* flag, PipeLaunch L) {
  if (!nq) return;
  hipLaunchKernelGGL(k_low_ac2_exact<0>, dim3((nq + kBlock - 1) / kBlock),
                     dim3(kBlock), 0, L.s, S, q, nq, ecol, eleft, tk, tedge,
                     flag, S.reach, S.margin);
}
void launch_edge_ac2(const Stage& S, int64_t ne, const uint32_t* ecol,
                     const uint32_t* eleft, const uint8_t* edge_in,
                     uint8_t* flag, QEntry* q, unsigned* qn, uint32_t qcap,
                     PipeLaunch L) {
  if (!ne) return;
  dim3 gr((unsigned)((ne + kBlock - 1) / kBlock));
  if (S.staged)
    hipLaunchKernelGGL(k_edge_ac2<true>, gr, dim3(kBlock), 0, L.s, S, ne, ecol,
                       eleft, edge_in, flag, q, qn, qcap, S.reach, S.margin);
  else
    hipLaunchKernelGGL(k_edge_ac2<false>, gr, dim3(kBlock), 0, L.s, S, ne, ecol,
                       eleft, edge_in, flag, q, qn, qcap, S.reach, S.margin);
}
void launch_edge_ac2_exact(const Stage& S, const QEntry* q, uint32_t nq,
                           const uint32_t* ecol, const uint32_t* eleft,
                           uint8_t* flag, PipeLaunch L) {
  if (!nq) return;
  hipLaunchKernelGGL(k_low_ac2_exact<1>, dim3((nq + kBlock - 1) / kBlock),
                     dim3(kBlock), 0, L.s, S, q, nq, ecol, eleft,
                     (const uint32_t*)nullptr, (const uint32_t*)nullptr, flag,
                     S.reach, S.margin);
}

/* ------------------------------------------------------------------ */
/* closure marking                                                     */
/* ------------------------------------------------------------------ */
__device__ __forceinline__ int64_t find_tri(const uint64_t* erow,
                                            const uint32_t* ecol,
                                            const uint64_t* trow,
                                            const uint32_t* tk, uint32_t a,
                                            uint32_t b, uint32_t c) {
  int64_t e = find_edge(erow, ecol, a, b);
  if (e < 0) return -1;
  uint64_t lo = trow[e], hi = trow[e + 1];
  int64_t p = find_in(tk, lo, hi, c);
  return ((uint64_t)p < hi && tk[p] == c) ? p : -1;
}
__global__ __launch_bounds__(kBlock) void k_mark_tet_faces(
    int64_t nk, const uint4* tets, const uint8_t* kflag, const uint64_t* erow,
    const uint32_t* ecol, const uint64_t* trow, const uint32_t* tk,
    uint8_t* tri_in) {
  int64_t id = gtid();
  if (id >= nk || !kflag[id]) return;
  uint4 t = tets[id];
  int64_t f;
  if ((f = find_tri(erow, ecol, trow, tk, t.x, t.y, t.z)) >= 0) tri_in[f] = 1;
  if ((f = find_tri(erow, ecol, trow, tk, t.x, t.y, t.w)) >= 0) tri_in[f] = 1;
  if ((f = find_tri(erow, ecol, trow, tk, t.x, t.z, t.w)) >= 0) tri_in[f] = 1;
  if ((f = find_tri(erow, ecol, trow, tk, t.y, t.z, t.w)) >= 0) tri_in[f] = 1;
}
void launch_mark_tet_faces(int64_t nk, const uint4* tets, const uint8_t* kflag,
                           const uint64_t* erow, const uint32_t* ecol,
                           const uint64_t* trow, const uint32_t* tk,
                           uint8_t* tri_in, PipeLaunch L) {
  if (!nk) return;
  hipLaunchKernelGGL(k_mark_tet_faces,
                     dim3((unsigned)((nk + kBlock - 1) / kBlock)), dim3(kBlock),
                     0, L.s, nk, tets, kflag, erow, ecol, trow, tk, tri_in);
}
__global__ __launch_bounds__(kBlock) void k_mark_tri_edges(
    int64_t nt, const uint32_t* ecol, const uint32_t* eleft,
    const uint32_t* tk, const uint32_t* tedge, const uint8_t* tflag,
    const uint64_t* erow, uint8_t* edge_in) {
  int64_t id = gtid();
  if (id >= nt || !tflag[id]) return;
  uint32_t e = tedge[id];
  uint32_t i = eleft[e], j = ecol[e], k = tk[id];
  edge_in[e] = 1;
  int64_t f;
  if ((f = find_edge(erow, ecol, i, k)) >= 0) edge_in[f] = 1;
  if ((f = find_edge(erow, ecol, j, k)) >= 0) edge_in[f] = 1;
}
void launch_mark_tri_edges(int64_t nt, const uint32_t* ecol,
                           const uint32_t* eleft, const uint32_t* tk,
                           const uint32_t* tedge, const uint8_t* tflag,
                           const uint64_t* erow, uint8_t* edge_in,
                           PipeLaunch L) {
  if (!nt) return;
  hipLaunchKernelGGL(k_mark_tri_edges,
                     dim3((unsigned)((nt + kBlock - 1) / kBlock)), dim3(kBlock),
                     0, L.s, nt, ecol, eleft, tk, tedge, tflag, erow, edge_in);
}

/* ------------------------------------------------------------------ */
/* emit final lists in original ids                                    */
/* ------------------------------------------------------------------ */
__global__ __launch_bounds__(kBlock) void k_emit_tets(
    int64_t nk, const uint4* tets, const uint8_t* flag, const uint64_t* pos,
    const uint32_t* perm, int32_t* out) {
  int64_t id = gtid();
  if (id >= nk || !flag[id]) return;
  uint4 t = tets[id];
  uint64_t w = pos[id] * 4;
  out[w] = (int32_t)perm[t.x];
  out[w + 1] = (int32_t)perm[t.y];
  out[w + 2] = (int32_t)perm[t.z];
  out[w + 3] = (int32_t)perm[t.w];
}
void launch_emit_tets(int64_t nk, const uint4* tets, const uint8_t* flag,
                      const uint64_t* pos, const uint32_t* perm, int32_t* out,
                      PipeLaunch L) {
  if (!nk) return;
  hipLaunchKernelGGL(k_emit_tets, dim3((unsigned)((nk + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, L.s, nk, tets, flag, pos, perm, out);
}
__global__ __launch_bounds__(kBlock) void k_emit_tris(
    int64_t nt, const uint32_t* ecol, const uint32_t* eleft,
    const uint32_t* tk, const uint32_t* tedge, const uint8_t* flag,
    const uint64_t* pos, const uint32_t* perm, int32_t* out) {
  int64_t id = gtid();
  if (id >= nt || !flag[id]) return;
  uint32_t e = tedge[id];
  uint64_t w = pos[id] * 3;
  out[w] = (int32_t)perm[eleft[e]];
  out[w + 1] = (int32_t)perm[ecol[e]];
  out[w + 2] = (int32_t)perm[tk[id]];
}
void launch_emit_tris(int64_t nt, const uint32_t* ecol, const uint32_t* eleft,
                      const uint32_t* tk, const uint32_t* tedge,
                      const uint8_t* flag, const uint64_t* pos,
                      const uint32_t* perm, int32_t* out, PipeLaunch L) {
  if (!nt) return;
  hipLaunchKernelGGL(k_emit_tris, dim3((unsigned)((nt + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, L.s, nt, ecol, eleft, tk, tedge, flag,
                     pos, perm, out);
}
__global__ __launch_bounds__(kBlock) void k_emit_edges(
    int64_t ne, const uint32_t* ecol, const uint32_t* eleft,
    const uint8_t* flag, const uint64_t* pos, const uint32_t* perm,
    int32_t* out) {
  int64_t id = gtid();
  if (id >= ne || !flag[id]) return;
  uint64_t w = pos[id] * 2;
  out[w] = (int32_t)perm[eleft[id]];
  out[w + 1] = (int32_t)perm[ecol[id]];
}
void launch_emit_edges(int64_t ne, const uint32_t* ecol, const uint32_t* eleft,
                       const uint8_t* flag, const uint64_t* pos,
                       const uint32_t* perm, int32_t* out, PipeLaunch L) {
  if (!ne) return;
  hipLaunchKernelGGL(k_emit_edges, dim3((unsigned)((ne + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, L.s, ne, ecol, eleft, flag, pos, perm,
                     out);
}

int sos_table_host(int32_t* out3, int cap) {
  for (int t = 0; t < PAC_SOS_TERMS && t < cap; ++t)
    for (int c = 0; c < 3; ++c) out3[3 * t + c] = pac_sos_table[t][c];
  return PAC_SOS_TERMS;
}

}  // namespace pac
