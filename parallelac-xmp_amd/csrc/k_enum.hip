/*
 * k_enum.hip -- binning, candidate enumeration and the alpha filters.
 *
 * One thread per ball / alpha edge / alpha triangle enumerates its
 * candidates (cells of the stencil, sorted-row intersections, three-way
 * merges), classifies each with one precision stage (FP32 on the staged
 * path, FP64 on the plain path) and appends undecided candidates to a
 * queue.  The escalation kernel re-decides them in FP64 then double-double;
 * what is still undecided goes to the wide-integer kernel.
 *
 * Reference kernels restated (AlphaComplexLargeData.cu unless noted):
 *   k_cell_key           computeCellIndex :65-78
 *   k_edge_count/fill    computePossibleEdgeCount / fillPotentialEdges
 *                        :80-165 + markRealEdges :167-190 (fused)
 *   k_tri_count/fill     computePossibleTriangleCount / fillPossibleTriangles
 *                        :192-296 + markRealTriangles :298-348 (fused)
 *   k_tet_count/fill     computePossibleTetCount / fillPossibleTets :350-529
 *                        + markRealTets :531-591 (fused)
 */
#include "pac_dev.h"
#include "pac_wide.h"

namespace pac {
namespace PAC_V {

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

__device__ __forceinline__ int neg_lsb_exp(double v) {
  uint64_t b = (uint64_t)__double_as_longlong(v) & 0x7fffffffffffffffull;
  if (b == 0) return 0;
  uint32_t E = (uint32_t)(b >> 52);
  uint64_t m = b & ((1ull << 52) - 1);
  int e;
  if (E == 0) {
    e = -1074;
  } else {
    m |= 1ull << 52;
    e = (int)E - 1075;
  }
  int lsb = e + __builtin_ctzll(m);
  return lsb < 0 ? -lsb : 0;
}
/* r = r_raw + solvent as the reference computes it (FP32 add,
 * AlphaComplexParallel.cu:1086), or in FP64 for FP64 balls */
__device__ __forceinline__ float add_r(float a, float s) { return __fadd_rn(a, s); }
__device__ __forceinline__ double add_r(double a, double s) { return a + s; }

struct BRed {
  Real v[9];  /* minx miny minz -maxx -maxy -maxz -maxrawr -maxr -maxabs */
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
  for (int q = 0; q < 9; ++q) a.v[q] = fmin(a.v[q], b.v[q]);
  a.shift = max(a.shift, b.shift);
  a.nonfinite += b.nonfinite;
}
__device__ void bred_ball(BRed& b, const Ball& a, Real solvent) {
  if (!isfinite(a.x) || !isfinite(a.y) || !isfinite(a.z) || !isfinite(a.w)) {
    b.nonfinite++;
    return;
  }
  Real r = add_r(a.w, solvent);
  b.v[0] = fmin(b.v[0], a.x);
  b.v[1] = fmin(b.v[1], a.y);
  b.v[2] = fmin(b.v[2], a.z);
  b.v[3] = fmin(b.v[3], -a.x);
  b.v[4] = fmin(b.v[4], -a.y);
  b.v[5] = fmin(b.v[5], -a.z);
  b.v[6] = fmin(b.v[6], -a.w);
  b.v[7] = fmin(b.v[7], -fabs(r));
  b.v[8] = fmin(b.v[8], -fmax(fabs(a.x), fmax(fabs(a.y), fabs(a.z))));
  int s = max(max(neg_lsb_exp(a.x), neg_lsb_exp(a.y)),
              max(neg_lsb_exp(a.z), neg_lsb_exp(r)));
  b.shift = max(b.shift, s);
}
__device__ Bounds bred_out(const BRed& b) {
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

  return o;
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

__global__ __launch_bounds__(kBlock) void k_bounds_part(const Ball* raw,
                                                        int64_t n,
                                                        Real solvent,
                                                        BRed* part) {
  __shared__ BRed sh[kBlock];
  BRed b;
  bred_init(b);
  for (int64_t i = gtid(); i < n; i += (int64_t)gridDim.x * blockDim.x)
    bred_ball(b, raw[i], solvent);
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
  if (threadIdx.x == 0) *out = bred_out(b);
}
void launch_bounds(const Ball* raw, int64_t n, Real solvent, Bounds* part,
                   Bounds* out, PipeLaunch L) {
  int nb = (int)std::min<int64_t>(1024, (n + kBlock - 1) / kBlock);
  if (nb < 1) nb = 1;
  hipLaunchKernelGGL(k_bounds_part, dim3(nb), dim3(kBlock), 0, L.s, raw, n,
                     solvent, (BRed*)part);
  hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(kBlock), 0, L.s,
                     (const BRed*)part, nb, out);
}

/* Per-molecule bounds of a batch: block m reduces the balls of molecule m
 * (a few hundred to tens of thousands each). */
__global__ __launch_bounds__(kBlock) void k_bounds_mol(const Ball* raw,
                                                       const uint32_t* moff,
                                                       Real solvent,
                                                       Bounds* out) {
  __shared__ BRed sh[kBlock];
  const uint32_t m = blockIdx.x, b0 = moff[m], b1 = moff[m + 1];
  BRed b;
  bred_init(b);
  for (uint32_t i = b0 + threadIdx.x; i < b1; i += blockDim.x)
    bred_ball(b, raw[i], solvent);
  bred_block(b, sh);
  if (threadIdx.x == 0) out[m] = bred_out(b);
}
void launch_bounds_mol(const Ball* raw, const uint32_t* moff, int64_t nmol,
                       Real solvent, Bounds* out, PipeLaunch L) {
  if (nmol <= 0) return;
  hipLaunchKernelGGL(k_bounds_mol, dim3((unsigned)nmol), dim3(kBlock), 0, L.s,
                     raw, moff, solvent, out);
}

__global__ __launch_bounds__(kBlock) void k_mol_ids(const uint32_t* moff,
                                                    int64_t nmol, int64_t n,
                                                    uint32_t* mol) {
  int64_t i = gtid();
  if (i >= n) return;
  /* last m with moff[m] <= i */
  int64_t lo = 0, hi = nmol - 1;
  while (lo < hi) {
    int64_t mid = (lo + hi + 1) >> 1;
    if ((int64_t)moff[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  mol[i] = (uint32_t)lo;
}
void launch_mol_ids(const uint32_t* moff, int64_t nmol, int64_t n,
                    uint32_t* mol, PipeLaunch L) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_mol_ids, dim3((unsigned)((n + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, L.s, moff, nmol, n, mol);
}

/* computeCellIndex (LD:65-78): FP32 subtract, IEEE divide, truncate; in a
 * batch the molecule's first global cell id is added. */
__global__ __launch_bounds__(kBlock) void k_cell_key(const Ball* raw,
                                                     int64_t n, GridP g1,
                                                     const GridP* grids,
                                                     const uint32_t* mol,
                                                     uint32_t* key,
                                                     uint32_t* idx) {
  int64_t i = gtid();
  if (i >= n) return;
  const GridP g = mol ? grids[mol[i]] : g1;
  Ball a = raw[i];
  /* FP64 balls are keyed by their FP32 roundings (monotone, so every key
   * stays inside the grid; the host widens the search margin by the
   * rounding) */
  int cx = (int)__fdiv_rn(__fsub_rn((float)a.x, g.minx), g.step);
  int cy = (int)__fdiv_rn(__fsub_rn((float)a.y, g.miny), g.step);
  int cz = (int)__fdiv_rn(__fsub_rn((float)a.z, g.minz), g.step);
  /* balls always key inside the grid planned from their own bounds; the
   * clamp only keeps a recorded pipeline in bounds after its input moved
   * (that run is then reported stale) */
  cx = min(max(cx, 0), g.ex - 1);
  cy = min(max(cy, 0), g.ey - 1);
  cz = min(max(cz, 0), g.ez - 1);
  key[i] = g.cbase + (uint32_t)(cx + g.ex * cy + g.ex * g.ey * cz);
  idx[i] = (uint32_t)i;
}
void launch_cell_key(const Ball* raw, int64_t n, GridP g,
                     const GridP* grids, const uint32_t* mol, uint32_t* key,
                     uint32_t* idx, PipeLaunch L) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_cell_key, dim3((unsigned)((n + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, L.s, raw, n, g, grids, mol, key, idx);
}

__global__ __launch_bounds__(kBlock) void k_gather(const Ball* raw,
                                                   const uint32_t* sidx,
                                                   const uint32_t* skey,
                                                   int64_t n, Real solvent,
                                                   Ball* ball, uint32_t* perm,
                                                   uint32_t* cellk) {
  int64_t s = gtid();
  if (s >= n) return;
  uint32_t o = sidx[s];
  Ball a = raw[o];
  ball[s] = Ball{a.x, a.y, a.z, add_r(a.w, solvent)};
  perm[s] = o;
  cellk[s] = skey[s];
}
void launch_gather(const Ball* raw, const uint32_t* sidx,
                   const uint32_t* skey, int64_t n, Real solvent,
                   Ball* ball, uint32_t* perm, uint32_t* cellk,
                   PipeLaunch L) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_gather, dim3((unsigned)((n + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, L.s, raw, sidx, skey, n, solvent, ball,
                     perm, cellk);
}

/* cstart[c] = first sorted ball with key >= c, for c in [0, G]: one
 * thread per cell, binary search over the sorted keys (neighbouring cells
 * share the first probes, so the search runs from cache; a grid that is
 * mostly empty -- a chunk's sub-input on the whole input's grid -- costs no
 * serial fill of its empty runs) */
__global__ __launch_bounds__(kBlock) void k_cell_starts(const uint32_t* cellk,
                                                        int64_t n, int64_t G,
                                                        uint32_t* cstart) {
  const int64_t c = gtid();
  if (c > G) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)cellk[mid] < c)
      lo = mid + 1;
    else
      hi = mid;
  }
  cstart[c] = (uint32_t)lo;
}
void launch_cell_starts(const uint32_t* cellk, int64_t n, int64_t G,
                        uint32_t* cstart, PipeLaunch L) {
  hipLaunchKernelGGL(k_cell_starts,
                     dim3((unsigned)((G + 1 + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, L.s, cellk, n, G, cstart);
}

/* Fine scan grid: each reference cell split into kx cells along x.  One
 * thread per reference cell partitions its balls (in order) by x
 * sub-cell, (x - minx) * kx / step in double clamped into the cell: a ball
 * the FP32 key put in a neighbouring cell is off by less than the grid's
 * margin, which every scan radius includes. */
__global__ __launch_bounds__(kBlock) void k_fine_split(
    const Ball* ball, const uint32_t* cstart, int64_t G, int kx, GridP g1,
    const GridP* grids, const uint32_t* mol, Ball* fball, uint32_t* fid,
    uint32_t* fstart) {
  const int64_t c = gtid();
  if (c >= G) return;
  const uint32_t b0 = cstart[c], b1 = cstart[c + 1];
  uint32_t cnt[kMaxFine];
#pragma unroll
  for (int q = 0; q < kMaxFine; ++q) cnt[q] = 0;
  int base = 0;
  GridP g = g1;
  if (b1 > b0) {
    if (mol) g = grids[mol[b0]];
    /* g is the fine grid: its cell ids and x extent are kx times the
     * reference grid's */
    const int64_t local = c - (int64_t)(g.cbase / (uint32_t)kx);
    base = (int)(local % (g.ex / kx)) * kx; /* first fine x index of the cell */
    for (uint32_t i = b0; i < b1; ++i) {
      const double q = ((double)ball[i].x - (double)g.minx) * g.inv_step - base;
      const int s = q >= 1.0 ? min((int)fmin(q, 64.0), kx - 1) : 0;
#pragma unroll
      for (int t = 0; t < kMaxFine; ++t) cnt[t] += (t == s);
    }
  }
  uint32_t off[kMaxFine];
  uint32_t acc = b0;
#pragma unroll
  for (int t = 0; t < kMaxFine; ++t) {
    off[t] = acc;
    if (t < kx) fstart[c * kx + t] = acc;
    acc += cnt[t];
  }
  if (c == G - 1) fstart[G * kx] = b1;
  for (uint32_t i = b0; i < b1; ++i) {
    const double q = ((double)ball[i].x - (double)g.minx) * g.inv_step - base;
    const int s = q >= 1.0 ? min((int)fmin(q, 64.0), kx - 1) : 0;
    uint32_t at = 0;
#pragma unroll
    for (int t = 0; t < kMaxFine; ++t)
      if (t == s) at = off[t]++;
    fball[at] = ball[i];
    fid[at] = i;
  }
}
void launch_fine_split(const Ball* ball, const uint32_t* cstart, int64_t G,
                       int kx, GridP g, const GridP* grids,
                       const uint32_t* mol, Ball* fball, uint32_t* fid,
                       uint32_t* fstart, PipeLaunch L) {
  if (G <= 0) return;
  hipLaunchKernelGGL(k_fine_split, dim3((unsigned)((G + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, L.s, ball, cstart, G, kx, g, grids, mol,
                     fball, fid, fstart);
}

/* ------------------------------------------------------------------ */
/* filter decisions (one stage) and exact decisions                    */
/* ------------------------------------------------------------------ */
template <class T>
__device__ int edge_code(const Ball& a, const Ball& b, Real alpha) {
  T v = edge_val(Ld<T>::p(a), Ld<T>::w(a), Ld<T>::p(b), Ld<T>::w(b),
                 Al<T>::c(alpha));
  return code_of(fe_sign(v));
}
/* exact decisions over the wavefront-cooperative integers (every lane of
 * a group of kTPI lanes calls them with the same balls) */
using XI = CW<kTPI>;
__device__ int edge_exact(const Ball& a, const Ball& b,
                          const ExactScale& X) {
  if (same_pos(a, b)) return REJECT;
  const LdCW<kTPI> L{X.S};
  XI v = edge_val(L.p(a), L.w(a), L.p(b), L.w(b), cw_of<kTPI>(X.alpha));
  return fe_sign(v) >= 0 ? KEEP : REJECT;
}

template <class T>
__device__ int tri_code(const Ball& a, const Ball& b, const Ball& c,
                        Real alpha) {
  T nn;
  T v = tri_val(Ld<T>::p(a), Ld<T>::w(a), Ld<T>::p(b), Ld<T>::w(b),
                Ld<T>::p(c), Ld<T>::w(c), Al<T>::c(alpha), &nn);
  if (fe_sign(nn) <= 0) return UNDEC;
  return code_of(fe_sign(v));
}
__device__ int tri_exact(const Ball& a, const Ball& b, const Ball& c,
                         const ExactScale& X) {
  const LdCW<kTPI> L{X.S};
  XI nn;
  XI v = tri_val(L.p(a), L.w(a), L.p(b), L.w(b), L.p(c), L.w(c),
                 cw_of<kTPI>(X.alpha), &nn);
  if (fe_sign(nn) == 0) return REJECT;
  return fe_sign(v) >= 0 ? KEEP : REJECT;
}

template <class T>
__device__ int tet_code(const Ball& a, const Ball& b, const Ball& c,
                        const Ball& d, Real alpha) {
  T D;
  V3<T> N;
  T wa = Ld<T>::w(a);
  tet_DN(Ld<T>::p(a), wa, Ld<T>::p(b), Ld<T>::w(b), Ld<T>::p(c), Ld<T>::w(c),
         Ld<T>::p(d), Ld<T>::w(d), &D, &N);
  if (fe_sign(D) == 0) return UNDEC;
  return code_of(fe_sign(tet_val(D, N, wa, Al<T>::c(alpha))));
}
/* The first stages of the edge and triangle filters, semi-static as
 * tet_dn_ss: each value evaluated once in precision R together with its
 * permanent (the same expression on absolute values, subtractions made
 * additions), the error bound gamma_h times the permanent (ss_fe) for
 * the expression tree's height h, leaves u, v = p - p_a and w = r^2
 * rounded once, alpha exact, products by 2 and 4 exact.
 *   edge: V = 4 |u|^2 (w_a + alpha) - A^2, A = |u|^2 + w_a - w_b; h 8;
 *   triangle: |n|^2 (n = u x v), h 6, and V = 4 |n|^2 (w_a + alpha) -
 *   ((A^2 |v|^2 - 2 A B u.v) + B^2 |u|^2), A = |u|^2 + w_a - w_b,
 *   B = |v|^2 + w_a - w_c; h 11.
 * The same values as edge_val / tri_val; the bound replaces a running
 * error per operation (~4x the instructions for triangles). */
template <class R>
__device__ __forceinline__ int edge_code_ss(const Ball& a, const Ball& b,
                                            Real alpha) {
  const R ux = (R)b.x - (R)a.x, uy = (R)b.y - (R)a.y, uz = (R)b.z - (R)a.z;
  const R wa = (R)a.w * (R)a.w, wb = (R)b.w * (R)b.w, al = (R)alpha;
  const R uu = (ux * ux + uy * uy) + uz * uz;
  const R A = (uu + wa) - wb, Ap = (uu + wa) + wb;
  const R V = ((R)4 * uu) * (wa + al) - A * A;
  const R Vp = ((R)4 * uu) * (wa + fabs(al)) + Ap * Ap;
  return code_of(fe_sign(ss_fe(SS<R>{V, Vp}, 8)));
}
template <class R>
__device__ __forceinline__ int tri_code_ss(const Ball& a, const Ball& b,
                                           const Ball& c, Real alpha) {
  const R ax = a.x, ay = a.y, az = a.z;
  const R ux = (R)b.x - ax, uy = (R)b.y - ay, uz = (R)b.z - az;
  const R vx = (R)c.x - ax, vy = (R)c.y - ay, vz = (R)c.z - az;
  const R wa = (R)a.w * (R)a.w, wb = (R)b.w * (R)b.w, wc = (R)c.w * (R)c.w;
  SS<R> n[3];
  cross_ss(ux, uy, uz, vx, vy, vz, n);
  const R nn = (n[0].v * n[0].v + n[1].v * n[1].v) + n[2].v * n[2].v;
  const R nnp = (n[0].p * n[0].p + n[1].p * n[1].p) + n[2].p * n[2].p;
  if (fe_sign(ss_fe(SS<R>{nn, nnp}, 6)) <= 0) return UNDEC;
  const R uu = (ux * ux + uy * uy) + uz * uz;
  const R vv = (vx * vx + vy * vy) + vz * vz;
  const R uv = (ux * vx + uy * vy) + uz * vz;
  const R uvp = (fabs(ux) * fabs(vx) + fabs(uy) * fabs(vy)) + fabs(uz) * fabs(vz);
  const R A = (uu + wa) - wb, Ap = (uu + wa) + wb;
  const R B = (vv + wa) - wc, Bp = (vv + wa) + wc;
  const R lhs = ((A * A) * vv - (((R)2 * A) * B) * uv) + (B * B) * uu;
  const R lhsp = ((Ap * Ap) * vv + (((R)2 * Ap) * Bp) * uvp) + (Bp * Bp) * uu;
  const R al = (R)alpha;
  const R V = ((R)4 * nn) * (wa + al) - lhs;
  const R Vp = ((R)4 * nnp) * (wa + fabs(al)) + lhsp;
  return code_of(fe_sign(ss_fe(SS<R>{V, Vp}, 11)));
}
/* the first stage of the tet filter: D and N with a-priori error bounds
 * (tet_dn_ss: one extra operation per operation, on absolute values),
 * the final 4 D^2 (w_a + alpha) - |N|^2 with one bound (tet_val_ss) */
template <class R>
__device__ __forceinline__ int tet_code_ss(const Ball* v, Real alpha) {
  TetCert<R> T;
  tet_dn_ss(v, T);
  if (fe_sign(T.D) == 0) return UNDEC;
  return code_of(tet_val_ss(T, (R)alpha));
}
__device__ int tet_exact(const Ball& a, const Ball& b, const Ball& c,
                         const Ball& d, const ExactScale& X) {
  const LdCW<kTPI> L{X.S};
  XI D;
  V3<XI> N;
  XI wa = L.w(a);
  tet_DN(L.p(a), wa, L.p(b), L.w(b), L.p(c), L.w(c), L.p(d), L.w(d), &D, &N);
  if (fe_sign(D) == 0) return REJECT;
  return fe_sign(tet_val(D, N, wa, cw_of<kTPI>(X.alpha))) >= 0 ? KEEP : REJECT;
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
  const GridP g = grid_of(S, (uint32_t)i);
  uint32_t c = S.cellk[i];
  int gx, gy, gz;
  cell_xyz(c - g.cbase, g, gx, gy, gz);
  const int k = g.k_enum;
  uint32_t total = 0;
  for (int z = max(gz - k, 0); z <= min(gz + k, g.ez - 1); ++z)
    for (int y = max(gy - k, 0); y <= min(gy + k, g.ey - 1); ++y) {
      int64_t row = (int64_t)g.cbase + (int64_t)g.ex * (y + (int64_t)g.ey * z);
      int x0 = max(gx - k, 0), x1 = min(gx + k, g.ex - 1);
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

/* ------------------------------------------------------------------ */
/* triangles: k in row(i) after j, intersected with row(j)             */
/* ------------------------------------------------------------------ */
/* ------------------------------------------------------------------ */
/* tets: l in T(i,j) after k, T(i,k), T(j,k)                           */
/* ------------------------------------------------------------------ */
/* ------------------------------------------------------------------ */
/* enumeration: candidates as (segment, new vertex) pairs              */
/* ------------------------------------------------------------------ */
/* Slot capacities that need no merge: a triangle from alpha edge (i,j) has
 * its third vertex in row(i) after j, a tet from alpha triangle (i,j,k) has
 * its fourth vertex in T(i,j) after k.  Unused slots carry seg = ~0. */
__global__ __launch_bounds__(kBlock) void k_tail_caps(Cnt c_m,
                                                      const uint64_t* row,
                                                      const uint32_t* owner,
                                                      uint32_t* cap) {
  const int64_t m = c_m.get();
  int64_t q = gtid();
  if (q >= m) return;
  cap[q] = (uint32_t)(row[owner[q] + 1] - (uint64_t)q - 1);
}
/* chunk sub-input: no slots for the triangles (i,j,k) none of whose
 * vertices is owned -- every tet they start has its three lowest vertices
 * unowned, and no owned simplex depends on it */
__global__ __launch_bounds__(kBlock) void k_tail_caps_own(
    Stage S, Cnt c_m, const uint64_t* row, const uint32_t* owner,
    const uint32_t* eleft, const uint32_t* ecol, const uint32_t* tk,
    uint32_t* cap) {
  const int64_t m = c_m.get();
  int64_t q = gtid();
  if (q >= m) return;
  const uint32_t e = owner[q];
  const uint32_t i = eleft[e], j = ecol[e], k = tk[q];
  const bool need = (i >= S.own_lo && i < S.own_hi) ||
                    (j >= S.own_lo && j < S.own_hi) ||
                    (k >= S.own_lo && k < S.own_hi);
  cap[q] = need ? (uint32_t)(row[e + 1] - (uint64_t)q - 1) : 0u;
}
void launch_tail_caps(Cnt m, const uint64_t* row, const uint32_t* owner,
                      uint32_t* cap, PipeLaunch L) {
  if (m.cap <= 0) return;
  hipLaunchKernelGGL(k_tail_caps, dim3((unsigned)((m.cap + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, L.s, m, row, owner, cap);
}
void launch_tet_caps(const Stage& S, Cnt m, const uint64_t* trow,
                     const uint32_t* tedge, const uint32_t* eleft,
                     const uint32_t* ecol, const uint32_t* tk, uint32_t* cap,
                     PipeLaunch L) {
  if (m.cap <= 0) return;
  if (!S.chunk) {
    launch_tail_caps(m, trow, tedge, cap, L);
    return;
  }
  hipLaunchKernelGGL(k_tail_caps_own,
                     dim3((unsigned)((m.cap + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, L.s, S, m, trow, tedge, eleft, ecol, tk,
                     cap);
}

/* number of triangles of alpha edge q = (i,j): the size of row(i) after j
 * intersected with row(j) (computePossibleTriangleCount, Parallel:232-266) */
template <bool BUF>
__global__ __launch_bounds__(kBlock) void k_tri_count(Cnt c_ne,
                                                     const uint64_t* erow,
                                                     const uint32_t* ecol,
                                                     const uint32_t* eleft,
                                                     uint32_t* cnt) {
  const int64_t ne = c_ne.get();
  int64_t q = gtid();
  if (q >= ne) return;
  uint32_t i = eleft[q], j = ecol[q];
  uint32_t c = 0;
  if constexpr (BUF) {
    /* byte offsets into ecol (below 4 GiB): no 64-bit address
     * arithmetic per load */
    const BufU32 E{ecol, (uint32_t)(c_ne.cap * 4)};
    uint32_t a = ((uint32_t)q + 1u) << 2, ae = (uint32_t)erow[i + 1] << 2;
    uint32_t b = (uint32_t)erow[j] << 2, be = (uint32_t)erow[j + 1] << 2;
    /* two merge steps a round: each list's next element is loaded with
     * its current one (~0u past the row's end, above every ball id) */
    while (a < ae && b < be) {
      const uint32_t x0 = E.at(a), y0 = E.at(b);
      const uint32_t x1 = a + 4u < ae ? E.at(a + 4u) : ~0u;
      const uint32_t y1 = b + 4u < be ? E.at(b + 4u) : ~0u;
      const bool sa = x0 <= y0, sb = y0 <= x0;
      const uint32_t x = sa ? x1 : x0, y = sb ? y1 : y0;
      c += (x0 == y0) + (x == y && x != ~0u);
      a += (sa ? 4u : 0u) + (x <= y ? 4u : 0u);
      b += (sb ? 4u : 0u) + (y <= x ? 4u : 0u);
    }
  } else {
    /* alpha-edge ids fit 32 bits (the candidate count is checked below
     * 2^32), and 32-bit cursors halve the merge's index arithmetic */
    uint32_t a = (uint32_t)q + 1, ae = (uint32_t)erow[i + 1];
    uint32_t b = (uint32_t)erow[j], be = (uint32_t)erow[j + 1];
    while (a < ae && b < be) {
      const uint32_t x = ecol[a], y = ecol[b];
      c += (x == y);
      a += (x <= y);
      b += (y <= x);
    }
  }
  cnt[q] = c;
}
void launch_tri_count(Cnt ne, const uint64_t* erow, const uint32_t* ecol,
                      const uint32_t* eleft, uint32_t* cnt, PipeLaunch L) {
  if (ne.cap <= 0) return;
  auto kf = ne.cap < ((int64_t)1 << 30) ? k_tri_count<true> : k_tri_count<false>;
  hipLaunchKernelGGL(kf, dim3((unsigned)((ne.cap + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, L.s, ne, erow, ecol, eleft, cnt);
}

/* ------------------------------------------------------------------ */
/* alpha filters, one thread per candidate (uniform work per lane)     */
/* ------------------------------------------------------------------ */
/* vertices of candidate (seg, val): KIND 1 edge (seg = ball), 2 triangle
 * (seg = alpha edge), 3 tet (seg = alpha triangle) */
template <int KIND>
__device__ __forceinline__ void cand_balls(const Stage& S, const CandCtx& C,
                                           int64_t o, uint32_t seg,
                                           uint32_t val, Ball* v) {
  if (KIND == 3) {
    const uint4 t = C.ctet[o];
    v[0] = S.ball[t.x];
    v[1] = S.ball[t.y];
    v[2] = S.ball[t.z];
    v[3] = S.ball[t.w];
  } else if (KIND == 1) {
    v[0] = S.ball[seg];
    v[1] = S.ball[val];
  } else if (KIND == 2) {
    v[0] = S.ball[C.eleft[seg]];
    v[1] = S.ball[C.ecol[seg]];
    v[2] = S.ball[val];
  }
}
template <int KIND, class T>
__device__ __forceinline__ int filter_code(const Ball* v, Real alpha) {
  if (KIND == 1) return same_pos(v[0], v[1]) ? REJECT : edge_code<T>(v[0], v[1], alpha);
  if (KIND == 2) return tri_code<T>(v[0], v[1], v[2], alpha);
  return tet_code<T>(v[0], v[1], v[2], v[3], alpha);
}
/* Degenerate candidates (collinear triangle centres: n = u x v = 0;
 * coplanar tet centres: D = u . (v x t) = 0), which the filter rejects,
 * decided in 128-bit integers on the coordinates scaled by 2^S when their
 * differences fit 40 bits (n's components stay below 2^81, D below 2^123):
 * 1 degenerate, 0 not, -1 out of that range.  Lattice inputs send millions
 * of such candidates past the float stages (D or |n|^2 exactly 0). */
template <int KIND>
__device__ __forceinline__ int degenerate_i128(const Ball* v, int S) {
  constexpr int NB = KIND == 3 ? 4 : 3;
  int64_t d[NB - 1][3];
  double o[3] = {(double)v[0].x, (double)v[0].y, (double)v[0].z};
  int64_t q0[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double sv = __builtin_ldexp(o[a], S);
    if (!(fabs(sv) < 0x1p62)) return -1;
    q0[a] = (int64_t)sv;
  }
#pragma unroll
  for (int i = 1; i < NB; ++i) {
    const double c[3] = {(double)v[i].x, (double)v[i].y, (double)v[i].z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double sv = __builtin_ldexp(c[a], S);
      if (!(fabs(sv) < 0x1p62)) return -1;
      d[i - 1][a] = (int64_t)sv - q0[a];
      if (d[i - 1][a] >= ((int64_t)1 << 40) || d[i - 1][a] <= -((int64_t)1 << 40))
        return -1;
    }
  }
  const int64_t *u = d[0], *w = d[1];
  const __int128 nx = (__int128)u[1] * w[2] - (__int128)u[2] * w[1];
  const __int128 ny = (__int128)u[2] * w[0] - (__int128)u[0] * w[2];
  const __int128 nz = (__int128)u[0] * w[1] - (__int128)u[1] * w[0];
  if (KIND == 2) return nx == 0 && ny == 0 && nz == 0;
  /* D = t . (u x w), the same determinant as u . (w x t) */
  const int64_t* t = d[NB - 2];
  const __int128 D = (__int128)t[0] * nx + (__int128)t[1] * ny +
                     (__int128)t[2] * nz;
  return D == 0;
}

template <int KIND>
__device__ __forceinline__ int filter_exact(const Ball* v,
                                            const ExactScale& X) {
  if (KIND == 1) return edge_exact(v[0], v[1], X);
  if (KIND == 2) return tri_exact(v[0], v[1], v[2], X);
  return tet_exact(v[0], v[1], v[2], v[3], X);
}
constexpr int fam_of(int KIND) {
  return KIND == 1 ? FAM_EDGE : (KIND == 2 ? FAM_TRI : FAM_TET);
}

/* Tets of alpha triangle q = (i,j,k): l in T(i,j) after k, T(i,k), T(j,k)
 * (computePossibleTetCount / fillPossibleTets, Parallel:418-543), fused
 * with the tet filter's first stage (markRealTets, :531-591).  A block's
 * candidate slots are one contiguous range, staged in LDS as (i, j, k, l)
 * during the three-list merges; the block then filters the staged slots
 * round robin (uniform work per lane), writes every slot's flag coalesced
 * and the four vertex ids only of the slots that are kept or undecided
 * (the later stages read no others).  Slots past the LDS stage were
 * written to global memory during the merges and are filtered from there. */
/* First slot of alpha triangle q without a per-triangle offset array:
 * the triangles of edge e occupy rows trow[e] .. trow[e+1]-1 and each has
 * its row's tail after it as slots, so triangle s + k of a row of L = trow
 * [e+1] - s triangles starts k (L - 1) - k (k - 1) / 2 slots into the
 * edge's range, which begins at roff[e] (the exclusive scan over edges of
 * L (L - 1) / 2). */
__device__ __forceinline__ uint64_t tet_slot_off(const uint64_t* roff,
                                                 const uint64_t* trow,
                                                 uint32_t e, uint32_t q) {
  const uint64_t s = trow[e], L = trow[e + 1] - s, k = (uint64_t)q - s;
  return roff[e] + k * (L - 1) - k * (k - 1) / 2;
}
/* FP32 first stage: 8 waves per SIMD (at most 64 VGPRs, no spills; the
 * LDS stage admits 8 blocks per CU); FP64: the compiler's choice */
template <class R, bool BUF>
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 8 : 1))) void k_tet_enum_filter(
    Stage S, Cnt c_nt, const uint64_t* trow, const uint32_t* tk,
    const uint32_t* tedge, const uint32_t* teik, const uint32_t* tejk,
    const uint32_t* eleft, const uint32_t* ecol, const uint64_t* off,
    const uint64_t* roff, uint64_t* bbase, uint4* ctet, uint8_t* flag,
    uint32_t* bcnt, uint32_t* bkeep, Queue q1, uint64_t lim) {
  const int64_t nt = c_nt.get();
  /* 8 blocks per CU fit in LDS with block_flush's 16 B (the kernel's 60
   * VGPRs allow 8 waves per SIMD); a block of 256 alpha triangles averages
   * ~600 slots */
  constexpr int CAP = 1272;
  __shared__ uint4 s_out[CAP];
  const int64_t q0 = (int64_t)blockIdx.x * kBlock;
  if (q0 >= nt) { /* whole block past the (device) count */
    if (threadIdx.x == 0) {
      bcnt[blockIdx.x] = bkeep[blockIdx.x] = 0;
      bbase[blockIdx.x] = 0;
    }
    return;
  }
  __shared__ unsigned s_und; /* an undecided candidate written */
  if (threadIdx.x == 0) s_und = 0;
  const int64_t q1e = q0 + kBlock < nt ? q0 + kBlock : nt;
  /* the triangle's slots [o, oe): off[] per triangle, or (roff) the
   * closed form, a triangle's slots being its row's tail; the block's
   * range runs from its first triangle's o to its last one's oe */
  const int64_t q = gtid();
  uint32_t eij = 0;
  uint64_t o = 0, oe = 0;
  if (q < nt) {
    eij = tedge[q];
    if (off) {
      o = off[q];
      oe = off[q + 1];
    } else {
      o = tet_slot_off(roff, trow, eij, (uint32_t)q);
      oe = o + (trow[eij + 1] - (uint64_t)q - 1);
    }
  }
  __shared__ uint64_t s_span[2];
  if (q == q0) s_span[0] = o;
  if (q == q1e - 1) s_span[1] = oe;
  __syncthreads();
  if (threadIdx.x == 0) bbase[blockIdx.x] = s_span[0];
  BlockOut4<CAP> B{s_out, s_span[0], s_span[1] - s_span[0], lim};
  unsigned found = 0;
  if (q < nt) {
    const uint32_t eik = teik[q], ejk = tejk[q];
    const uint32_t i = eleft[eij], j = ecol[eij], k = tk[q];
    if constexpr (BUF) {
      /* the same merge over byte offsets into tk (below 4 GiB): no 64-bit
       * address arithmetic per load.  A triangle's slots are its row's
       * tail after it, so o < oe holds while a < ae, unless it was given
       * none (chunk sub-inputs): then the merge is empty. */
      const BufU32 K{tk, (uint32_t)(c_nt.cap * 4)};
      uint32_t a = ((uint32_t)q + 1u) << 2, ae = (uint32_t)trow[eij + 1] << 2;
      uint32_t b = (uint32_t)trow[eik] << 2, be = (uint32_t)trow[eik + 1] << 2;
      uint32_t c = (uint32_t)trow[ejk] << 2, ce = (uint32_t)trow[ejk + 1] << 2;
      if (o == oe) ae = a;
      /* two merge steps a round: each list's next element is loaded with
       * its current one (~0u past the row's end, above every ball id) */
      while (a < ae && b < be && c < ce) {
        const uint32_t x0 = K.at(a), y0 = K.at(b), z0 = K.at(c);
        const uint32_t x1 = a + 4u < ae ? K.at(a + 4u) : ~0u;
        const uint32_t y1 = b + 4u < be ? K.at(b + 4u) : ~0u;
        const uint32_t z1 = c + 4u < ce ? K.at(c + 4u) : ~0u;
        const uint32_t m0 = max(x0, max(y0, z0));
        const bool all0 = x0 == m0 && y0 == m0 && z0 == m0;
        if (all0) {
          B.put(ctet, o, make_uint4(i, j, k, m0));
          ++o;
          ++found;
        }
        const bool sa = all0 || x0 < m0, sb = all0 || y0 < m0,
                   sc = all0 || z0 < m0;
        const uint32_t x = sa ? x1 : x0, y = sb ? y1 : y0, z = sc ? z1 : z0;
        const uint32_t m = max(x, max(y, z));
        const bool all = x == m && y == m && z == m;
        if (all && m != ~0u) {
          B.put(ctet, o, make_uint4(i, j, k, m));
          ++o;
          ++found;
        }
        a += (sa ? 4u : 0u) + ((all || x < m) ? 4u : 0u);
        b += (sb ? 4u : 0u) + ((all || y < m) ? 4u : 0u);
        c += (sc ? 4u : 0u) + ((all || z < m) ? 4u : 0u);
      }
    } else {
      /* alpha-triangle ids fit 32 bits (the slot count is checked below
       * 2^32): 32-bit cursors halve the merge's index arithmetic; the
       * slots end with T(i,j) (o < oe adds nothing then) unless the
       * triangle was given none (chunk sub-inputs) */
      uint32_t a = (uint32_t)q + 1, ae = (uint32_t)trow[eij + 1];
      uint32_t b = (uint32_t)trow[eik], be = (uint32_t)trow[eik + 1];
      uint32_t c = (uint32_t)trow[ejk], ce = (uint32_t)trow[ejk + 1];
      while (a < ae && b < be && c < ce && o < oe) {
        const uint32_t x = tk[a], y = tk[b], z = tk[c];
        const uint32_t m = max(x, max(y, z));
        if (x == m && y == m && z == m) {
          B.put(ctet, o, make_uint4(i, j, k, m));
          ++o;
          ++found;
          ++a;
          ++b;
          ++c;
        } else {
          a += (x < m);
          b += (y < m);
          c += (z < m);
        }
      }
    }
    for (; o < oe; ++o) B.put(ctet, o, make_uint4(0xffffffffu, 0u, 0u, 0u));
  }
  __syncthreads(); /* LDS stage and the overflow slots written */
  /* First precision stage over the block's slots, 256 per round; the kept
   * and undecided ones are written compacted, in slot order, from the
   * block's first slot on: position base + (kept before them) <= their
   * own slot, and every slot of this round and the earlier ones was read
   * before the round's first barrier, so no unread overflow slot is
   * overwritten. */
  __shared__ unsigned s_wsum[kBlock / 64];
  const uint64_t base = B.base;
  const uint64_t m = base >= lim ? 0 : (B.total < lim - base ? B.total : lim - base);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  uint32_t run = 0; /* kept + undecided written so far (block-uniform) */
  for (uint64_t r0 = 0; r0 < m; r0 += kBlock) {
    const uint64_t t = r0 + threadIdx.x;
    uint4 v = make_uint4(0xffffffffu, 0u, 0u, 0u);
    int r = REJECT;
    if (t < m) {
      v = t < (uint64_t)CAP ? s_out[t] : ctet[base + t];
      if (v.x != 0xffffffffu) {
        const Ball bv[4] = {S.ball[v.x], S.ball[v.y], S.ball[v.z], S.ball[v.w]};
        /* the weights are first used after the test of D's sign; left
         * alone, the compiler loads them there, a second round trip per
         * round; used here (an empty asm), each ball is one 16-byte load */
#pragma unroll
        for (int k = 0; k < 4; ++k) asm volatile("" ::"v"(bv[k].w));
        r = tet_code_ss<R>(bv, S.alpha);
      }
    }
    const uint64_t bm = __ballot(r != REJECT);
    if (lane == 0) s_wsum[w] = (unsigned)__popcll(bm);
    __syncthreads();
    uint32_t pre = run + (uint32_t)__popcll(bm & lt), tot = 0;
#pragma unroll
    for (int q = 0; q < kBlock / 64; ++q) {
      const unsigned c = s_wsum[q];
      if (q < w) pre += c;
      tot += c;
    }
    if (r != REJECT) {
      const uint64_t at = base + pre;
      ctet[at] = v;
      flag[at] = (uint8_t)r;
      if (r == UNDEC) {
        push_q(q1, (uint32_t)at, v.x);
        s_und = 1;
      }
    }
    run += tot;
    __syncthreads(); /* s_wsum is rewritten next round */
  }
  if (threadIdx.x == 0) {
    bcnt[blockIdx.x] = run;
    bkeep[blockIdx.x] = s_und ? kRecount : run;
  }
  block_flush(S.ctr, CTR_CAND + 2, found);
}
void launch_tet_enum_filter(const Stage& S, Cnt nt, const uint64_t* trow,
                            const uint32_t* tk, const uint32_t* tedge,
                            const uint32_t* teik, const uint32_t* tejk,
                            const uint32_t* eleft, const uint32_t* ecol,
                            const uint64_t* off, const uint64_t* roff,
                            uint64_t* bbase, uint4* ctet, uint8_t* flag,
                            uint32_t* bcnt, uint32_t* bkeep, Queue q1,
                            PipeLaunch L, uint64_t lim) {
  if (nt.cap <= 0) return;
  dim3 gr((unsigned)((nt.cap + kBlock - 1) / kBlock));
  /* byte offsets into tk while it stays below 4 GiB */
  const bool buf = nt.cap < ((int64_t)1 << 30);
  if constexpr (!kBallF64) {
    if (S.staged) {
      auto kf = buf ? k_tet_enum_filter<float, true>
                    : k_tet_enum_filter<float, false>;
      hipLaunchKernelGGL(kf, gr, dim3(kBlock), 0, L.s, S, nt, trow, tk, tedge,
                         teik, tejk, eleft, ecol, off, roff, bbase, ctet, flag,
                         bcnt, bkeep, q1, lim);
      return;
    }
  }
  auto kd = buf ? k_tet_enum_filter<double, true>
                : k_tet_enum_filter<double, false>;
  hipLaunchKernelGGL(kd, gr, dim3(kBlock), 0, L.s, S, nt, trow, tk, tedge, teik,
                     tejk, eleft, ecol, off, roff, bbase, ctet, flag, bcnt,
                     bkeep, q1, lim);
}

/* After the ladder: the kept tets of each enumeration block (its first
 * bcnt[b] positions from off[256 b] hold its kept and formerly undecided
 * tets), one wave per block. */
__global__ __launch_bounds__(kBlock) void k_tet_block_keep(
    Cnt c_nt, const uint64_t* off, int stride, const uint32_t* bcnt,
    const uint8_t* flag, uint32_t* bkeep) {
  const int64_t nt = c_nt.get();
  const int64_t nblk = (c_nt.cap + kBlock - 1) / kBlock;
  const int64_t b = gtid() >> 6;
  const int lane = threadIdx.x & 63;
  if (b >= nblk) return;
  /* a block that queued no undecided candidate kept all it wrote: its
   * enumeration left that count here */
  if (bkeep[b] != kRecount) return;
  unsigned c = 0;
  if (b * kBlock < nt) {
    const uint64_t base = off[b * stride];
    const uint32_t n = bcnt[b];
    for (uint32_t t = lane; t < n; t += 64) c += flag[base + t] == KEEP;
  }
  c = wave_sum(c);
  if (lane == 0) bkeep[b] = c;
}
/* the kept tets of block b to tets[bout[b], ...), in order; one wave per
 * block (a block holds ~100 entries: with 256-thread workgroups the copy
 * waits on four times fewer blocks in flight, each a chain of dependent
 * loads).  Grid as k_tet_block_keep's. */
__global__ __launch_bounds__(kBlock) void k_tet_block_copy(
    Cnt c_nt, const uint64_t* off, int stride, const uint32_t* bcnt,
    const uint4* ctet, const uint8_t* flag, const uint64_t* bout, uint4* tets,
    uint64_t ocap) {
  const int64_t nt = c_nt.get();
  const int64_t b = gtid() >> 6;
  const int lane = threadIdx.x & 63;
  if (b * kBlock >= nt) return; /* wave-uniform */
  const uint64_t base = off[b * stride];
  const uint32_t n = bcnt[b];
  uint64_t o = bout[b];
  if (bout[b + 1] - o == n) { /* every entry kept: a straight copy */
    for (uint32_t t = lane; t < n; t += 64)
      if (o + t < ocap) tets[o + t] = ctet[base + t];
    return;
  }
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  for (uint32_t r0 = 0; r0 < n; r0 += 64) {
    const uint32_t t = r0 + lane;
    bool keep = false;
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if (t < n) { /* both loads in flight together */
      v = ctet[base + t];
      keep = flag[base + t] == KEEP;
    }
    const uint64_t bm = __ballot(keep);
    const uint64_t pre = o + (uint64_t)__popcll(bm & lt);
    if (keep && pre < ocap) tets[pre] = v;
    o += (uint64_t)__popcll(bm);
  }
}
void launch_tet_block_keep(Cnt nt, const uint64_t* off, int stride,
                           const uint32_t* bcnt, const uint8_t* flag,
                           uint32_t* bkeep, PipeLaunch L) {
  if (nt.cap <= 0) return;
  const int64_t nblk = (nt.cap + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_tet_block_keep,
                     dim3((unsigned)((nblk * 64 + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, L.s, nt, off, stride, bcnt, flag, bkeep);
}
void launch_tet_block_copy(Cnt nt, const uint64_t* off, int stride,
                           const uint32_t* bcnt, const uint4* ctet,
                           const uint8_t* flag, const uint64_t* bout,
                           uint4* tets, PipeLaunch L, uint64_t ocap) {
  if (nt.cap <= 0) return;
  const int64_t nblk = (nt.cap + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_tet_block_copy,
                     dim3((unsigned)((nblk * 64 + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, L.s, nt, off, stride, bcnt, ctet, flag,
                     bout, tets, ocap);
}

/* Edge candidates of ball i (j > i in the stencil rows, Parallel:122-201)
 * fused with the edge filter's first stage (markRealEdges,
 * Parallel:1229-1250).  One wave per ball: lane r finds stencil row r's
 * candidate range (64 rows per pass), a wave prefix sum lays the ranges
 * end to end, and the lanes test the concatenation 64 candidates at a
 * time (a row of ~15 candidates would leave most lanes idle); the kept
 * and undecided ones are written compacted, in order, from the ball's
 * first slot on (a ballot prefix per 64), with their flags; nothing is
 * written for the others. */
constexpr int kEdgeBallsPerBlock = kBlock / 64;
template <class R>
__global__ __launch_bounds__(kBlock) void k_edge_enum_filter(
    Stage S, const uint64_t* off, uint32_t* cval, uint32_t* cseg,
    uint8_t* flag, uint32_t* bcnt, Queue q1, uint64_t lim) {
  const int64_t iw = (int64_t)blockIdx.x * kEdgeBallsPerBlock + (threadIdx.x >> 6);
  if (iw >= S.n) return; /* a whole wave */
  const uint32_t i = __builtin_amdgcn_readfirstlane((uint32_t)iw);
  const int lane = threadIdx.x & 63;
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  const GridP g = grid_of(S, i);
  const uint32_t c = S.cellk[i];
  int gx, gy, gz;
  cell_xyz(c - g.cbase, g, gx, gy, gz);
  const int k = g.k_enum;
  const int z0 = max(gz - k, 0), y0 = max(gy - k, 0);
  const int nz = min(gz + k, g.ez - 1) - z0 + 1, ny = min(gy + k, g.ey - 1) - y0 + 1;
  const int x0 = max(gx - k, 0), x1 = min(gx + k, g.ex - 1);
  const Ball bi = S.ball[i];
  const uint64_t o0 = off[i];
  uint64_t o = o0; /* next compacted position */
  __shared__ uint32_t s_end[kEdgeBallsPerBlock][64], s_beg[kEdgeBallsPerBlock][64];
  const int wv = threadIdx.x >> 6;
  for (int r0 = 0; r0 < ny * nz; r0 += 64) {
    /* this lane's row: candidates [b, b + len) */
    uint32_t b = 0, len = 0;
    const int r = r0 + lane;
    if (r < ny * nz) {
      const int y = y0 + r % ny, z = z0 + r / ny;
      const int64_t row = (int64_t)g.cbase + (int64_t)g.ex * (y + (int64_t)g.ey * z);
      const int64_t c0 = row + x0, c1 = row + x1;
      if (c1 >= (int64_t)c) {
        b = S.cstart[c0];
        const uint32_t e = S.cstart[c1 + 1];
        if (b < i + 1) b = i + 1;
        len = e > b ? e - b : 0u;
      }
    }
    /* inclusive prefix of the lengths: row r's candidates are positions
     * [incl - len, incl) of the pass; the rows' ends and starts go to LDS
     * for the lanes' searches */
    uint32_t incl = len;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t y = (uint32_t)__shfl_up((int)incl, d, 64);
      if (lane >= d) incl += y;
    }
    const uint32_t total = (uint32_t)__shfl((int)incl, 63, 64);
    s_end[wv][lane] = incl;
    s_beg[wv][lane] = b - (incl - len); /* j = s_beg[row] + f */
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    int rr = 0; /* the row of this lane's candidate: first with end > f */
    for (uint32_t f0 = 0; f0 < total; f0 += 64) {
      const uint32_t f = f0 + (uint32_t)lane;
      int code = REJECT;
      uint32_t j = 0;
      if (f < total) {
        while (s_end[wv][rr] <= f) ++rr; /* f < total = s_end[63] */
        j = s_beg[wv][rr] + f;
        const Ball bj = S.ball[j];
        code = same_pos(bi, bj) ? REJECT : edge_code_ss<R>(bi, bj, S.alpha);
      }
      const uint64_t m = __ballot(code != REJECT);
      if (code != REJECT) {
        const uint64_t w = o + (uint64_t)__popcll(m & lt);
        if (w < lim) {
          cval[w] = j;
          cseg[w] = i;
          flag[w] = (uint8_t)code;
          if (code == UNDEC) push_q(q1, (uint32_t)w, i);
        }
      }
      o += (uint64_t)__popcll(m);
    }
    /* the next pass rewrites the tables */
    __builtin_amdgcn_wave_barrier();
  }
  if (lane == 0) bcnt[i] = (uint32_t)(o - o0);
}
void launch_edge_enum_filter(const Stage& S, const uint64_t* off,
                             uint32_t* cval, uint32_t* cseg, uint8_t* flag,
                             uint32_t* bcnt, Queue q1, PipeLaunch L,
                             uint64_t lim) {
  if (S.n <= 0) return;
  dim3 gr((unsigned)((S.n + kEdgeBallsPerBlock - 1) / kEdgeBallsPerBlock));
  if (S.staged)
    hipLaunchKernelGGL(k_edge_enum_filter<float>, gr, dim3(kBlock), 0, L.s, S,
                       off, cval, cseg, flag, bcnt, q1, lim);
  else
    hipLaunchKernelGGL(k_edge_enum_filter<double>, gr, dim3(kBlock), 0, L.s, S,
                       off, cval, cseg, flag, bcnt, q1, lim);
}
/* kept candidates of segment i among its bcnt[i] compacted ones */
__global__ __launch_bounds__(kBlock) void k_seg_keep(int64_t n,
                                                     const uint64_t* off,
                                                     const uint32_t* bcnt,
                                                     const uint8_t* flag,
                                                     uint32_t* keep) {
  const int64_t i = gtid();
  if (i >= n) return;
  const uint64_t base = off[i];
  const uint32_t m = bcnt[i];
  uint32_t c = 0;
  for (uint32_t t = 0; t < m; ++t) c += flag[base + t] == KEEP;
  keep[i] = c;
}
void launch_seg_keep(int64_t n, const uint64_t* off, const uint32_t* bcnt,
                     const uint8_t* flag, uint32_t* keep, PipeLaunch L) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_seg_keep, dim3((unsigned)((n + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, L.s, n, off, bcnt, flag, keep);
}
/* ball i's kept edges (one wave per ball) to (ecol, eleft)[erow[i], ...) */
__global__ __launch_bounds__(kBlock) void k_edge_copy(
    int64_t n, const uint64_t* off, const uint32_t* bcnt, const uint32_t* cval,
    const uint8_t* flag, const uint64_t* erow, uint32_t* ecol, uint32_t* eleft,
    uint64_t ocap) {
  const int64_t iw = (int64_t)blockIdx.x * kEdgeBallsPerBlock + (threadIdx.x >> 6);
  if (iw >= n) return;
  const uint32_t i = __builtin_amdgcn_readfirstlane((uint32_t)iw);
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  const uint64_t base = off[i];
  const uint32_t m = bcnt[i];
  uint64_t o = erow[i];
  for (uint32_t t0 = 0; t0 < m; t0 += 64) {
    const uint32_t t = t0 + lane;
    const bool keep = t < m && flag[base + t] == KEEP;
    const uint64_t bm = __ballot(keep);
    if (keep) {
      const uint64_t w = o + (uint64_t)__popcll(bm & lt);
      if (w < ocap) {
        ecol[w] = cval[base + t];
        eleft[w] = i;
      }
    }
    o += (uint64_t)__popcll(bm);
  }
}
void launch_edge_copy(int64_t n, const uint64_t* off, const uint32_t* bcnt,
                      const uint32_t* cval, const uint8_t* flag,
                      const uint64_t* erow, uint32_t* ecol, uint32_t* eleft,
                      PipeLaunch L, uint64_t ocap) {
  if (n <= 0) return;
  hipLaunchKernelGGL(
      k_edge_copy,
      dim3((unsigned)((n + kEdgeBallsPerBlock - 1) / kEdgeBallsPerBlock)),
      dim3(kBlock), 0, L.s, n, off, bcnt, cval, flag, erow, ecol, eleft, ocap);
}

/* Triangles of alpha edge q = (i, j) -- k in row(i) after j and in row(j),
 * the positions of k in the two rows being the alpha-edge ids of (i, k)
 * and (j, k), recorded for the tet stage and the closure -- fused with the
 * triangle filter's first stage (markRealTriangles, Parallel:1294-1340):
 * the block's candidate slots are staged in LDS during the row merges,
 * filtered round robin (uniform work per lane) and the kept and undecided
 * ones written compacted, in slot order, from the block's first slot on
 * (as in k_tet_enum_filter), four arrays and the flag; the filter never
 * reads an array the enumeration wrote for every slot. */
template <class R, bool BUF>
__global__ __launch_bounds__(kBlock) void k_tri_enum_filter(
    Stage S, Cnt c_ne, const uint64_t* erow, const uint32_t* ecol,
    const uint32_t* eleft, const uint64_t* off, uint32_t* cval,
    uint32_t* cseg, uint32_t* cea, uint32_t* ceb, uint8_t* flag,
    uint32_t* bcnt, uint32_t* bkeep, Queue q1, uint64_t lim) {
  const int64_t ne = c_ne.get();
  /* 6 blocks per CU fit in LDS with the edge table and block_flush's 16 B
   * (a block averages ~1,800 slots on the 100k globule; the rest go to
   * global memory: measured 6 > 5 > 8 blocks per CU) */
  constexpr int CAP = 1570;
  __shared__ uint32_t s_out[4 * CAP];
  __shared__ uint2 s_ij[kBlock]; /* the block's edges (i, j) */
  const int64_t q0 = (int64_t)blockIdx.x * kBlock;
  if (q0 >= ne) { /* whole block past the (device) count */
    if (threadIdx.x == 0) bcnt[blockIdx.x] = bkeep[blockIdx.x] = 0;
    return;
  }
  __shared__ unsigned s_und; /* an undecided candidate written */
  if (threadIdx.x == 0) s_und = 0;
  const int64_t q1e = q0 + kBlock < ne ? q0 + kBlock : ne;
  BlockOut<4, CAP> B{s_out, off[q0], off[q1e] - off[q0], lim};
  uint32_t* const g[4] = {cval, cseg, cea, ceb};
  const int64_t q = gtid();
  unsigned found = 0;
  if (q < ne) {
    const uint32_t i = eleft[q], j = ecol[q];
    s_ij[threadIdx.x] = make_uint2(i, j);
    uint64_t o = off[q];
    if constexpr (BUF) { /* byte offsets into ecol, as in k_tri_count */
      const BufU32 E{ecol, (uint32_t)(c_ne.cap * 4)};
      uint32_t a = ((uint32_t)q + 1u) << 2, ae = (uint32_t)erow[i + 1] << 2;
      uint32_t b = (uint32_t)erow[j] << 2, be = (uint32_t)erow[j + 1] << 2;
      /* two merge steps a round, as in k_tri_count */
      while (a < ae && b < be) {
        const uint32_t x0 = E.at(a), y0 = E.at(b);
        const uint32_t x1 = a + 4u < ae ? E.at(a + 4u) : ~0u;
        const uint32_t y1 = b + 4u < be ? E.at(b + 4u) : ~0u;
        if (x0 == y0) {
          const uint32_t v[4] = {x0, (uint32_t)q, a >> 2, b >> 2};
          B.put(g, o, v);
          ++o;
          ++found;
        }
        const bool sa = x0 <= y0, sb = y0 <= x0;
        a += sa ? 4u : 0u;
        b += sb ? 4u : 0u;
        const uint32_t x = sa ? x1 : x0, y = sb ? y1 : y0;
        if (x == y && x != ~0u) {
          const uint32_t v[4] = {x, (uint32_t)q, a >> 2, b >> 2};
          B.put(g, o, v);
          ++o;
          ++found;
        }
        a += x <= y ? 4u : 0u;
        b += y <= x ? 4u : 0u;
      }
    } else {
      uint32_t a = (uint32_t)q + 1, ae = (uint32_t)erow[i + 1];
      uint32_t b = (uint32_t)erow[j], be = (uint32_t)erow[j + 1];
      while (a < ae && b < be) {
        const uint32_t x = ecol[a], y = ecol[b];
        if (x == y) {
          const uint32_t v[4] = {x, (uint32_t)q, (uint32_t)a, (uint32_t)b};
          B.put(g, o, v);
          ++o;
          ++found;
        }
        a += (x <= y);
        b += (y <= x);
      }
    }
  }
  __syncthreads(); /* LDS stage, edge table and overflow slots written */
  __shared__ unsigned s_wsum[kBlock / 64];
  const uint64_t base = B.base;
  const uint64_t m = base >= lim ? 0 : (B.total < lim - base ? B.total : lim - base);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  uint32_t run = 0; /* kept + undecided written so far (block-uniform) */
  for (uint64_t r0 = 0; r0 < m; r0 += kBlock) {
    const uint64_t t = r0 + threadIdx.x;
    uint32_t v[4] = {0u, 0u, 0u, 0u};
    int r = REJECT;
    if (t < m) {
#pragma unroll
      for (int k = 0; k < 4; ++k)
        v[k] = t < (uint64_t)CAP ? s_out[k * CAP + t] : g[k][base + t];
      const uint2 ij = s_ij[v[1] - (uint32_t)q0];
      const Ball b0 = S.ball[ij.x], b1 = S.ball[ij.y], b2 = S.ball[v[0]];
      /* weights used here: one 16-byte load per ball (as for tets) */
      asm volatile("" ::"v"(b0.w), "v"(b1.w), "v"(b2.w));
      r = tri_code_ss<R>(b0, b1, b2, S.alpha);
    }
    const uint64_t bm = __ballot(r != REJECT);
    if (lane == 0) s_wsum[w] = (unsigned)__popcll(bm);
    __syncthreads();
    uint32_t pre = run + (uint32_t)__popcll(bm & lt), tot = 0;
#pragma unroll
    for (int k = 0; k < kBlock / 64; ++k) {
      const unsigned c = s_wsum[k];
      if (k < w) pre += c;
      tot += c;
    }
    if (r != REJECT) {
      const uint64_t at = base + pre;
#pragma unroll
      for (int k = 0; k < 4; ++k) g[k][at] = v[k];
      flag[at] = (uint8_t)r;
      if (r == UNDEC) {
        push_q(q1, (uint32_t)at, v[1]);
        s_und = 1;
      }
    }
    run += tot;
    __syncthreads(); /* s_wsum is rewritten next round */
  }
  if (threadIdx.x == 0) {
    bcnt[blockIdx.x] = run;
    bkeep[blockIdx.x] = s_und ? kRecount : run;
  }
  block_flush(S.ctr, CTR_CAND + 1, found);
}
void launch_tri_enum_filter(const Stage& S, Cnt ne, const uint64_t* erow,
                            const uint32_t* ecol, const uint32_t* eleft,
                            const uint64_t* off, uint32_t* cval,
                            uint32_t* cseg, uint32_t* cea, uint32_t* ceb,
                            uint8_t* flag, uint32_t* bcnt, uint32_t* bkeep,
                            Queue q1, PipeLaunch L, uint64_t lim) {
  if (ne.cap <= 0) return;
  dim3 gr((unsigned)((ne.cap + kBlock - 1) / kBlock));
  /* byte offsets into ecol while it stays below 4 GiB */
  const bool buf = ne.cap < ((int64_t)1 << 30);
  auto kf = S.staged ? (buf ? k_tri_enum_filter<float, true>
                            : k_tri_enum_filter<float, false>)
                     : (buf ? k_tri_enum_filter<double, true>
                            : k_tri_enum_filter<double, false>);
  hipLaunchKernelGGL(kf, gr, dim3(kBlock), 0, L.s, S, ne, erow, ecol, eleft, off,
                     cval, cseg, cea, ceb, flag, bcnt, bkeep, q1, lim);
}
/* the kept triangles of block b to (tk, tedge, teik, tejk)[bout[b], ...)
 * in order; trow[e] for the block's edges e (and trow[ne], by the block
 * holding index ne) = bout[b] + the block's kept triangles of edges
 * before e (counted per edge in LDS, then scanned).  Grid: nblk + 1
 * blocks, nblk = ceil(ne.cap / 256); bout has nblk + 1 entries. */
__global__ __launch_bounds__(kBlock) void k_tri_block_copy(
    Cnt c_ne, const uint64_t* off, const uint32_t* bcnt, const uint32_t* cval,
    const uint32_t* cseg, const uint32_t* cea, const uint32_t* ceb,
    const uint8_t* flag, const uint64_t* bout, uint32_t* tk, uint32_t* tedge,
    uint32_t* teik, uint32_t* tejk, uint64_t* trow, uint64_t* tpair,
    uint64_t ocap) {
  const int64_t ne = c_ne.get();
  const int64_t b = blockIdx.x;
  const int64_t q0 = b * kBlock;
  if (q0 > ne) return;
  /* wave sums double-buffered by round parity: a round's sums are
   * rewritten two rounds later, after the next round's barrier */
  __shared__ unsigned s_wsum[2][kBlock / 64];
  __shared__ uint32_t s_cnt[kBlock];
  s_cnt[threadIdx.x] = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  const uint64_t o0 = bout[b];
  if (q0 < ne) {
    const uint64_t base = off[q0];
    const uint32_t n = bcnt[b];
    /* every entry kept (bout's step equals the written count): positions
     * follow without ranking, and no flag is read */
    const bool all = bout[b + 1] - o0 == n;
    uint64_t o = o0;
    int par = 0;
    for (uint32_t r0 = 0; r0 < n; r0 += kBlock, par ^= 1) {
      const uint32_t t = r0 + threadIdx.x;
      /* every load of the round in flight together: an entry's arrays are
       * read whether or not it is kept (all were written) */
      uint32_t fv = KEEP, sv = 0xffffffffu, xv = 0u, av = 0u, bv = 0u;
      if (t < n) {
        if (!all) fv = flag[base + t];
        sv = cseg[base + t];
        xv = cval[base + t];
        av = cea[base + t];
        bv = ceb[base + t];
      }
      const bool keep = t < n && fv == KEEP;
      const uint32_t seg = keep ? sv : 0xffffffffu;
      /* entries come in edge order: one LDS add per run of lanes with
       * the same edge (its first lane adds the run's length) instead of
       * one per lane on the same counter */
      const uint32_t prev = (uint32_t)__shfl_up((int)seg, 1, 64);
      const uint64_t heads = __ballot(lane == 0 || prev != seg);
      if (keep && ((heads >> lane) & 1ull)) {
        const uint64_t later = lane == 63 ? 0ull : heads >> (lane + 1);
        const unsigned len = later ? (unsigned)__builtin_ctzll(later) + 1u
                                   : 64u - (unsigned)lane;
        atomicAdd(&s_cnt[seg - (uint32_t)q0], len);
      }
      if (all) {
        if (keep && o + t - r0 < ocap) {
          const uint64_t at = o + t - r0;
          tk[at] = xv;
          tedge[at] = seg;
          teik[at] = av;
          tejk[at] = bv;
        }
        o += min((uint32_t)kBlock, n - r0);
        continue;
      }
      const uint64_t bm = __ballot(keep);
      if (lane == 0) s_wsum[par][w] = (unsigned)__popcll(bm);
      __syncthreads();
      uint64_t pre = o + (uint64_t)__popcll(bm & lt), tot = 0;
#pragma unroll
      for (int k = 0; k < kBlock / 64; ++k) {
        const unsigned c = s_wsum[par][k];
        if (k < w) pre += c;
        tot += c;
      }
      if (keep && pre < ocap) {
        tk[pre] = xv;
        tedge[pre] = seg;
        teik[pre] = av;
        tejk[pre] = bv;
      }
      o += tot;
    }
  }
  __syncthreads(); /* the per-edge counts complete */
  /* exclusive scan of the per-edge counts (one wave per 64 edges, then
   * the wave totals) */
  const uint32_t c = s_cnt[threadIdx.x];
  uint32_t incl = c;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = (uint32_t)__shfl_up((int)incl, d, 64);
    if (lane >= d) incl += y;
  }
  if (lane == 63) s_wsum[0][w] = incl;
  __syncthreads();
  uint64_t before = o0;
#pragma unroll
  for (int k = 0; k < kBlock / 64; ++k)
    if (k < w) before += s_wsum[0][k];
  const int64_t e = q0 + threadIdx.x;
  if (e <= ne) {
    const uint64_t v = before + incl - c;
    trow[e] = v < ocap ? v : ocap;
  }
  /* the tet slots of the edge's triangles (each one's row tail) */
  if (tpair && e < ne) tpair[e] = c ? (uint64_t)c * (c - 1) / 2 : 0;
}
void launch_tri_block_copy(Cnt ne, const uint64_t* off, const uint32_t* bcnt,
                           const uint32_t* cval, const uint32_t* cseg,
                           const uint32_t* cea, const uint32_t* ceb,
                           const uint8_t* flag, const uint64_t* bout,
                           uint32_t* tk, uint32_t* tedge, uint32_t* teik,
                           uint32_t* tejk, uint64_t* trow, uint64_t* tpair,
                           PipeLaunch L, uint64_t ocap) {
  const int64_t nblk = (std::max<int64_t>(ne.cap, 0) + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_tri_block_copy, dim3((unsigned)(nblk + 1)), dim3(kBlock),
                     0, L.s, ne, off, bcnt, cval, cseg, cea, ceb, flag, bout,
                     tk, tedge, teik, tejk, trow, tpair, ocap);
}

template <int KIND>
__global__ __launch_bounds__(kBlock) void k_filter_esc(Stage S, CandCtx C,
                                                       QRef q,
                                                       const uint32_t* cval,
                                                       uint8_t* flag,
                                                       Queue q2) {
  StageCnt sc;
  for_queue(q, [&](uint32_t t) {
    QEntry e = q.at(t);
    Ball v[4];
    cand_balls<KIND>(S, C, e.item, e.aux, KIND == 3 ? 0u : cval[e.item], v);
    sc.c[1]++;
    int r = filter_code<KIND, FE<double>>(v, S.alpha);
    if (KIND > 1 && r == UNDEC && degenerate_i128<KIND>(v, S.ex.S) == 1) {
      sc.c[3]++; /* decided exactly, in 128-bit integers */
      r = REJECT;
    }
    if (r == UNDEC) {
      sc.c[2]++;
      r = filter_code<KIND, FEDD>(v, S.alpha);
    }
    if (r == UNDEC)
      push_q(q2, e.item, e.aux);
    else
      flag[e.item] = (uint8_t)r;
  });
  sc.flush(S.ctr, fam_of(KIND));
}
/* Exact filter decisions in two steps per wave: every lane takes one
 * queued candidate and settles degenerate ones (collinear / coplanar
 * centres) in 128-bit integers; the wave's other candidates are listed in
 * LDS and taken 4 at a time by its groups of kTPI lanes for the
 * wavefront-cooperative wide-integer evaluation. */
template <int KIND>
__global__ __launch_bounds__(kBlock) void k_filter_exact(Stage S, CandCtx C,
                                                         QRef q,
                                                         const uint32_t* cval,
                                                         uint8_t* flag) {
  __shared__ uint32_t s_pend[kBlock / 64][64];
  unsigned cnt = 0;
  const uint32_t n = q.size();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x >> 6);
  uint32_t* const pend_list = s_pend[w];
  /* wave-uniform loop over 64 entries at a time */
  for (uint64_t base = ((uint64_t)blockIdx.x * (blockDim.x >> 6) + w) * 64;
       base < n; base += waves * 64) {
    const uint32_t t = (uint32_t)(base + lane);
    bool pend = false;
    if (t < n) {
      pend = true;
      if (KIND > 1) {
        const QEntry e = q.at(t);
        Ball v[4];
        cand_balls<KIND>(S, C, e.item, e.aux, KIND == 3 ? 0u : cval[e.item], v);
        if (degenerate_i128<KIND>(v, S.ex.S) == 1) {
          flag[e.item] = REJECT;
          ++cnt;
          pend = false;
        }
      }
    }
    const uint64_t m = __ballot(pend);
    if (pend) pend_list[__popcll(m & lt)] = t;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int np = __popcll(m);
    /* group-uniform loop: group g takes the pending entries g, g + 4, ... */
    for (int k = lane / kTPI; k < np; k += 64 / kTPI) {
      const QEntry e = q.at(pend_list[k]);
      Ball v[4];
      cand_balls<KIND>(S, C, e.item, e.aux, KIND == 3 ? 0u : cval[e.item], v);
      const int r = filter_exact<KIND>(v, S.ex);
      if (Grp<kTPI>::idx() == 0) {
        flag[e.item] = (uint8_t)r;
        ++cnt;
      }
    }
    /* the list is rewritten next round: every group is done with it */
    __builtin_amdgcn_wave_barrier();
  }
  flush_ctr(S.ctr, CTR_STAGE + fam_of(KIND) * 4 + 3, cnt);
}

void launch_filter_esc(const Stage& S, int kind, CandCtx C, QRef q, const uint32_t* cval, uint8_t* flag,
                       Queue q2, PipeLaunch L) {
  if (!q.launch_items()) return;
  dim3 gr = qgrid(q, kEscBlocks);
  if (kind == 1)
    hipLaunchKernelGGL(k_filter_esc<1>, gr, dim3(kBlock), 0, L.s, S, C, q,
                       cval, flag, q2);
  else if (kind == 2)
    hipLaunchKernelGGL(k_filter_esc<2>, gr, dim3(kBlock), 0, L.s, S, C, q,
                       cval, flag, q2);
  else
    hipLaunchKernelGGL(k_filter_esc<3>, gr, dim3(kBlock), 0, L.s, S, C, q,
                       cval, flag, q2);
}
void launch_filter_exact(const Stage& S, int kind, CandCtx C, QRef q, const uint32_t* cval, uint8_t* flag,
                         PipeLaunch L) {
  if (!q.launch_items()) return;
  dim3 gr = qgrid(q, kExactBlocks); /* one entry per lane in the first step */
  if (kind == 1)
    hipLaunchKernelGGL(k_filter_exact<1>, gr, dim3(kBlock), 0, L.s, S, C, q,
                       cval, flag);
  else if (kind == 2)
    hipLaunchKernelGGL(k_filter_exact<2>, gr, dim3(kBlock), 0, L.s, S, C, q,
                       cval, flag);
  else
    hipLaunchKernelGGL(k_filter_exact<3>, gr, dim3(kBlock), 0, L.s, S, C, q,
                       cval, flag);
}

/* ------------------------------------------------------------------ */
/* survivors: segment starts and flat stable compaction                */
/* ------------------------------------------------------------------ */
/* orow[s] = number of survivors before segment s = pos[off[s]] */
}  // namespace PAC_V
}  // namespace pac
