/*
 * k_output.hip -- per-simplex alpha values, their adjoint, and canonical
 * row order.
 *
 * Alpha values: rho = |x|^2 - w_0 with x = c - p_0 the orthocentre offset,
 * evaluated in FP64 with the closed forms the filters use (u, v, t edge
 * vectors from p_0, A = |u|^2 + w_0 - w_1 and so on):
 *   edge      x = A u / (2|u|^2)
 *   triangle  x = (A (v x n) - B (u x n)) / (2|n|^2),  n = u x v
 *   tet       x = (A (v x t) + B (t x u) + C (u x v)) / (2 D),  D = u.(v x t)
 * Canonical order: vertices of each row sorted by original id, rows sorted
 * lexicographically (LSD radix over the columns).
 */
#include "pac_dev.h"

namespace pac {
using namespace PAC_V; /* type-independent device helpers */

/* Orthocentre of a simplex relative to its first vertex, x = c - p_0, and
 * the barycentric coordinates lam of c (c = sum lam_q p_q), in FP64. */
template <int DIM>
__device__ __forceinline__ void ortho_bary(const double4* b, const int32_t* r,
                                           double x[3], double lam[4],
                                           double* rho) {
  double4 a = b[r[0]];
  double wa = a.w * a.w;
  double u[3] = {b[r[1]].x - a.x, b[r[1]].y - a.y, b[r[1]].z - a.z};
  double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  double A = uu + wa - b[r[1]].w * b[r[1]].w;
  lam[2] = lam[3] = 0;
  if (DIM == 1) {
    double t = A / (2.0 * uu);
    for (int q = 0; q < 3; ++q) x[q] = t * u[q];
    lam[1] = t;
  } else if (DIM == 2) {
    double v[3] = {b[r[2]].x - a.x, b[r[2]].y - a.y, b[r[2]].z - a.z};
    double B = v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + wa - b[r[2]].w * b[r[2]].w;
    double n[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2],
                   u[0] * v[1] - u[1] * v[0]};
    double nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    double vn[3] = {v[1] * n[2] - v[2] * n[1], v[2] * n[0] - v[0] * n[2],
                    v[0] * n[1] - v[1] * n[0]};
    double un[3] = {u[1] * n[2] - u[2] * n[1], u[2] * n[0] - u[0] * n[2],
                    u[0] * n[1] - u[1] * n[0]};
    for (int q = 0; q < 3; ++q) x[q] = (A * vn[q] - B * un[q]) / (2.0 * nn);
    /* x = l1 u + l2 v:  l1 = (x x v).n / |n|^2,  l2 = (u x x).n / |n|^2 */
    double xv[3] = {x[1] * v[2] - x[2] * v[1], x[2] * v[0] - x[0] * v[2],
                    x[0] * v[1] - x[1] * v[0]};
    double ux[3] = {u[1] * x[2] - u[2] * x[1], u[2] * x[0] - u[0] * x[2],
                    u[0] * x[1] - u[1] * x[0]};
    lam[1] = (xv[0] * n[0] + xv[1] * n[1] + xv[2] * n[2]) / nn;
    lam[2] = (ux[0] * n[0] + ux[1] * n[1] + ux[2] * n[2]) / nn;
  } else {
    double v[3] = {b[r[2]].x - a.x, b[r[2]].y - a.y, b[r[2]].z - a.z};
    double t[3] = {b[r[3]].x - a.x, b[r[3]].y - a.y, b[r[3]].z - a.z};
    double B = v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + wa - b[r[2]].w * b[r[2]].w;
    double C = t[0] * t[0] + t[1] * t[1] + t[2] * t[2] + wa - b[r[3]].w * b[r[3]].w;
    double vt[3] = {v[1] * t[2] - v[2] * t[1], v[2] * t[0] - v[0] * t[2],
                    v[0] * t[1] - v[1] * t[0]};
    double tu[3] = {t[1] * u[2] - t[2] * u[1], t[2] * u[0] - t[0] * u[2],
                    t[0] * u[1] - t[1] * u[0]};
    double uv[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2],
                    u[0] * v[1] - u[1] * v[0]};
    double D = u[0] * vt[0] + u[1] * vt[1] + u[2] * vt[2];
    for (int q = 0; q < 3; ++q) x[q] = (A * vt[q] + B * tu[q] + C * uv[q]) / (2.0 * D);
    /* Cramer: x = l1 u + l2 v + l3 t */
    lam[1] = (x[0] * vt[0] + x[1] * vt[1] + x[2] * vt[2]) / D;
    lam[2] = (x[0] * tu[0] + x[1] * tu[1] + x[2] * tu[2]) / D;
    lam[3] = (x[0] * uv[0] + x[1] * uv[1] + x[2] * uv[2]) / D;
  }
  lam[0] = 1.0 - lam[1] - lam[2] - lam[3];
  *rho = x[0] * x[0] + x[1] * x[1] + x[2] * x[2] - wa;
}

/* rho (the power of the smallest orthosphere, the quantity the reference
 * compares with alpha and stores in float4 .w, AlphaComplexParallel.cu:370,
 * :603) of every row, from the original-order balls */
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_rho(int64_t m, const int32_t* rows,
                                                const double4* balls,
                                                double* rho) {
  int64_t id = gtid();
  if (id >= m) return;
  double x[3], lam[4], r;
  ortho_bary<DIM>(balls, rows + (int64_t)(DIM + 1) * id, x, lam, &r);
  rho[id] = r;
}
void launch_rho(int dim, int64_t m, const int32_t* rows, const double4* balls,
                double* rho, hipStream_t s) {
  if (m <= 0) return;
  dim3 gr((unsigned)((m + kBlock - 1) / kBlock));
  if (dim == 1)
    hipLaunchKernelGGL(k_rho<1>, gr, dim3(kBlock), 0, s, m, rows, balls, rho);
  else if (dim == 2)
    hipLaunchKernelGGL(k_rho<2>, gr, dim3(kBlock), 0, s, m, rows, balls, rho);
  else
    hipLaunchKernelGGL(k_rho<3>, gr, dim3(kBlock), 0, s, m, rows, balls, rho);
}

/* ---- filtration values ----
 * The alpha value at which a simplex enters the filtration of weighted
 * alpha complexes: a ball at -w; a simplex whose smallest orthosphere is
 * empty (no other ball of power <= rho, the reference's emptiness rule,
 * AlphaComplexLargeData.cu:838, :971) at its own rho; any other at the
 * smallest value among the simplices of the complex it is a face of.  Tets
 * are always their own rho.  Values are handled as order-preserving
 * 64-bit keys (atomic minimum, radix sort). */
__device__ __forceinline__ uint64_t ord64(double v) {
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double unord64(uint64_t k) {
  const uint64_t b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double((long long)b);
}
constexpr uint64_t kNoBirth = ~0ull;

/* key[i] = rho of row i if it is its own (gab == nullptr or gab[i]), else
 * "not yet" (kNoBirth) */
__global__ __launch_bounds__(kBlock) void k_birth_init(int64_t m,
                                                       const double* rho,
                                                       const uint8_t* gab,
                                                       uint64_t* key) {
  const int64_t i = gtid();
  if (i >= m) return;
  key[i] = (!gab || gab[i]) ? ord64(rho[i]) : kNoBirth;
}
/* balls: -w = -r^2 */
__global__ __launch_bounds__(kBlock) void k_birth_vertex(int64_t n,
                                                         const double4* balls,
                                                         uint64_t* key) {
  const int64_t i = gtid();
  if (i >= n) return;
  const double r = balls[i].w;
  key[i] = ord64(-(r * r));
}

/* Row of `face` (K columns, rows lexicographic by rank) equal to the K
 * ranks f[], or -1. */
template <int K>
__device__ __forceinline__ int64_t find_row(const int32_t* face, int64_t mf,
                                            const uint32_t* rank,
                                            const uint32_t* f) {
  int64_t lo = 0, hi = mf; /* first row >= f */
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    int c = 0;
    for (int q = 0; q < K && c == 0; ++q) {
      const uint32_t v = rank[face[mid * K + q]];
      c = v < f[q] ? -1 : (v > f[q] ? 1 : 0);
    }
    if (c < 0)
      lo = mid + 1;
    else
      hi = mid;
  }
  if (lo >= mf) return -1;
  for (int q = 0; q < K; ++q)
    if (rank[face[lo * K + q]] != f[q]) return -1;
  return lo;
}

/* Every coface row (K+1 columns) lowers the key of each of its faces that
 * is not its own value (gab == 0) to its key.  Rows of both lists are in
 * rank order within the row and lexicographic by rank (reference order). */
template <int K>
__global__ __launch_bounds__(kBlock) void k_birth_faces(
    int64_t mc, const int32_t* cof, const uint64_t* ckey, int64_t mf,
    const int32_t* face, const uint8_t* fgab, uint64_t* fkey,
    const uint32_t* rank, unsigned* missing) {
  const int64_t c = gtid();
  if (c >= mc) return;
  uint32_t r[K + 1];
  for (int q = 0; q <= K; ++q) r[q] = rank[cof[c * (K + 1) + q]];
  const uint64_t k = ckey[c];
  for (int drop = 0; drop <= K; ++drop) {
    uint32_t f[K];
    int w = 0;
    for (int q = 0; q <= K; ++q)
      if (q != drop) f[w++] = r[q];
    const int64_t row = find_row<K>(face, mf, rank, f);
    if (row < 0) {
      atomicAdd(missing, 1u);
      continue;
    }
    if (!fgab[row]) atomicMin((unsigned long long*)&fkey[row], (unsigned long long)k);
  }
}

/* one key array over all simplices (balls, edges, triangles, tets) to
 * values; counts keys left without a value */
__global__ __launch_bounds__(kBlock) void k_birth_values(int64_t m,
                                                         const uint64_t* key,
                                                         double* out,
                                                         unsigned* missing) {
  const int64_t i = gtid();
  if (i >= m) return;
  const uint64_t k = key[i];
  if (k == kNoBirth) atomicAdd(missing, 1u);
  out[i] = unord64(k);
}

__global__ __launch_bounds__(kBlock) void k_split_key(int64_t m,
                                                      const uint64_t* key,
                                                      uint32_t* lo,
                                                      uint32_t* idx) {
  const int64_t i = gtid();
  if (i >= m) return;
  lo[i] = (uint32_t)key[i];
  idx[i] = (uint32_t)i;
}
__global__ __launch_bounds__(kBlock) void k_high_key(int64_t m,
                                                     const uint64_t* key,
                                                     const uint32_t* idx,
                                                     uint32_t* hi) {
  const int64_t i = gtid();
  if (i >= m) return;
  hi[i] = (uint32_t)(key[idx[i]] >> 32);
}
/* sorted position p -> (dim, index in its list, value) */
__global__ __launch_bounds__(kBlock) void k_filtration_out(
    int64_t m, const uint32_t* idx, const uint64_t* key, int64_t b1,
    int64_t b2, int64_t b3, int32_t* dim, int64_t* index, double* value) {
  const int64_t p = gtid();
  if (p >= m) return;
  const int64_t g = idx[p];
  const int d = g >= b3 ? 3 : (g >= b2 ? 2 : (g >= b1 ? 1 : 0));
  const int64_t base = d == 3 ? b3 : (d == 2 ? b2 : (d == 1 ? b1 : 0));
  dim[p] = d;
  index[p] = g - base;
  value[p] = unord64(key[g]);
}

void birth_keys(const int32_t* const lists[3], const int64_t counts[4],
                const double4* balls, const uint8_t* const gab[2],
                const uint32_t* rank, uint64_t* key, unsigned* missing,
                hipStream_t s) {
  const int64_t n = counts[0];
  const int64_t off[4] = {0, n, n + counts[1], n + counts[1] + counts[2]};
  auto grid = [](int64_t m) { return dim3((unsigned)((m + kBlock - 1) / kBlock)); };
  if (n > 0)
    hipLaunchKernelGGL(k_birth_vertex, grid(n), dim3(kBlock), 0, s, n, balls, key);
  for (int d = 1; d <= 3; ++d) {
    const int64_t m = counts[d];
    if (m <= 0) continue;
    launch_rho(d, m, lists[d - 1], balls, (double*)(key + off[d]), s);
    /* rho in place, then to keys */
    hipLaunchKernelGGL(k_birth_init, grid(m), dim3(kBlock), 0, s, m,
                       (const double*)(key + off[d]), d == 3 ? nullptr : gab[d - 1],
                       key + off[d]);
  }
  /* tets -> triangles, then triangles -> edges */
  if (counts[3] > 0 && counts[2] > 0)
    hipLaunchKernelGGL(k_birth_faces<3>, grid(counts[3]), dim3(kBlock), 0, s,
                       counts[3], lists[2], key + off[3], counts[2], lists[1],
                       gab[1], key + off[2], rank, missing);
  if (counts[2] > 0 && counts[1] > 0)
    hipLaunchKernelGGL(k_birth_faces<2>, grid(counts[2]), dim3(kBlock), 0, s,
                       counts[2], lists[1], key + off[2], counts[1], lists[0],
                       gab[0], key + off[1], rank, missing);
}

void birth_values(int64_t m, const uint64_t* key, double* out,
                  unsigned* missing, hipStream_t s) {
  if (m <= 0) return;
  hipLaunchKernelGGL(k_birth_values, dim3((unsigned)((m + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, s, m, key, out, missing);
}

void filtration_order(int64_t m, const uint64_t* key, const int64_t counts[4],
                      int32_t* dim, int64_t* index, double* value,
                      uint32_t* a, uint32_t* b, uint32_t* c, uint32_t* e,
                      void* rtmp, size_t rtb, hipStream_t s) {
  if (m <= 0) return;
  const dim3 gr((unsigned)((m + kBlock - 1) / kBlock));
  /* stable LSD over the two 32-bit halves of the keys: ties keep the input
   * order (balls, edges, triangles, tets; list order within), so a face
   * entering with its coface comes first */
  hipLaunchKernelGGL(k_split_key, gr, dim3(kBlock), 0, s, m, key, a, b);
  bool alt = false;
  radix_sort_pairs(a, b, c, e, m, 32, rtmp, rtb, s, &alt);
  uint32_t *rk = alt ? c : a, *rv = alt ? e : b;
  uint32_t *fk = alt ? a : c, *fv = alt ? b : e;
  hipLaunchKernelGGL(k_high_key, gr, dim3(kBlock), 0, s, m, key, rv, rk);
  bool alt2 = false;
  radix_sort_pairs(rk, rv, fk, fv, m, 32, rtmp, rtb, s, &alt2);
  const uint32_t* fin = alt2 ? fv : rv;
  const int64_t b1 = counts[0], b2 = b1 + counts[1], b3 = b2 + counts[2];
  hipLaunchKernelGGL(k_filtration_out, gr, dim3(kBlock), 0, s, m, fin, key, b1,
                     b2, b3, dim, index, value);
}

/* ---- adjoint of rho with the combinatorics fixed ----
 * Every vertex q of a simplex satisfies rho = |c - p_q|^2 - w_q.  With
 * c = sum lam_q p_q (sum lam_q = 1, c in the affine hull), summing the
 * differentials with weights lam_q removes dc:
 *     drho = sum_q lam_q (2 (p_q - c) . dp_q - dw_q),   dw_q = 2 r_q dr_q,
 * so d rho / d p_q = 2 lam_q (p_q - c) and d rho / d r_q = -2 lam_q r_q.
 * Contributions are written per (simplex, vertex) slot and summed per ball
 * in slot order after a stable sort by ball: the result is deterministic. */
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_adj_contrib(
    int64_t m, const int32_t* rows, const double4* balls, const double* g,
    int64_t slot0, uint32_t* key, uint32_t* val, double4* contrib) {
  int64_t id = gtid();
  if (id >= m) return;
  const int32_t* r = rows + (int64_t)(DIM + 1) * id;
  double x[3], lam[4], rho;
  ortho_bary<DIM>(balls, r, x, lam, &rho);
  const double gs = g[id];
  double4 a = balls[r[0]];
#pragma unroll
  for (int q = 0; q <= DIM; ++q) {
    double4 p = balls[r[q]];
    /* p_q - c = (p_q - p_0) - x */
    double dx = (p.x - a.x) - x[0], dy = (p.y - a.y) - x[1], dz = (p.z - a.z) - x[2];
    double k = 2.0 * lam[q] * gs;
    int64_t sl = slot0 + (int64_t)(DIM + 1) * id + q;
    contrib[sl] = make_double4(k * dx, k * dy, k * dz, -k * p.w);
    key[sl] = (uint32_t)r[q];
    val[sl] = (uint32_t)sl;
  }
}
__global__ __launch_bounds__(kBlock) void k_adj_gather(
    int64_t n, const uint32_t* vstart, const uint32_t* sval,
    const double4* contrib, double* dxyz, double* dr) {
  int64_t i = gtid();
  if (i >= n) return;
  double sx = 0, sy = 0, sz = 0, sr = 0;
  for (uint32_t q = vstart[i], e = vstart[i + 1]; q < e; ++q) {
    double4 c = contrib[sval[q]];
    sx += c.x;
    sy += c.y;
    sz += c.z;
    sr += c.w;
  }
  dxyz[3 * i] = sx;
  dxyz[3 * i + 1] = sy;
  dxyz[3 * i + 2] = sz;
  dr[i] = sr;
}
void launch_adj_contrib(int dim, int64_t m, const int32_t* rows,
                        const double4* balls, const double* g, int64_t slot0,
                        uint32_t* key, uint32_t* val, double4* contrib,
                        hipStream_t s) {
  if (m <= 0) return;
  dim3 gr((unsigned)((m + kBlock - 1) / kBlock));
  if (dim == 1)
    hipLaunchKernelGGL(k_adj_contrib<1>, gr, dim3(kBlock), 0, s, m, rows, balls,
                       g, slot0, key, val, contrib);
  else if (dim == 2)
    hipLaunchKernelGGL(k_adj_contrib<2>, gr, dim3(kBlock), 0, s, m, rows, balls,
                       g, slot0, key, val, contrib);
  else
    hipLaunchKernelGGL(k_adj_contrib<3>, gr, dim3(kBlock), 0, s, m, rows, balls,
                       g, slot0, key, val, contrib);
}
void launch_adj_gather(int64_t n, const uint32_t* vstart, const uint32_t* sval,
                       const double4* contrib, double* dxyz, double* dr,
                       hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_adj_gather, dim3((unsigned)((n + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, s, n, vstart, sval, contrib, dxyz, dr);
}

/* sort the vertices of each row ascending (rows of k = DIM+1 ids) */
template <int K>
__global__ __launch_bounds__(kBlock) void k_sort_rows(int64_t m, int32_t* rows) {
  int64_t id = gtid();
  if (id >= m) return;
  int32_t v[K];
#pragma unroll
  for (int q = 0; q < K; ++q) v[q] = rows[K * id + q];
#pragma unroll
  for (int a = 0; a < K; ++a)
#pragma unroll
    for (int b = 0; b + 1 < K - a; ++b)
      if (v[b] > v[b + 1]) {
        int32_t t = v[b];
        v[b] = v[b + 1];
        v[b + 1] = t;
      }
#pragma unroll
  for (int q = 0; q < K; ++q) rows[K * id + q] = v[q];
}
/* key[i] = rows[K*idx[i] + col]; val stays idx */
__global__ __launch_bounds__(kBlock) void k_col_key(int64_t m, int K,
                                                    const int32_t* rows,
                                                    const uint32_t* idx,
                                                    int col, uint32_t* key) {
  int64_t i = gtid();
  if (i >= m) return;
  key[i] = (uint32_t)rows[(int64_t)K * idx[i] + col];
}
__global__ __launch_bounds__(kBlock) void k_iota(int64_t m, uint32_t* idx) {
  int64_t i = gtid();
  if (i < m) idx[i] = (uint32_t)i;
}
__global__ __launch_bounds__(kBlock) void k_gather_rows(int64_t m, int K,
                                                        const int32_t* in,
                                                        const uint32_t* idx,
                                                        int32_t* out) {
  int64_t i = gtid();
  if (i >= m) return;
  for (int q = 0; q < K; ++q) out[(int64_t)K * i + q] = in[(int64_t)K * idx[i] + q];
}

/* canonical order of m rows of K ids (< 2^idbits) in place; tmp buffers are
 * provided by the caller: idx/key/idx2/key2 (m each), radix tmp, rows copy */
void canonical_rows(int K, int64_t m, int32_t* rows, int idbits,
                    uint32_t* idx, uint32_t* key, uint32_t* idx2,
                    uint32_t* key2, void* rtmp, size_t rtb, int32_t* copy,
                    hipStream_t s) {
  if (m <= 0) return;
  dim3 gr((unsigned)((m + kBlock - 1) / kBlock));
  if (K == 2)
    hipLaunchKernelGGL(k_sort_rows<2>, gr, dim3(kBlock), 0, s, m, rows);
  else if (K == 3)
    hipLaunchKernelGGL(k_sort_rows<3>, gr, dim3(kBlock), 0, s, m, rows);
  else
    hipLaunchKernelGGL(k_sort_rows<4>, gr, dim3(kBlock), 0, s, m, rows);
  hipLaunchKernelGGL(k_iota, gr, dim3(kBlock), 0, s, m, idx);
  for (int col = K - 1; col >= 0; --col) {
    hipLaunchKernelGGL(k_col_key, gr, dim3(kBlock), 0, s, m, K,
                       (const int32_t*)rows, (const uint32_t*)idx, col, key);
    bool alt = false;
    radix_sort_pairs(key, idx, key2, idx2, m, idbits, rtmp, rtb, s, &alt);
    if (alt) {
      uint32_t* t = idx;
      idx = idx2;
      idx2 = t;
      t = key;
      key = key2;
      key2 = t;
    }
  }
  hipLaunchKernelGGL(k_gather_rows, gr, dim3(kBlock), 0, s, m, K,
                     (const int32_t*)rows, (const uint32_t*)idx, copy);
  (void)hipMemcpyAsync(rows, copy, (size_t)m * K * sizeof(int32_t),
                       hipMemcpyDeviceToDevice, s);
}

/* Rows of a batch's list are grouped by molecule in molecule order and the
 * first id of a row lies in its molecule's id range, so "first id >=
 * moff[m]" is monotone over the rows: out[m] = first row of molecule m. */
__global__ __launch_bounds__(kBlock) void k_mol_rows(const int32_t* rows,
                                                    Cnt c_m, int K,
                                                    const uint32_t* moff,
                                                    int64_t nmol,
                                                    uint64_t* out) {
  const int64_t m = c_m.get();
  int64_t q = gtid();
  if (q > nmol) return;
  uint32_t key = moff[q];
  int64_t lo = 0, hi = m;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if ((uint32_t)rows[(int64_t)K * mid] < key) lo = mid + 1;
    else hi = mid;
  }
  out[q] = (uint64_t)lo;
}
void launch_mol_rows(const int32_t* rows, Cnt m, int K,
                     const uint32_t* moff, int64_t nmol, uint64_t* out,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_mol_rows, dim3((unsigned)((nmol + 1 + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, s, rows, m, K, moff, nmol, out);
}

}  // namespace pac
