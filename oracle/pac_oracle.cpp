/*
 * pac_oracle.cpp -- TEST INFRASTRUCTURE ONLY (see pac_oracle.h).
 *
 * CPU restatement of the reference pipeline with exact decisions.  Each
 * predicate is evaluated first in FP64 with a running (Wilkinson-style)
 * absolute error bound; when the sign is not certified it is recomputed in
 * fixed-width two's-complement integers on inputs scaled by a global power
 * of two, which is exact.  Ties in the tet emptiness test are broken by
 * Edelsbrunner-Muecke simulation of simplicity; the term list is enumerated
 * here from the definition (perturbation eps^(2^(4(i+1)-j)) of coordinate j
 * of the i-th point in index order), not transcribed.
 *
 * Reference correspondence:
 *   cell key                 AlphaComplexLargeData.cu:65-78, Parallel.cu:1120-1124
 *   candidate edges          AlphaComplexLargeData.cu:80-165
 *   markRealEdges            AlphaComplexLargeData.cu:167-190
 *   candidate triangles      AlphaComplexLargeData.cu:192-296
 *   markRealTriangles        AlphaComplexLargeData.cu:298-348
 *   candidate tets           AlphaComplexLargeData.cu:350-529
 *   markRealTets             AlphaComplexLargeData.cu:531-591
 *   checkTetsOrthosphere     AlphaComplexLargeData.cu:597-735, predicates.h:2015,4250,4669
 *   closure (tris / edges)   AlphaComplexParallel.cu:1480-1788
 *   checkDanglingTris        AlphaComplexLargeData.cu:807-881
 *   checkDanglingEdges       AlphaComplexLargeData.cu:928-1015
 */
#include "pac_oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <new>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

typedef __int128 i128;
typedef unsigned __int128 u128;

/* ------------------------------------------------------------------ */
/* Fixed-width two's-complement integers, L 64-bit limbs.              */
/* ------------------------------------------------------------------ */
template <int L>
struct Big {
  uint64_t d[L];
  Big() { std::memset(d, 0, sizeof(d)); }
  static Big from_i128(i128 v) {
    Big r;
    r.d[0] = (uint64_t)v;
    if (L > 1) r.d[1] = (uint64_t)(v >> 64);
    uint64_t ext = v < 0 ? ~0ull : 0ull;
    for (int i = 2; i < L; ++i) r.d[i] = ext;
    return r;
  }
  static Big from_int(long long v) { return from_i128((i128)v); }
  int sign() const {
    if (d[L - 1] >> 63) return -1;
    for (int i = 0; i < L; ++i)
      if (d[i]) return 1;
    return 0;
  }
  bool operator==(const Big& o) const {
    return std::memcmp(d, o.d, sizeof(d)) == 0;
  }
};
template <int L>
inline Big<L> operator+(const Big<L>& a, const Big<L>& b) {
  Big<L> r;
  u128 c = 0;
  for (int i = 0; i < L; ++i) {
    c += (u128)a.d[i] + b.d[i];
    r.d[i] = (uint64_t)c;
    c >>= 64;
  }
  return r;
}
template <int L>
inline Big<L> operator-(const Big<L>& a, const Big<L>& b) {
  Big<L> r;
  uint64_t borrow = 0;
  for (int i = 0; i < L; ++i) {
    uint64_t ai = a.d[i], bi = b.d[i];
    uint64_t t = ai - bi;
    uint64_t b1 = ai < bi;
    uint64_t t2 = t - borrow;
    uint64_t b2 = t < borrow;
    r.d[i] = t2;
    borrow = b1 | b2;
  }
  return r;
}
template <int L>
inline Big<L> operator-(const Big<L>& a) {
  return Big<L>() - a;
}
template <int L>
inline Big<L> operator*(const Big<L>& a, const Big<L>& b) {
  /* Work on magnitudes so that small operands stay cheap. */
  bool na = a.sign() < 0, nb = b.sign() < 0;
  Big<L> ma = na ? -a : a, mb = nb ? -b : b;
  int la = L, lb = L;
  while (la > 0 && ma.d[la - 1] == 0) --la;
  while (lb > 0 && mb.d[lb - 1] == 0) --lb;
  Big<L> r;
  for (int i = 0; i < la; ++i) {
    u128 carry = 0;
    uint64_t ai = ma.d[i];
    if (!ai) continue;
    for (int j = 0; j < lb && i + j < L; ++j) {
      u128 t = (u128)ai * mb.d[j] + r.d[i + j] + carry;
      r.d[i + j] = (uint64_t)t;
      carry = t >> 64;
    }
    if (i + lb < L) r.d[i + lb] = (uint64_t)carry;
  }
  return (na != nb) ? -r : r;
}

/* v * 2^shift as an exact integer (v*2^shift must be integral). */
template <int L>
Big<L> big_from_scaled(double v, int shift) {
  if (v == 0.0) return Big<L>();
  int e;
  double m = std::frexp(v, &e); /* v = m * 2^e, 0.5 <= |m| < 1 */
  long long mi = (long long)std::ldexp(m, 53);
  int ex = e - 53 + shift; /* v*2^shift = mi * 2^ex */
  while (ex < 0 && (mi % 2) == 0) {
    mi /= 2;
    ++ex;
  }
  if (ex < 0) return Big<L>(); /* not integral: caller checks scale */
  Big<L> r = Big<L>::from_int(mi < 0 ? -mi : mi);
  /* shift left by ex */
  int limbs = ex / 64, bits = ex % 64;
  Big<L> s;
  for (int i = L - 1; i >= 0; --i) {
    int src = i - limbs;
    uint64_t lo = (src >= 0) ? r.d[src] : 0;
    uint64_t lo2 = (src - 1 >= 0) ? r.d[src - 1] : 0;
    s.d[i] = bits ? ((lo << bits) | (lo2 >> (64 - bits))) : lo;
  }
  return mi < 0 ? -s : s;
}

/* ------------------------------------------------------------------ */
/* FP64 value with a certified absolute error bound.                   */
/* ------------------------------------------------------------------ */
struct FE {
  double v, e;
};
const double kU = 0x1p-53;
const double kTiny = 0x1p-1000; /* covers underflow of any single op */
inline FE fe(double x) { return FE{x, 0.0}; }
inline FE operator+(FE a, FE b) {
  double s = a.v + b.v;
  return FE{s, a.e + b.e + kU * std::fabs(s) + kTiny};
}
inline FE operator-(FE a, FE b) {
  double s = a.v - b.v;
  return FE{s, a.e + b.e + kU * std::fabs(s) + kTiny};
}
inline FE operator-(FE a) { return FE{-a.v, a.e}; }
inline FE operator*(FE a, FE b) {
  double p = a.v * b.v;
  return FE{p, std::fabs(a.v) * b.e + std::fabs(b.v) * a.e + a.e * b.e +
                   kU * std::fabs(p) + kTiny};
}
const int UNDECIDED = 2;
inline int fe_sign(FE a) {
  double bound = a.e * (1.0 + 0x1p-20);
  if (a.v > bound) return 1;
  if (a.v < -bound) return -1;
  return UNDECIDED;
}

/* ------------------------------------------------------------------ */
/* Generic 3-vector helpers over any ring type.                        */
/* ------------------------------------------------------------------ */
template <class T>
struct V3 {
  T x, y, z;
};
template <class T>
inline V3<T> vsub(const V3<T>& a, const V3<T>& b) {
  return V3<T>{a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class T>
inline T vdot(const V3<T>& a, const V3<T>& b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class T>
inline V3<T> vcross(const V3<T>& a, const V3<T>& b) {
  return V3<T>{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z,
               a.x * b.y - a.y * b.x};
}
template <class T>
inline V3<T> vscale(const T& s, const V3<T>& a) {
  return V3<T>{s * a.x, s * a.y, s * a.z};
}
template <class T>
inline V3<T> vadd(const V3<T>& a, const V3<T>& b) {
  return V3<T>{a.x + b.x, a.y + b.y, a.z + b.z};
}

/* ------------------------------------------------------------------ */
/* Predicate polynomials.  All take absolute positions p and weights w */
/* in the ring T and the alpha value al.                               */
/* ------------------------------------------------------------------ */

/* Edge (i,j): rho <= alpha  <=>  4|u|^2 (w_i + al) - A^2 >= 0,
 * u = p_j - p_i, A = |u|^2 + w_i - w_j   (LD:167-190 multiplied through). */
template <class T>
inline T edge_val(const V3<T>& pi, const T& wi, const V3<T>& pj, const T& wj,
                  const T& al, const T& four, T* uu_out) {
  V3<T> u = vsub(pj, pi);
  T uu = vdot(u, u);
  T A = uu + wi - wj;
  if (uu_out) *uu_out = uu;
  return four * uu * (wi + al) - A * A;
}

/* Triangle (i,j,k): orthocentre x = (A (v x n) - B (u x n)) / (2|n|^2),
 * rho = |x|^2 - w_i, so rho <= alpha <=>
 * 4|n|^2 (w_i+al) - (A^2|v|^2 - 2AB u.v + B^2|u|^2) >= 0   (LD:298-348). */
template <class T>
inline T tri_val(const V3<T>& pi, const T& wi, const V3<T>& pj, const T& wj,
                 const V3<T>& pk, const T& wk, const T& al, const T& two,
                 const T& four, T* nn_out) {
  V3<T> u = vsub(pj, pi), v = vsub(pk, pi);
  V3<T> n = vcross(u, v);
  T nn = vdot(n, n);
  T uu = vdot(u, u), vv = vdot(v, v), uv = vdot(u, v);
  T A = uu + wi - wj, B = vv + wi - wk;
  if (nn_out) *nn_out = nn;
  T lhs = A * A * vv - two * A * B * uv + B * B * uu;
  return four * nn * (wi + al) - lhs;
}

/* Tet (i,j,k,l): x = N / (2D), N = A(v x t) + B(t x u) + C(u x v),
 * D = u.(v x t); rho <= alpha <=> 4D^2(w_i+al) - |N|^2 >= 0  (LD:531-591). */
template <class T>
inline T tet_val(const V3<T>& pi, const T& wi, const V3<T>& pj, const T& wj,
                 const V3<T>& pk, const T& wk, const V3<T>& pl, const T& wl,
                 const T& al, const T& four, T* D_out, V3<T>* N_out) {
  V3<T> u = vsub(pj, pi), v = vsub(pk, pi), t = vsub(pl, pi);
  V3<T> vt = vcross(v, t), tu = vcross(t, u), uv = vcross(u, v);
  T D = vdot(u, vt);
  T A = vdot(u, u) + wi - wj, B = vdot(v, v) + wi - wk,
    C = vdot(t, t) + wi - wl;
  V3<T> N = vadd(vadd(vscale(A, vt), vscale(B, tu)), vscale(C, uv));
  if (D_out) *D_out = D;
  if (N_out) *N_out = N;
  return four * D * D * (wi + al) - vdot(N, N);
}

/* ------------------------------------------------------------------ */
/* SoS term list for det[x y z l 1] of 5 points (Edelsbrunner-Muecke).  */
/* ------------------------------------------------------------------ */
struct SosTerm {
  int rows;  /* bitmask of kept rows (of 5) */
  int cols;  /* bitmask of kept columns (of 5: x y z l 1) */
  int sign;  /* sign of the eps coefficient relative to the minor */
};

std::vector<SosTerm> build_sos_terms() {
  struct Cand {
    long long ex;
    int R[4], C[4], k;
  };
  std::vector<Cand> all;
  /* partial matchings rows -> perturbed columns (0..3) */
  std::function<void(int, int, int, Cand&)> rec = [&](int r0, int usedR,
                                                      int usedC, Cand& c) {
    all.push_back(c);
    if (c.k == 4) return;
    for (int r = r0; r < 5; ++r) {
      if (usedR & (1 << r)) continue;
      for (int col = 0; col < 4; ++col) {
        if (usedC & (1 << col)) continue;
        Cand n = c;
        n.R[n.k] = r;
        n.C[n.k] = col;
        n.k++;
        n.ex += 1ll << ((r + 1) * 4 - (col + 1));
        rec(r + 1, usedR | (1 << r), usedC | (1 << col), n);
      }
    }
  };
  Cand root{};
  root.ex = 0;
  root.k = 0;
  rec(0, 0, 0, root);
  std::sort(all.begin(), all.end(),
            [](const Cand& a, const Cand& b) { return a.ex < b.ex; });
  std::vector<SosTerm> out;
  std::vector<std::pair<int, int>> seen;
  for (const Cand& c : all) {
    if (c.k == 0) continue;
    int rows = 0x1f, cols = 0x1f;
    int sumrc = 0;
    for (int q = 0; q < c.k; ++q) {
      rows &= ~(1 << c.R[q]);
      cols &= ~(1 << c.C[q]);
      sumrc += c.R[q] + c.C[q];
    }
    bool dup = false;
    for (auto& s : seen)
      if (s.first == rows && s.second == cols) dup = true;
    if (dup) continue;
    seen.push_back({rows, cols});
    /* sign of sigma: rows are listed ascending; columns as assigned */
    int inv = 0;
    for (int a = 0; a < c.k; ++a)
      for (int b = a + 1; b < c.k; ++b)
        if (c.C[a] > c.C[b]) ++inv;
    int sign = ((sumrc + inv) % 2) ? -1 : 1;
    out.push_back(SosTerm{rows, cols, sign});
    if (__builtin_popcount(cols) == 1) break; /* constant minor: done */
  }
  return out;
}
const std::vector<SosTerm>& sos_terms() {
  static const std::vector<SosTerm> t = build_sos_terms();
  return t;
}

template <int L>
Big<L> minor_det(const Big<L> M[5][5], const int* rows, const int* cols,
                 int m) {
  if (m == 1) return M[rows[0]][cols[0]];
  Big<L> acc;
  int sub_cols[5];
  for (int c = 0; c < m; ++c) {
    int q = 0;
    for (int cc = 0; cc < m; ++cc)
      if (cc != c) sub_cols[q++] = cols[cc];
    Big<L> term = M[rows[0]][cols[c]] * minor_det(M, rows + 1, sub_cols, m - 1);
    acc = (c % 2) ? acc - term : acc + term;
  }
  return acc;
}

/* Sign of the SoS-perturbed det[x y z l 1] with rows in argument order.
 * P[r] = {X, Y, Z, L} exact, ids give the perturbation order. */
template <int L>
int sos_det5(const Big<L> P[5][4], const int64_t ids[5], bool* was_zero) {
  int ord[5] = {0, 1, 2, 3, 4};
  int swaps = 0;
  for (int a = 0; a < 5; ++a)
    for (int b = 0; b + 1 < 5 - a; ++b)
      if (ids[ord[b + 1]] < ids[ord[b]]) {
        std::swap(ord[b], ord[b + 1]);
        ++swaps;
      }
  Big<L> M[5][5];
  for (int r = 0; r < 5; ++r) {
    for (int c = 0; c < 4; ++c) M[r][c] = P[ord[r]][c];
    M[r][4] = Big<L>::from_int(1);
  }
  int all_rows[5] = {0, 1, 2, 3, 4}, all_cols[5] = {0, 1, 2, 3, 4};
  int s0 = minor_det(M, all_rows, all_cols, 5).sign();
  int parity = (swaps % 2) ? -1 : 1;
  if (was_zero) *was_zero = (s0 == 0);
  if (s0 != 0) return parity * s0;
  for (const SosTerm& t : sos_terms()) {
    int rows[5], cols[5], m = 0, mc = 0;
    for (int r = 0; r < 5; ++r)
      if (t.rows & (1 << r)) rows[m++] = r;
    for (int c = 0; c < 5; ++c)
      if (t.cols & (1 << c)) cols[mc++] = c;
    int s = minor_det(M, rows, cols, m).sign();
    if (s != 0) return parity * t.sign * s;
  }
  return parity; /* unreachable: last term is the constant 1 */
}

/* ------------------------------------------------------------------ */
/* Input embedding.                                                    */
/* ------------------------------------------------------------------ */
template <class F>
int lsb_exponent(F v) {
  /* exponent of the lowest set bit of a finite nonzero value */
  int e;
  double m = std::frexp((double)v, &e);
  long long mi = (long long)std::ldexp(std::fabs(m), 53);
  int tz = __builtin_ctzll((unsigned long long)mi);
  return e - 53 + tz;
}

template <class F, int L>
struct Balls {
  int64_t n = 0;
  std::vector<F> x, y, z, r;       /* sorted order */
  std::vector<double> w;           /* approximate r^2 (exact for float) */
  std::vector<int64_t> orig;       /* sorted -> original id */
  std::vector<i128> X, Y, Z;       /* exact scaled coordinates */
  std::vector<Big<L>> W;           /* exact scaled weights */
  Big<L> AL;                       /* exact scaled alpha */
  double alpha = 0;
  int S = 0;
  /* grid */
  F minx = 0, miny = 0, minz = 0, step = 0;
  int ex = 0, ey = 0, ez = 0;
  std::vector<uint32_t> cell;      /* per sorted ball */
  std::vector<uint32_t> cstart;    /* G+1 */
  int k_enum = 1, k_ac2 = 1;
  bool no_edges = false;

  V3<FE> pf(uint32_t i) const {
    return V3<FE>{fe((double)x[i]), fe((double)y[i]), fe((double)z[i])};
  }
  FE wf(uint32_t i) const {
    FE rr = fe((double)r[i]);
    return rr * rr;
  }
  V3<Big<L>> pb(uint32_t i) const {
    return V3<Big<L>>{Big<L>::from_i128(X[i]), Big<L>::from_i128(Y[i]),
                      Big<L>::from_i128(Z[i])};
  }
};

struct Counters {
  std::atomic<int64_t> exact_filter{0}, exact_ac2{0}, sos{0}, pairs{0};
};

/* ---------------- exact decisions (FE filter, then integers) -------- */
template <class F, int L>
bool same_pos(const Balls<F, L>& B, uint32_t i, uint32_t j) {
  return B.x[i] == B.x[j] && B.y[i] == B.y[j] && B.z[i] == B.z[j];
}

template <class F, int L>
bool edge_keep(const Balls<F, L>& B, uint32_t i, uint32_t j, Counters& C) {
  if (same_pos(B, i, j)) return false;
  FE v = edge_val(B.pf(i), B.wf(i), B.pf(j), B.wf(j), fe(B.alpha), fe(4.0),
                  (FE*)nullptr);
  int s = fe_sign(v);
  if (s != UNDECIDED) return s > 0;
  C.exact_filter++;
  typedef Big<L> I;
  I vb = edge_val(B.pb(i), B.W[i], B.pb(j), B.W[j], B.AL, I::from_int(4),
                  (I*)nullptr);
  return vb.sign() >= 0;
}

template <class F, int L>
bool tri_keep(const Balls<F, L>& B, uint32_t i, uint32_t j, uint32_t k,
              Counters& C) {
  FE nn;
  FE v = tri_val(B.pf(i), B.wf(i), B.pf(j), B.wf(j), B.pf(k), B.wf(k),
                 fe(B.alpha), fe(2.0), fe(4.0), &nn);
  int sn = fe_sign(nn), s = fe_sign(v);
  if (sn == 1 && s != UNDECIDED) return s > 0;
  C.exact_filter++;
  typedef Big<L> I;
  I nb;
  I vb = tri_val(B.pb(i), B.W[i], B.pb(j), B.W[j], B.pb(k), B.W[k], B.AL,
                 I::from_int(2), I::from_int(4), &nb);
  if (nb.sign() == 0) return false;
  return vb.sign() >= 0;
}

template <class F, int L>
bool tet_keep(const Balls<F, L>& B, uint32_t i, uint32_t j, uint32_t k,
              uint32_t l, Counters& C) {
  FE D;
  FE v = tet_val(B.pf(i), B.wf(i), B.pf(j), B.wf(j), B.pf(k), B.wf(k),
                 B.pf(l), B.wf(l), fe(B.alpha), fe(4.0), &D, (V3<FE>*)nullptr);
  int sd = fe_sign(D), s = fe_sign(v);
  if (sd != UNDECIDED && s != UNDECIDED) return s > 0;
  C.exact_filter++;
  typedef Big<L> I;
  I Db;
  I vb = tet_val(B.pb(i), B.W[i], B.pb(j), B.W[j], B.pb(k), B.W[k], B.pb(l),
                 B.W[l], B.AL, I::from_int(4), &Db, (V3<I>*)nullptr);
  if (Db.sign() == 0) return false;
  return vb.sign() >= 0;
}

/* cell range around a point for the AC2 scans */
template <class F, int L>
void cell_range(const Balls<F, L>& B, double cx, double cy, double cz, int k,
                int lo[3], int hi[3]) {
  double q[3] = {((double)cx - (double)B.minx) / (double)B.step,
                 ((double)cy - (double)B.miny) / (double)B.step,
                 ((double)cz - (double)B.minz) / (double)B.step};
  int ext[3] = {B.ex, B.ey, B.ez};
  for (int a = 0; a < 3; ++a) {
    double f = std::floor(q[a]);
    if (!(f == f)) f = 0;
    if (f < -1e9) f = -1e9;
    if (f > 1e9) f = 1e9;
    long long c = (long long)f;
    long long l0 = c - k, h0 = c + k;
    if (l0 < 0) l0 = 0;
    if (h0 > ext[a] - 1) h0 = ext[a] - 1;
    lo[a] = (int)l0;
    hi[a] = (int)h0;
  }
}

template <class F, int L>
bool tri_ac2_pass(const Balls<F, L>& B, uint32_t i, uint32_t j, uint32_t k,
                  Counters& C) {
  typedef Big<L> I;
  V3<FE> pi = B.pf(i), pj = B.pf(j), pk = B.pf(k);
  FE wi = B.wf(i), wj = B.wf(j), wk = B.wf(k);
  V3<FE> u = vsub(pj, pi), v = vsub(pk, pi), n = vcross(u, v);
  FE nn = vdot(n, n);
  FE A = vdot(u, u) + wi - wj, Bq = vdot(v, v) + wi - wk;
  V3<FE> Xv = vsub(vscale(A, vcross(v, n)), vscale(Bq, vcross(u, n)));
  /* approximate orthocentre for the cell search */
  double cx = pi.x.v + Xv.x.v / (2 * nn.v), cy = pi.y.v + Xv.y.v / (2 * nn.v),
         cz = pi.z.v + Xv.z.v / (2 * nn.v);
  int lo[3], hi[3];
  cell_range(B, cx, cy, cz, B.k_ac2, lo, hi);
  bool have_exact = false;
  V3<I> piE, XE;
  I nnE, wiE;
  int64_t npairs = 0;
  bool pass = true;
  for (int gz = lo[2]; gz <= hi[2] && pass; ++gz)
    for (int gy = lo[1]; gy <= hi[1] && pass; ++gy)
      for (int gx = lo[0]; gx <= hi[0] && pass; ++gx) {
        int64_t c = gx + (int64_t)B.ex * (gy + (int64_t)B.ey * gz);
        for (uint32_t e = B.cstart[c]; e < B.cstart[c + 1]; ++e) {
          if (e == i || e == j || e == k) continue;
          ++npairs;
          V3<FE> y = vsub(B.pf(e), pi);
          FE Le = vdot(y, y) - B.wf(e) + wi;
          FE val = Le * nn - vdot(y, Xv);
          int s = fe_sign(val);
          if (s == UNDECIDED) {
            C.exact_ac2++;
            if (!have_exact) {
              piE = B.pb(i);
              V3<I> uE = vsub(B.pb(j), piE), vE = vsub(B.pb(k), piE);
              V3<I> nE = vcross(uE, vE);
              nnE = vdot(nE, nE);
              wiE = B.W[i];
              I AE = vdot(uE, uE) + wiE - B.W[j];
              I BE = vdot(vE, vE) + wiE - B.W[k];
              XE = vsub(vscale(AE, vcross(vE, nE)), vscale(BE, vcross(uE, nE)));
              have_exact = true;
            }
            V3<I> yE = vsub(B.pb(e), piE);
            I LeE = vdot(yE, yE) - B.W[e] + wiE;
            s = (LeE * nnE - vdot(yE, XE)).sign();
          }
          if (s <= 0) {
            pass = false;
            break;
          }
        }
      }
  C.pairs += npairs;
  return pass;
}

template <class F, int L>
bool edge_ac2_pass(const Balls<F, L>& B, uint32_t i, uint32_t j, Counters& C) {
  typedef Big<L> I;
  V3<FE> pi = B.pf(i), pj = B.pf(j);
  FE wi = B.wf(i), wj = B.wf(j);
  V3<FE> u = vsub(pj, pi);
  FE uu = vdot(u, u);
  FE A = uu + wi - wj;
  double f = A.v / (2 * uu.v);
  double cx = pi.x.v + u.x.v * f, cy = pi.y.v + u.y.v * f,
         cz = pi.z.v + u.z.v * f;
  int lo[3], hi[3];
  cell_range(B, cx, cy, cz, B.k_ac2, lo, hi);
  bool have_exact = false;
  V3<I> piE, uE;
  I uuE, AE, wiE;
  int64_t npairs = 0;
  bool pass = true;
  for (int gz = lo[2]; gz <= hi[2] && pass; ++gz)
    for (int gy = lo[1]; gy <= hi[1] && pass; ++gy)
      for (int gx = lo[0]; gx <= hi[0] && pass; ++gx) {
        int64_t c = gx + (int64_t)B.ex * (gy + (int64_t)B.ey * gz);
        for (uint32_t e = B.cstart[c]; e < B.cstart[c + 1]; ++e) {
          if (e == i || e == j) continue;
          ++npairs;
          V3<FE> y = vsub(B.pf(e), pi);
          FE Le = vdot(y, y) - B.wf(e) + wi;
          FE val = Le * uu - vdot(y, u) * A;
          int s = fe_sign(val);
          if (s == UNDECIDED) {
            C.exact_ac2++;
            if (!have_exact) {
              piE = B.pb(i);
              uE = vsub(B.pb(j), piE);
              uuE = vdot(uE, uE);
              wiE = B.W[i];
              AE = uuE + wiE - B.W[j];
              have_exact = true;
            }
            V3<I> yE = vsub(B.pb(e), piE);
            I LeE = vdot(yE, yE) - B.W[e] + wiE;
            s = (LeE * uuE - vdot(yE, uE) * AE).sign();
          }
          if (s <= 0) {
            pass = false;
            break;
          }
        }
      }
  C.pairs += npairs;
  return pass;
}

/* exact lifted rows {X, Y, Z, X^2+Y^2+Z^2 - W} */
template <class F, int L>
void lifted_row(const Balls<F, L>& B, uint32_t i, Big<L> out[4]) {
  typedef Big<L> I;
  out[0] = I::from_i128(B.X[i]);
  out[1] = I::from_i128(B.Y[i]);
  out[2] = I::from_i128(B.Z[i]);
  out[3] = out[0] * out[0] + out[1] * out[1] + out[2] * out[2] - B.W[i];
}

template <class F, int L>
bool tet_ac2_pass(const Balls<F, L>& B, const uint32_t v[4], Counters& C) {
  typedef Big<L> I;
  uint32_t i = v[0];
  V3<FE> pi = B.pf(i);
  FE wi = B.wf(i);
  FE D;
  V3<FE> N;
  tet_val(pi, wi, B.pf(v[1]), B.wf(v[1]), B.pf(v[2]), B.wf(v[2]), B.pf(v[3]),
          B.wf(v[3]), fe(B.alpha), fe(4.0), &D, &N);
  int sD = fe_sign(D);
  double cx = pi.x.v + N.x.v / (2 * D.v), cy = pi.y.v + N.y.v / (2 * D.v),
         cz = pi.z.v + N.z.v / (2 * D.v);
  int lo[3], hi[3];
  cell_range(B, cx, cy, cz, B.k_ac2, lo, hi);
  bool have_exact = false;
  V3<I> piE, NE;
  I DE, wiE;
  auto need_exact = [&]() {
    if (have_exact) return;
    piE = B.pb(i);
    wiE = B.W[i];
    tet_val(piE, wiE, B.pb(v[1]), B.W[v[1]], B.pb(v[2]), B.W[v[2]],
            B.pb(v[3]), B.W[v[3]], B.AL, I::from_int(4), &DE, &NE);
    have_exact = true;
  };
  if (sD == UNDECIDED) {
    need_exact();
    sD = DE.sign();
  }
  int64_t npairs = 0;
  bool pass = true;
  for (int gz = lo[2]; gz <= hi[2] && pass; ++gz)
    for (int gy = lo[1]; gy <= hi[1] && pass; ++gy)
      for (int gx = lo[0]; gx <= hi[0] && pass; ++gx) {
        int64_t c = gx + (int64_t)B.ex * (gy + (int64_t)B.ey * gz);
        for (uint32_t e = B.cstart[c]; e < B.cstart[c + 1]; ++e) {
          if (e == v[0] || e == v[1] || e == v[2] || e == v[3]) continue;
          ++npairs;
          V3<FE> y = vsub(B.pf(e), pi);
          FE Le = vdot(y, y) - B.wf(e) + wi;
          FE val = Le * D - vdot(y, N);
          int s = fe_sign(val);
          int viol;
          if (s != UNDECIDED) {
            viol = (s * sD) < 0; /* pi_e - rho < 0 */
          } else {
            C.exact_ac2++;
            need_exact();
            V3<I> yE = vsub(B.pb(e), piE);
            I LeE = vdot(yE, yE) - B.W[e] + wiE;
            int se = (LeE * DE - vdot(yE, NE)).sign();
            if (se != 0) {
              viol = (se * sD) < 0;
            } else {
              /* tie: SoS of insphere_w = det[x y z l 1](a,b,c,d,e),
               * orient3d(a,b,c,d) = -sign(D); violated iff SoS*orient > 0
               * (AlphaComplexParallel.cu:653-683). */
              C.sos++;
              I P[5][4];
              int64_t ids[5];
              for (int q = 0; q < 4; ++q) {
                lifted_row(B, v[q], P[q]);
                ids[q] = B.orig[v[q]];
              }
              lifted_row(B, e, P[4]);
              ids[4] = B.orig[e];
              int ss = sos_det5<L>(P, ids, nullptr);
              int orient = -sD;
              viol = (ss * orient) > 0;
            }
          }
          if (viol) {
            pass = false;
            break;
          }
        }
      }
  C.pairs += npairs;
  return pass;
}

/* ------------------------------------------------------------------ */
/* Parallel helpers: deterministic chunked map with ordered concat.     */
/* ------------------------------------------------------------------ */
struct Chunk {
  std::vector<uint32_t> vals;   /* emitted values */
  std::vector<uint32_t> cnt;    /* per item counts */
  int64_t cand = 0;
};

template <class Fn>
void chunked(int64_t nitems, int64_t chunk, int nthreads,
             std::vector<Chunk>& out, Fn fn) {
  int64_t nch = (nitems + chunk - 1) / chunk;
  out.assign((size_t)nch, Chunk());
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (int64_t c = 0; c < nch; ++c) {
    int64_t b = c * chunk, e = std::min(nitems, b + chunk);
    fn(b, e, out[(size_t)c]);
  }
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

}  // namespace

/* ------------------------------------------------------------------ */
struct pac_oracle_result {
  std::vector<int32_t> edges, tris, tets; /* original ids, reference order */
  std::vector<int32_t> perm;
  int64_t stats[17] = {0};
  double times[8] = {0};
};

namespace {

enum { PAC_OK = 0, PAC_ERR_INVALID = -2, PAC_ERR_RANGE = -7, PAC_ERR_NOMEM = -4 };

template <class F, int L>
int run_pipeline(const F* xyzr, int64_t n, F solvent, F alpha, int skip_tet,
                 int policy, int nthreads, pac_oracle_result* R) {
  typedef Big<L> I;
  if (nthreads <= 0) {
#ifdef _OPENMP
    nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
  }
  Balls<F, L> B;
  B.n = n;
  R->stats[0] = n;
  if (n == 0) return PAC_OK;
  double t0 = now_ms();
  /* ---- grid (AlphaComplexParallel.cu:1082-1124), input float type ---- */
  std::vector<F> rx(n), ry(n), rz(n), rr(n);
  F minX = 0, minY = 0, minZ = 0, maxX = 0, maxY = 0, maxZ = 0, maxR = 0;
  for (int64_t i = 0; i < n; ++i) {
    F x = xyzr[4 * i], y = xyzr[4 * i + 1], z = xyzr[4 * i + 2],
      r = xyzr[4 * i + 3];
    if (!std::isfinite(x) || !std::isfinite(y) || !std::isfinite(z) ||
        !std::isfinite(r))
      return PAC_ERR_INVALID;
    rx[i] = x;
    ry[i] = y;
    rz[i] = z;
    volatile F rs = r + solvent; /* fl(r_raw + s) */
    rr[i] = rs;
    if (i == 0) {
      minX = maxX = x;
      minY = maxY = y;
      minZ = maxZ = z;
      maxR = r;
    } else {
      minX = std::min(minX, x);
      maxX = std::max(maxX, x);
      minY = std::min(minY, y);
      maxY = std::max(maxY, y);
      minZ = std::min(minZ, z);
      maxZ = std::max(maxZ, z);
      maxR = std::max(maxR, r);
    }
  }
  /* The cell key is the reference's FP32 computation (LD:65-78).  FP64
   * balls are keyed by their FP32 roundings: the grid values and the key
   * arithmetic are FP32 for both input types. */
  const float fminX = (float)minX, fminY = (float)minY, fminZ = (float)minZ;
  const float fmaxR = (float)maxR;
  volatile float m2 = fmaxR * fmaxR;
  volatile float m2a = m2 + (float)alpha;
  float step = 2.0f * std::sqrt((float)m2a);
  if (!(step > 0) || !std::isfinite(step)) return PAC_ERR_INVALID;
  volatile float qx = ((float)maxX - fminX) / step,
                 qy = ((float)maxY - fminY) / step,
                 qz = ((float)maxZ - fminZ) / step;
  int64_t ex = (int64_t)(float)qx + 1, ey = (int64_t)(float)qy + 1,
          ez = (int64_t)(float)qz + 1;
  if (ex * ey * ez > (int64_t)1 << 31) return PAC_ERR_RANGE;
  B.minx = fminX;
  B.miny = fminY;
  B.minz = fminZ;
  B.step = step;
  B.ex = (int)ex;
  B.ey = (int)ey;
  B.ez = (int)ez;
  std::vector<uint32_t> key(n);
  double maxabs = 0;
  for (int64_t i = 0; i < n; ++i) {
    volatile float dx = (float)rx[i] - fminX, dy = (float)ry[i] - fminY,
                   dz = (float)rz[i] - fminZ;
    volatile float fx = dx / step, fy = dy / step, fz = dz / step;
    int cx = (int)(float)fx, cy = (int)(float)fy, cz = (int)(float)fz;
    key[i] = (uint32_t)(cx + ex * cy + ex * ey * cz);
    maxabs = std::max({maxabs, std::fabs((double)rx[i]), std::fabs((double)ry[i]),
                       std::fabs((double)rz[i])});
  }
  std::vector<int64_t> perm(n);
  for (int64_t i = 0; i < n; ++i) perm[i] = i;
  std::stable_sort(perm.begin(), perm.end(),
                   [&](int64_t a, int64_t b) { return key[a] < key[b]; });
  B.x.resize(n);
  B.y.resize(n);
  B.z.resize(n);
  B.r.resize(n);
  B.w.resize(n);
  B.orig = perm;
  B.cell.resize(n);
  for (int64_t s = 0; s < n; ++s) {
    int64_t o = perm[s];
    B.x[s] = rx[o];
    B.y[s] = ry[o];
    B.z[s] = rz[o];
    B.r[s] = rr[o];
    B.w[s] = (double)rr[o] * (double)rr[o];
    B.cell[s] = key[o];
  }
  int64_t G = ex * ey * ez;
  B.cstart.assign((size_t)G + 1, 0);
  for (int64_t s = 0; s < n; ++s) B.cstart[B.cell[s] + 1]++;
  for (int64_t c = 0; c < G; ++c) B.cstart[c + 1] += B.cstart[c];
  R->perm.resize(n);
  for (int64_t s = 0; s < n; ++s) R->perm[s] = (int32_t)perm[s];

  /* ---- stencil ---- */
  double maxw = 0;
  for (int64_t s = 0; s < n; ++s) maxw = std::max(maxw, B.w[s]);
  B.alpha = (double)alpha;
  double reach2 = maxw + (double)alpha;
  if (reach2 < 0) B.no_edges = true;
  if (policy == 1) {
    B.k_enum = B.k_ac2 = 1;
  } else {
    double dmax = reach2 > 0 ? 2.0 * std::sqrt(reach2) * (1 + 1e-12) : 0.0;
    /* key rounding in cells: FP32 subtract + divide, and for FP64 balls
     * the rounding of the coordinate to FP32 */
    double extm = (double)std::max(ex, std::max(ey, ez)) + 2;
    if (!std::is_same<F, float>::value) extm += maxabs / (double)step;
    double margin = 8.0 * extm * 0x1p-24 + 1e-9;
    B.k_enum = (int)std::ceil(dmax / (double)step + margin);
    B.k_ac2 = (int)std::ceil(0.5 * dmax / (double)step + margin);
    if (B.k_enum < 1) B.k_enum = 1;
    if (B.k_ac2 < 1) B.k_ac2 = 1;
  }
  R->stats[14] = B.k_enum;
  R->stats[15] = B.k_ac2;

  /* ---- exact embedding ---- */
  int S = 0;
  auto upd = [&](double v, int mult) {
    if (v == 0) return;
    int e = lsb_exponent(v);
    int need = (-e + mult - 1) / mult;
    if (-e > 0 && need > S) S = need;
  };
  for (int64_t s = 0; s < n; ++s) {
    upd((double)B.x[s], 1);
    upd((double)B.y[s], 1);
    upd((double)B.z[s], 1);
    upd((double)B.r[s], 1);
  }
  upd((double)alpha, 2);
  B.S = S;
  R->stats[16] = S;
  double maxr = 0;
  for (int64_t s = 0; s < n; ++s) {
    maxabs = std::max(maxabs, std::fabs((double)B.x[s]));
    maxabs = std::max(maxabs, std::fabs((double)B.y[s]));
    maxabs = std::max(maxabs, std::fabs((double)B.z[s]));
    maxr = std::max(maxr, std::fabs((double)B.r[s]));
  }
  double bits = std::log2(std::max({maxabs, maxr, std::sqrt(std::fabs((double)alpha)), 1.0})) + S + 2;
  if (bits > 120 || 8 * bits + 16 > 64.0 * L - 2) return PAC_ERR_RANGE;
  B.X.resize(n);
  B.Y.resize(n);
  B.Z.resize(n);
  B.W.resize(n);
  for (int64_t s = 0; s < n; ++s) {
    B.X[s] = (i128)std::ldexp((double)B.x[s], S);
    B.Y[s] = (i128)std::ldexp((double)B.y[s], S);
    B.Z[s] = (i128)std::ldexp((double)B.z[s], S);
    I Rb = I::from_i128((i128)std::ldexp((double)B.r[s], S));
    B.W[s] = Rb * Rb;
  }
  B.AL = big_from_scaled<L>((double)alpha, 2 * S);
  double t1 = now_ms();
  R->times[1] = t1 - t0; /* Grid computation */

  Counters C;
  /* ---- edges (LD:80-190) ---- */
  std::vector<Chunk> ch;
  std::vector<uint32_t> erow(n + 1, 0), ecol;
  int64_t Ecand = 0;
  if (!B.no_edges) {
    int k = B.k_enum;
    chunked(n, 256, nthreads, ch, [&](int64_t b, int64_t e, Chunk& out) {
      for (int64_t i = b; i < e; ++i) {
        uint32_t c = B.cell[i];
        int gx = (int)(c % (uint32_t)B.ex);
        int gy = (int)((c / (uint32_t)B.ex) % (uint32_t)B.ey);
        int gz = (int)(c / ((uint32_t)B.ex * (uint32_t)B.ey));
        size_t before = out.vals.size();
        for (int dz = -k; dz <= k; ++dz) {
          int z = gz + dz;
          if (z < 0 || z >= B.ez) continue;
          for (int dy = -k; dy <= k; ++dy) {
            int y = gy + dy;
            if (y < 0 || y >= B.ey) continue;
            for (int dx = -k; dx <= k; ++dx) {
              int x = gx + dx;
              if (x < 0 || x >= B.ex) continue;
              int64_t cc = x + (int64_t)B.ex * (y + (int64_t)B.ey * z);
              if ((uint32_t)cc < c) continue;
              for (uint32_t j = B.cstart[cc]; j < B.cstart[cc + 1]; ++j) {
                if (j <= (uint32_t)i) continue;
                out.cand++;
                if (edge_keep(B, (uint32_t)i, j, C)) out.vals.push_back(j);
              }
            }
          }
        }
        std::sort(out.vals.begin() + before, out.vals.end());
        out.cnt.push_back((uint32_t)(out.vals.size() - before));
      }
    });
    int64_t row = 0;
    for (auto& c : ch) {
      Ecand += c.cand;
      size_t p = 0;
      for (uint32_t cnt : c.cnt) {
        erow[row + 1] = erow[row] + cnt;
        ++row;
      }
      ecol.insert(ecol.end(), c.vals.begin(), c.vals.end());
      (void)p;
    }
  }
  int64_t Ealpha = (int64_t)ecol.size();
  std::vector<uint32_t> eleft(Ealpha);
  for (int64_t i = 0; i < n; ++i)
    for (uint32_t q = erow[i]; q < erow[i + 1]; ++q) eleft[q] = (uint32_t)i;
  double t2 = now_ms();
  R->times[2] = t2 - t1;

  auto find_edge = [&](uint32_t a, uint32_t b) -> int64_t {
    auto lo = ecol.begin() + erow[a], hi = ecol.begin() + erow[a + 1];
    auto it = std::lower_bound(lo, hi, b);
    if (it == hi || *it != b) return -1;
    return it - ecol.begin();
  };

  /* ---- triangles (LD:192-348): per alpha edge, row(i) & row(j) ---- */
  chunked(Ealpha, 1024, nthreads, ch, [&](int64_t b, int64_t e, Chunk& out) {
    for (int64_t q = b; q < e; ++q) {
      uint32_t i = eleft[q], j = ecol[q];
      uint32_t a0 = erow[i], a1 = erow[i + 1], b0 = erow[j], b1 = erow[j + 1];
      size_t before = out.vals.size();
      while (a0 < a1 && b0 < b1) {
        uint32_t x = ecol[a0], y = ecol[b0];
        if (x < y)
          ++a0;
        else if (y < x)
          ++b0;
        else {
          out.cand++;
          if (tri_keep(B, i, j, x, C)) out.vals.push_back(x);
          ++a0;
          ++b0;
        }
      }
      out.cnt.push_back((uint32_t)(out.vals.size() - before));
    }
  });
  std::vector<uint32_t> trow(Ealpha + 1, 0), tk;
  int64_t Tcand = 0;
  {
    int64_t row = 0;
    for (auto& c : ch) {
      Tcand += c.cand;
      for (uint32_t cnt : c.cnt) {
        trow[row + 1] = trow[row] + cnt;
        ++row;
      }
      tk.insert(tk.end(), c.vals.begin(), c.vals.end());
    }
  }
  int64_t Talpha = (int64_t)tk.size();
  std::vector<uint32_t> tedge(Talpha);
  for (int64_t q = 0; q < Ealpha; ++q)
    for (uint32_t p = trow[q]; p < trow[q + 1]; ++p) tedge[p] = (uint32_t)q;
  auto find_tri = [&](uint32_t a, uint32_t b, uint32_t c) -> int64_t {
    int64_t e = find_edge(a, b);
    if (e < 0) return -1;
    auto lo = tk.begin() + trow[e], hi = tk.begin() + trow[e + 1];
    auto it = std::lower_bound(lo, hi, c);
    if (it == hi || *it != c) return -1;
    return it - tk.begin();
  };
  double t3 = now_ms();
  R->times[3] = t3 - t2;

  /* ---- tets (LD:350-591): per alpha triangle (i,j,k), l in T(i,j), l>k,
   *      with (i,k,l) and (j,k,l) alpha triangles ---- */
  chunked(Talpha, 1024, nthreads, ch, [&](int64_t b, int64_t e, Chunk& out) {
    for (int64_t q = b; q < e; ++q) {
      uint32_t eij = tedge[q];
      uint32_t i = eleft[eij], j = ecol[eij], k = tk[q];
      size_t before = out.vals.size();
      int64_t eik = -1, ejk = -1;
      for (uint32_t p = (uint32_t)q + 1; p < trow[eij + 1]; ++p) {
        uint32_t l = tk[p];
        if (eik < 0) {
          eik = find_edge(i, k);
          ejk = find_edge(j, k);
          if (eik < 0 || ejk < 0) break;
        }
        if (!std::binary_search(tk.begin() + trow[eik], tk.begin() + trow[eik + 1], l))
          continue;
        if (!std::binary_search(tk.begin() + trow[ejk], tk.begin() + trow[ejk + 1], l))
          continue;
        out.cand++;
        if (tet_keep(B, i, j, k, l, C)) out.vals.push_back(l);
      }
      out.cnt.push_back((uint32_t)(out.vals.size() - before));
    }
  });
  std::vector<uint32_t> tets; /* 4 per tet, sorted space */
  int64_t Kcand = 0;
  {
    int64_t q = 0;
    for (auto& c : ch) {
      Kcand += c.cand;
      size_t p = 0;
      for (uint32_t cnt : c.cnt) {
        uint32_t eij = tedge[q];
        for (uint32_t s = 0; s < cnt; ++s) {
          tets.push_back(eleft[eij]);
          tets.push_back(ecol[eij]);
          tets.push_back(tk[q]);
          tets.push_back(c.vals[p++]);
        }
        ++q;
      }
    }
  }
  int64_t Kalpha = (int64_t)tets.size() / 4;
  double t4 = now_ms();
  R->times[4] = t4 - t3;

  /* ---- tet AC2 (Parallel:611-749) ---- */
  std::vector<uint8_t> tet_ok(Kalpha, 1);
  if (!skip_tet) {
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
    for (int64_t q = 0; q < Kalpha; ++q)
      tet_ok[q] = tet_ac2_pass(B, &tets[4 * q], C) ? 1 : 0;
  }
  std::vector<uint32_t> ftets;
  for (int64_t q = 0; q < Kalpha; ++q)
    if (tet_ok[q]) ftets.insert(ftets.end(), &tets[4 * q], &tets[4 * q + 4]);
  int64_t Kfinal = (int64_t)ftets.size() / 4;
  double t5 = now_ms();
  R->times[5] = t5 - t4;

  /* ---- triangle closure + dangling AC2 (Parallel:1490-1632) ---- */
  std::vector<uint8_t> tri_in(Talpha, 0);
  for (int64_t q = 0; q < Kfinal; ++q) {
    const uint32_t* t = &ftets[4 * q];
    const uint32_t f[4][3] = {{t[0], t[1], t[2]}, {t[0], t[1], t[3]},
                              {t[0], t[2], t[3]}, {t[1], t[2], t[3]}};
    for (auto& ff : f) {
      int64_t id = find_tri(ff[0], ff[1], ff[2]);
      if (id >= 0) tri_in[id] = 1;
    }
  }
  std::vector<uint8_t> tri_ok(Talpha, 0);
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
  for (int64_t q = 0; q < Talpha; ++q) {
    if (tri_in[q]) {
      tri_ok[q] = 1;
    } else {
      uint32_t e = tedge[q];
      tri_ok[q] = tri_ac2_pass(B, eleft[e], ecol[e], tk[q], C) ? 1 : 0;
    }
  }
  std::vector<uint32_t> ftris;
  for (int64_t q = 0; q < Talpha; ++q)
    if (tri_ok[q]) {
      uint32_t e = tedge[q];
      ftris.push_back(eleft[e]);
      ftris.push_back(ecol[e]);
      ftris.push_back(tk[q]);
    }
  int64_t Tfinal = (int64_t)ftris.size() / 3;
  double t6 = now_ms();
  R->times[6] = t6 - t5;

  /* ---- edge closure + dangling AC2 (Parallel:1646-1788) ---- */
  std::vector<uint8_t> edge_in(Ealpha, 0);
  for (int64_t q = 0; q < Tfinal; ++q) {
    const uint32_t* t = &ftris[3 * q];
    int64_t a = find_edge(t[0], t[1]), b = find_edge(t[0], t[2]),
            c = find_edge(t[1], t[2]);
    if (a >= 0) edge_in[a] = 1;
    if (b >= 0) edge_in[b] = 1;
    if (c >= 0) edge_in[c] = 1;
  }
  std::vector<uint8_t> edge_ok(Ealpha, 0);
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
  for (int64_t q = 0; q < Ealpha; ++q) {
    edge_ok[q] = edge_in[q] ? 1 : (edge_ac2_pass(B, eleft[q], ecol[q], C) ? 1 : 0);
  }
  double t7 = now_ms();
  R->times[7] = t7 - t6;

  /* ---- outputs: original ids, reference order ---- */
  for (int64_t q = 0; q < Ealpha; ++q)
    if (edge_ok[q]) {
      R->edges.push_back((int32_t)perm[eleft[q]]);
      R->edges.push_back((int32_t)perm[ecol[q]]);
    }
  for (uint32_t v : ftris) R->tris.push_back((int32_t)perm[v]);
  for (uint32_t v : ftets) R->tets.push_back((int32_t)perm[v]);
  R->stats[1] = Ecand;
  R->stats[2] = Ealpha;
  R->stats[3] = Tcand;
  R->stats[4] = Talpha;
  R->stats[5] = Kcand;
  R->stats[6] = Kalpha;
  R->stats[7] = Kfinal;
  R->stats[8] = Tfinal;
  R->stats[9] = (int64_t)R->edges.size() / 2;
  R->stats[10] = C.exact_filter;
  R->stats[11] = C.exact_ac2;
  R->stats[12] = C.sos;
  R->stats[13] = C.pairs;
  return PAC_OK;
}

/* single-predicate helpers: build a tiny embedding for the given balls */
template <int L>
struct Tiny {
  int S = 0;
  void add(double v, int mult) {
    if (v == 0) return;
    int e = lsb_exponent(v);
    int need = (-e + mult - 1) / mult;
    if (-e > 0 && need > S) S = need;
  }
  V3<Big<L>> p(const float* b) const {
    return V3<Big<L>>{big_from_scaled<L>(b[0], S), big_from_scaled<L>(b[1], S),
                      big_from_scaled<L>(b[2], S)};
  }
  Big<L> w(const float* b) const {
    Big<L> r = big_from_scaled<L>(b[3], S);
    return r * r;
  }
  void lifted(const float* b, Big<L> out[4]) const {
    V3<Big<L>> q = p(b);
    out[0] = q.x;
    out[1] = q.y;
    out[2] = q.z;
    out[3] = q.x * q.x + q.y * q.y + q.z * q.z - w(b);
  }
};
typedef Big<8> B8;
Tiny<8> tiny_of(std::initializer_list<const float*> bs, float alpha, bool use_alpha) {
  Tiny<8> t;
  for (const float* b : bs)
    for (int q = 0; q < 4; ++q) t.add(b[q], 1);
  if (use_alpha) t.add(alpha, 2);
  return t;
}

}  // namespace

extern "C" {

int pac_oracle_run_f32(const float* xyzr, int64_t n, float solvent, float alpha,
                       int skip_tet_ac2, int cell_policy, int nthreads,
                       pac_oracle_result** out) {
  if (!out || (n > 0 && !xyzr) || n < 0 || n > (int64_t)INT32_MAX)
    return PAC_ERR_INVALID;
  pac_oracle_result* R = new (std::nothrow) pac_oracle_result();
  if (!R) return PAC_ERR_NOMEM;
  int rc = run_pipeline<float, 8>(xyzr, n, solvent, alpha, skip_tet_ac2,
                                  cell_policy, nthreads, R);
  if (rc) {
    delete R;
    return rc;
  }
  *out = R;
  return PAC_OK;
}

int pac_oracle_run_f64(const double* xyzr, int64_t n, double solvent,
                       double alpha, int skip_tet_ac2, int cell_policy,
                       int nthreads, pac_oracle_result** out) {
  if (!out || (n > 0 && !xyzr) || n < 0 || n > (int64_t)INT32_MAX)
    return PAC_ERR_INVALID;
  pac_oracle_result* R = new (std::nothrow) pac_oracle_result();
  if (!R) return PAC_ERR_NOMEM;
  int rc = run_pipeline<double, 16>(xyzr, n, solvent, alpha, skip_tet_ac2,
                                    cell_policy, nthreads, R);
  if (rc) {
    delete R;
    return rc;
  }
  *out = R;
  return PAC_OK;
}

int64_t pac_oracle_count(const pac_oracle_result* r, int dim) {
  if (!r) return -1;
  if (dim == 1) return (int64_t)r->edges.size() / 2;
  if (dim == 2) return (int64_t)r->tris.size() / 3;
  if (dim == 3) return (int64_t)r->tets.size() / 4;
  return -1;
}
void pac_oracle_copy(const pac_oracle_result* r, int dim, int32_t* dst) {
  const std::vector<int32_t>* v =
      dim == 1 ? &r->edges : dim == 2 ? &r->tris : &r->tets;
  if (!v->empty()) std::memcpy(dst, v->data(), v->size() * sizeof(int32_t));
}
void pac_oracle_perm(const pac_oracle_result* r, int32_t* dst) {
  if (!r->perm.empty())
    std::memcpy(dst, r->perm.data(), r->perm.size() * sizeof(int32_t));
}
void pac_oracle_stats(const pac_oracle_result* r, int64_t* out) {
  std::memcpy(out, r->stats, sizeof(r->stats));
}
void pac_oracle_times(const pac_oracle_result* r, double* out) {
  std::memcpy(out, r->times, sizeof(r->times));
}
void pac_oracle_free(pac_oracle_result* r) { delete r; }

int pac_oracle_edge_keep(const float* b0, const float* b1, float alpha) {
  if (b0[0] == b1[0] && b0[1] == b1[1] && b0[2] == b1[2]) return 0;
  Tiny<8> t = tiny_of({b0, b1}, alpha, true);
  B8 v = edge_val(t.p(b0), t.w(b0), t.p(b1), t.w(b1),
                  big_from_scaled<8>(alpha, 2 * t.S), B8::from_int(4), (B8*)nullptr);
  return v.sign() >= 0;
}
int pac_oracle_tri_keep(const float* b0, const float* b1, const float* b2,
                        float alpha) {
  Tiny<8> t = tiny_of({b0, b1, b2}, alpha, true);
  B8 nn;
  B8 v = tri_val(t.p(b0), t.w(b0), t.p(b1), t.w(b1), t.p(b2), t.w(b2),
                 big_from_scaled<8>(alpha, 2 * t.S), B8::from_int(2),
                 B8::from_int(4), &nn);
  if (nn.sign() == 0) return 0;
  return v.sign() >= 0;
}
int pac_oracle_tet_keep(const float* b0, const float* b1, const float* b2,
                        const float* b3, float alpha) {
  Tiny<8> t = tiny_of({b0, b1, b2, b3}, alpha, true);
  B8 D;
  B8 v = tet_val(t.p(b0), t.w(b0), t.p(b1), t.w(b1), t.p(b2), t.w(b2), t.p(b3),
                 t.w(b3), big_from_scaled<8>(alpha, 2 * t.S), B8::from_int(4),
                 &D, (V3<B8>*)nullptr);
  if (D.sign() == 0) return 0;
  return v.sign() >= 0;
}
int pac_oracle_tri_violated(const float* b0, const float* b1, const float* b2,
                            const float* e) {
  Tiny<8> t = tiny_of({b0, b1, b2, e}, 0, false);
  V3<B8> pi = t.p(b0), u = vsub(t.p(b1), pi), v = vsub(t.p(b2), pi);
  V3<B8> n = vcross(u, v);
  B8 nn = vdot(n, n), wi = t.w(b0);
  B8 A = vdot(u, u) + wi - t.w(b1), Bq = vdot(v, v) + wi - t.w(b2);
  V3<B8> X = vsub(vscale(A, vcross(v, n)), vscale(Bq, vcross(u, n)));
  V3<B8> y = vsub(t.p(e), pi);
  B8 Le = vdot(y, y) - t.w(e) + wi;
  return (Le * nn - vdot(y, X)).sign() <= 0;
}
int pac_oracle_edge_violated(const float* b0, const float* b1, const float* e) {
  Tiny<8> t = tiny_of({b0, b1, e}, 0, false);
  V3<B8> pi = t.p(b0), u = vsub(t.p(b1), pi);
  B8 uu = vdot(u, u), wi = t.w(b0);
  B8 A = uu + wi - t.w(b1);
  V3<B8> y = vsub(t.p(e), pi);
  B8 Le = vdot(y, y) - t.w(e) + wi;
  return (Le * uu - vdot(y, u) * A).sign() <= 0;
}
int pac_oracle_tet_violated(const float* const* b4, const float* e,
                            const int64_t* ids5) {
  Tiny<8> t = tiny_of({b4[0], b4[1], b4[2], b4[3], e}, 0, false);
  B8 D;
  V3<B8> N;
  V3<B8> pi = t.p(b4[0]);
  B8 wi = t.w(b4[0]);
  tet_val(pi, wi, t.p(b4[1]), t.w(b4[1]), t.p(b4[2]), t.w(b4[2]), t.p(b4[3]),
          t.w(b4[3]), B8(), B8::from_int(4), &D, &N);
  int sD = D.sign();
  if (sD == 0) return -1;
  V3<B8> y = vsub(t.p(e), pi);
  B8 Le = vdot(y, y) - t.w(e) + wi;
  int s = (Le * D - vdot(y, N)).sign();
  if (s != 0) return (s * sD) < 0;
  B8 P[5][4];
  for (int q = 0; q < 4; ++q) t.lifted(b4[q], P[q]);
  t.lifted(e, P[4]);
  int ss = sos_det5<8>(P, ids5, nullptr);
  return (ss * -sD) > 0;
}
int pac_oracle_sos_det5(const float* const* b5, const int64_t* ids5) {
  Tiny<8> t = tiny_of({b5[0], b5[1], b5[2], b5[3], b5[4]}, 0, false);
  B8 P[5][4];
  for (int q = 0; q < 5; ++q) t.lifted(b5[q], P[q]);
  return sos_det5<8>(P, ids5, nullptr);
}
int pac_oracle_det5(const float* const* b5) {
  Tiny<8> t = tiny_of({b5[0], b5[1], b5[2], b5[3], b5[4]}, 0, false);
  B8 M[5][5];
  for (int q = 0; q < 5; ++q) {
    B8 row[4];
    t.lifted(b5[q], row);
    for (int c = 0; c < 4; ++c) M[q][c] = row[c];
    M[q][4] = B8::from_int(1);
  }
  int rows[5] = {0, 1, 2, 3, 4}, cols[5] = {0, 1, 2, 3, 4};
  return minor_det(M, rows, cols, 5).sign();
}
int pac_oracle_orient3d(const float* a, const float* b, const float* c,
                        const float* d) {
  Tiny<8> t = tiny_of({a, b, c, d}, 0, false);
  V3<B8> pd = t.p(d);
  V3<B8> ad = vsub(t.p(a), pd), bd = vsub(t.p(b), pd), cd = vsub(t.p(c), pd);
  return vdot(ad, vcross(bd, cd)).sign();
}
int pac_oracle_sos_table(int32_t* out3, int cap) {
  const auto& t = sos_terms();
  for (int q = 0; q < (int)t.size() && q < cap; ++q) {
    out3[3 * q] = t[q].rows;
    out3[3 * q + 1] = t[q].cols;
    out3[3 * q + 2] = t[q].sign;
  }
  return (int)t.size();
}

}  // extern "C"
