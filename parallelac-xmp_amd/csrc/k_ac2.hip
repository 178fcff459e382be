/*
 * k_ac2.hip -- emptiness ("AC2") tests of alpha tetrahedra, dangling
 * triangles and dangling edges.
 *
 * Hot kernels (one thread per simplex) compute the orthocentre once with a
 * certified error, build the cell box that must contain every possible
 * violator, and scan it centre cell first with the semi-static power-gap
 * test in one precision (FP32 staged / FP64 plain).  The first certain
 * violator ends the scan; undecided (simplex, ball) pairs, or simplices
 * whose centre could not be certified, are queued.  The escalation kernel
 * retries them in FP64 then double-double, and the exact kernel settles the
 * rest in 512-bit integers, breaking exact tet ties by Edelsbrunner-Muecke
 * simulation of simplicity over the original input order.
 *
 * Reference: checkTetsOrthosphere AlphaComplexLargeData.cu:597-735 (with
 * orient3d / insphere_w / orientation4SoS_w, predicates.h:2015, :4250,
 * :4669), checkDanglingTris :807-881, checkDanglingEdges :928-1015.
 */
#include "pac_dev.h"
#include "pac_wide.h"
#include "pac_sos_table.h"

namespace pac {
namespace PAC_V {

/* Exact stage: every lane of a group of kTPI lanes evaluates the same
 * item over the wavefront-cooperative integers (pac_wide.h). */
using XI = CW<kTPI>;

/* Sign of the SoS-perturbed det[x y z l 1] of the five balls (l = |p|^2 -
 * w), rows in argument order, perturbation order given by ids (original
 * input ids, as the reference's Atom.index, AlphaComplexParallel.cu:1087).
 * Every term of the expansion (pac_sos_table) is a minor that keeps the
 * column of ones, and such an m x m minor equals the (m-1) x (m-1)
 * determinant of its other columns with the last row subtracted from the
 * rest (subtracting a row changes no determinant; expanding along the
 * ones column, now (0, ..., 0, 1), leaves that block with sign +1).  The
 * lifted matrix lives in LDS (lm: this wave's 20 x 64 words, entry (r, c)
 * at (4 r + c) * 64 + lane), so that the table's run-time rows and columns
 * index it without private-memory arrays. */
__device__ int sos_det5(const Ball b[5], const uint32_t ids[5],
                        const LdCW<kTPI>& L, uint32_t* lm) {
  /* rows in increasing id order by a 9-comparator sorting network (all
   * indices known at compile time: no private-memory array), the
   * permutation's parity counted by its swaps */
  uint32_t k[5] = {ids[0], ids[1], ids[2], ids[3], ids[4]};
  int ord[5] = {0, 1, 2, 3, 4};
  int swaps = 0;
  constexpr int net[9][2] = {{0, 1}, {3, 4}, {2, 4}, {2, 3}, {0, 3},
                             {0, 2}, {1, 4}, {1, 3}, {1, 2}};
#pragma unroll
  for (int c = 0; c < 9; ++c) {
    const int i = net[c][0], j = net[c][1];
    const bool sw = k[j] < k[i];
    const uint32_t ki = k[i], kj = k[j];
    const int oi = ord[i], oj = ord[j];
    k[i] = sw ? kj : ki;
    k[j] = sw ? ki : kj;
    ord[i] = sw ? oj : oi;
    ord[j] = sw ? oi : oj;
    swaps += sw;
  }
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int r = 0; r < 5; ++r) {
    Ball br = b[0];
#pragma unroll
    for (int q = 1; q < 5; ++q)
      if (ord[r] == q) br = b[q];
    const V3<XI> p = L.p(br);
    const XI l = p.x * p.x + p.y * p.y + p.z * p.z - L.w(br);
    lm[(4 * r + 0) * 64 + lane] = p.x.v;
    lm[(4 * r + 1) * 64 + lane] = p.y.v;
    lm[(4 * r + 2) * 64 + lane] = p.z.v;
    lm[(4 * r + 3) * 64 + lane] = l.v;
  }
  auto at = [&](int r, int c) { return XI{lm[(4 * r + c) * 64 + lane]}; };
  const int parity = (swaps & 1) ? -1 : 1;
#pragma unroll 1
  for (int t = 0; t < PAC_SOS_TERMS; ++t) {
    unsigned rm = (unsigned)pac_sos_table[t][0];
    /* the ones column (bit 4) is in every term */
    unsigned cm = (unsigned)pac_sos_table[t][1] & 0xfu;
    const int m = __builtin_popcount(rm);
    int s;
    if (m == 1) {
      s = 1; /* the ones column alone */
    } else {
      /* rows r0 < r1 < ... < z (z the last), columns c0 < c1 < c2 */
      const int z = 31 - __builtin_clz(rm);
      const int r0 = __builtin_ctz(rm);
      rm &= rm - 1;
      const int r1 = __builtin_ctz(rm);
      rm &= rm - 1;
      const int r2 = __builtin_ctz(rm);
      const int c0 = __builtin_ctz(cm);
      cm &= cm - 1;
      const int c1 = cm ? __builtin_ctz(cm) : 0;
      cm &= cm - 1;
      const int c2 = cm ? __builtin_ctz(cm) : 0;
      auto d = [&](int r, int c) { return at(r, c) - at(z, c); };
      if (m == 2) {
        s = fe_sign(d(r0, c0));
      } else if (m == 3) {
        s = fe_sign(d(r0, c0) * d(r1, c1) - d(r0, c1) * d(r1, c0));
      } else {
        const XI a00 = d(r0, c0), a01 = d(r0, c1), a02 = d(r0, c2);
        const XI a10 = d(r1, c0), a11 = d(r1, c1), a12 = d(r1, c2);
        const XI a20 = d(r2, c0), a21 = d(r2, c1), a22 = d(r2, c2);
        s = fe_sign(a00 * (a11 * a22 - a12 * a21) -
                    a01 * (a10 * a22 - a12 * a20) +
                    a02 * (a10 * a21 - a11 * a20));
      }
    }
    if (s != 0) return parity * pac_sos_table[t][2] * s;
  }
  return parity;
}

/* ---------------- tet emptiness ---------------- */
template <class T>
struct TetQ {
  T D, wa;
  V3<T> pa, N;
};
template <class T>
__device__ __forceinline__ void tetq_init(TetQ<T>& t, const Ball v[4]) {
  t.pa = Ld<T>::p(v[0]);
  t.wa = Ld<T>::w(v[0]);
  tet_DN(t.pa, t.wa, Ld<T>::p(v[1]), Ld<T>::w(v[1]), Ld<T>::p(v[2]),
         Ld<T>::w(v[2]), Ld<T>::p(v[3]), Ld<T>::w(v[3]), &t.D, &t.N);
}
/* 1 violated, 0 not violated, 2 undecided (needs exact) */
template <class T>
__device__ __forceinline__ int tet_pair_code(const TetQ<T>& t, int sD,
                                             const Ball& e) {
  T g = power_gap(t.pa, t.wa, Ld<T>::p(e), Ld<T>::w(e), t.D, t.N);
  int s = fe_sign(g);
  if (s == 0) return 2;
  return (s * sD) < 0 ? 1 : 0;
}


/* per-simplex setup at precision R: D, N with a-priori error bounds
 * (tet_dn_ss), the certified centre box and the semi-static pair
 * constants */
template <class R>
__device__ __forceinline__ bool tet_pairs(const Ball v[4], PairSS<R>& P,
                                          TetCert<R>& T) {
  tet_dn_ss(v, T);
  V3<FE<R>> pa{FE<R>{(R)v[0].x, (R)0}, FE<R>{(R)v[0].y, (R)0},
               FE<R>{(R)v[0].z, (R)0}};
  return pairss_init(P, pa, T.wa, T.D, T.N);
}
/* the pair constants and the certified box that holds every violator */
template <class R, class Box>
__device__ __forceinline__ bool tet_setup(const GridP& g, const Ball v[4],
                                          PairSS<R>& P, Box& box) {
  TetCert<R> T;
  if (!tet_pairs(v, P, T)) return false;
  /* centre, reach and box in the stage's own precision */
  R c[3], e[3], rho_up;
  if (!centre_err<R>(T.D, T.N, v[0], T.wa, c, e, &rho_up)) return false;
  return place_box<R>(g, simplex_reach<R>(g, rho_up), c, e, box);
}

/* Quick rejection before the certified box is built: the balls of the
 * fine-grid row through an FP32 estimate of the centre, within the grid's
 * reach of it in x, go through the certified pair test, and a certain
 * violator rejects the simplex.  The region needs no guarantee (nothing is
 * concluded from its other balls), so the estimate need not be certified;
 * a vertex of the simplex never tests as a certain violator.  Most alpha
 * simplices are rejected by a ball near their centre: this spares them the
 * FP64 centre bounds and box. */
template <class R>
__device__ __forceinline__ bool quick_reject(const Stage& S, const GridP& g,
                                             const PairSS<R>& P,
                                             unsigned& pairs) {
  const float h = 0.5f / (float)P.den;
  const float cx = (float)P.ax + (float)P.nx * h;
  const float cy = (float)P.ay + (float)P.ny * h;
  const float cz = (float)P.az + (float)P.nz * h;
  if (!(fabsf(cx) < 1e30f && fabsf(cy) < 1e30f && fabsf(cz) < 1e30f))
    return false;
  const float iyz = (float)g.inv_yz, ix = (float)g.inv_step, rf = (float)g.reach;
  auto cell = [](float q, int e) {
    return (int)floorf(fminf(fmaxf(q, 0.f), (float)(e - 1)));
  };
  const int y = cell((cy - g.miny) * iyz, g.ey);
  const int z = cell((cz - g.minz) * iyz, g.ez);
  const int xl = cell((cx - rf - g.minx) * ix, g.ex);
  const int xh = cell((cx + rf - g.minx) * ix, g.ex);
  const int64_t row = (int64_t)g.cbase + (int64_t)g.ex * (y + (int64_t)g.ey * z);
  const uint32_t lo = S.fstart[row + xl], hi = S.fstart[row + xh + 1];
  bool dead = false;
  auto visit = [&](uint32_t, const Ball& E) {
    ++pairs;
    if (pairss_sign(P, E) < 0) {
      dead = true;
      return true;
    }
    return false;
  };
  scan_range4(S.fball, lo, hi, visit);
  return dead;
}

__device__ __forceinline__ bool owned(const Stage& S, uint32_t v) {
  return v >= S.own_lo && v < S.own_hi;
}
/* chunk sub-input: does an owned simplex depend on this tet? (one of its
 * faces or edges owned: its lowest, second or third vertex owned) */
__device__ __forceinline__ bool tet_needed(const Stage& S, const uint4& t) {
  return !S.chunk || owned(S, t.x) || owned(S, t.y) || owned(S, t.z);
}

/* chunk sub-input without the quick pass: the tets that need a test.
 * (The decision is carried in a bool: written as a sentinel id assigned in
 * one branch of an if/else whose other branch stores, hipcc dropped the
 * assignment.) */
__global__ __launch_bounds__(kBlock) void k_tet_need(Stage S, Cnt c_nk,
                                                     const uint4* tets,
                                                     uint8_t* flag,
                                                     uint32_t* next,
                                                     unsigned* nnext) {
  const int64_t nk = c_nk.get();
  const int64_t id = gtid();
  bool need = false;
  if (id < nk) {
    need = tet_needed(S, tets[id]);
    if (!need) flag[id] = 0;
  }
  list_append(next, nnext, S.lcap, need, (uint32_t)id);
}

/* Quick pass over the alpha tets (fine scan grid): FP32 pair constants
 * and quick_reject; the tets it does not reject (or whose pair constants
 * are not certified) are listed in `next` for the certified scan.  A
 * kernel of its own: the certified scan's registers stay off this path. */
template <class R, bool MOL>
__global__ __launch_bounds__(kBlock) void k_tet_ac2_quick(
    Stage S, Cnt c_nk, const uint4* tets, uint8_t* flag, uint32_t* next,
    unsigned* nnext) {
  unsigned pairs = 0;
  const int64_t nk = c_nk.get();
  const int64_t id = gtid();
  uint32_t keep = 0xffffffffu;
  if (id < nk) {
    const uint4 t = tets[id];
    const Ball v[4] = {S.ball[t.x], S.ball[t.y], S.ball[t.z], S.ball[t.w]};
    const GridP g = sgrid_at<MOL>(S, t.x);
    PairSS<R> P;
    TetCert<R> T;
    if (!tet_needed(S, t) || (tet_pairs(v, P, T) && quick_reject(S, g, P, pairs)))
      flag[id] = 0;
    else
      keep = (uint32_t)id;
  }
  list_append(next, nnext, S.lcap, keep != 0xffffffffu, keep);
  block_flush(S.ctr, CTR_PAIRS + 0, pairs);
}

/* Phase A (budget > 0): at most `budget` pair tests per tet, in scan order;
 * tets not settled within the budget are appended to `rest`.  Phase B
 * (grid-stride over `rest`): the remaining pairs of those tets, skipping the
 * first `budget` ones.  Splitting keeps lanes of a wave busy with similar
 * amounts of work: most tets are rejected by one of their first few
 * neighbours, while accepted tets must test every ball of their box.
 * FINE: the scan runs over the fine grid's cells that meet the sphere
 * (balls in fine-cell order, fid their ids).  A vertex of the simplex lies
 * on its orthosphere, so its certified test is always undecided: the
 * fine scan looks a ball's id up only then, to skip the vertices. */
template <class R, bool PHASE_B, bool FINE, bool MOL>
__global__ __launch_bounds__(kBlock) void k_tet_ac2_hot(
    Stage S, Cnt c_nk, const uint32_t* items, const unsigned* nitems,
    const uint4* tets, uint8_t* flag, Queue q1, uint32_t* rest,
    unsigned* nrest, int budget) {
  unsigned evals = 0, pairs = 0;
  /* phase A: the alpha tets, or the list `items` of them (those the quick
   * pass left); phase B: `rest` */
  __shared__ uint32_t s_pre[LSLOTS + 1];
  if (PHASE_B) list_table(nrest, S.lcap, s_pre);
  else if (items) list_table(nitems, S.lcap, s_pre);
  const int64_t total =
      PHASE_B || items ? (int64_t)s_pre[LSLOTS] : c_nk.get();
  uint32_t rest_id = 0xffffffffu; /* phase A: at most one item per thread */
  auto item = [&](int64_t w) {
    const int64_t id =
        PHASE_B ? (int64_t)list_at(rest, S.lcap, s_pre, (uint32_t)w)
                : (items ? (int64_t)list_at(items, S.lcap, s_pre, (uint32_t)w)
                         : w);
    uint4 t = tets[id];
    Ball v[4] = {S.ball[t.x], S.ball[t.y], S.ball[t.z], S.ball[t.w]};
    const GridP g = FINE ? sgrid_at<MOL>(S, t.x) : grid_at<MOL>(S, t.x);
    PairSS<R> P;
    typename std::conditional<FINE, SphBox, CellBox>::type box;
    if (!tet_setup(g, v, P, box)) {
      if (!PHASE_B) {
        push_q(q1, (uint32_t)id, 0xffffffffu);
        flag[id] = 2;
      }
      return;
    }
    /* why the scan stopped early: 1 a violator, 2 the budget (one
     * variable: two flags set in different exits were kept in scratch
     * memory, their stores merged through a select of addresses) */
    bool pending = false;
    int stop = 0;
    int seen = 0;
    auto visit = [&](uint32_t j, const Ball& E) {
      if (!FINE && (j == t.x || j == t.y || j == t.z || j == t.w)) return false;
      if (PHASE_B) {
        if (seen < budget) {
          ++seen;
          return false;
        }
      } else if (budget > 0 && seen == budget) {
        stop = 2;
        return true;
      } else {
        ++seen;
      }
      ++pairs;
      ++evals;
      int s = pairss_sign(P, E);
      if (s == 0) {
        const uint32_t p = FINE ? S.fid[j] : j;
        if (FINE && (p == t.x || p == t.y || p == t.z || p == t.w)) return false;
        pending = true;
        push_q(q1, (uint32_t)id, p);
        return false;
      }
      if (s < 0) { /* pi_e < rho: violator */
        stop = 1;
        return true;
      }
      return false;
    };
    /* fine grid: up to four balls at a time.  Phase A stops before a
     * group once `budget` balls are seen, phase B skips exactly the groups
     * phase A tested (the same traversal, the same counts).  The tests of
     * a group all run; a certain violator among them ends the scan (its
     * ties need no queueing: the tet is rejected). */
    auto group = [&](uint32_t p, uint32_t n) -> bool {
      if (PHASE_B) {
        if (seen < budget) {
          seen += (int)n;
          return false;
        }
      } else if (budget > 0 && seen >= budget) {
        stop = 2;
        return true;
      }
      seen += (int)n;
      Ball E[4];
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k) E[k] = S.fball[p + min(k, n - 1)];
      /* a ball repeated past n (E[k] = E[n-1]) repeats ball n-1's bits:
       * the first violator is unchanged, the ties are cut to n balls */
      unsigned viol, ties;
      pairss_masks(P, E, viol, ties);
      ties &= (1u << n) - 1u;
      /* pairs counted as a scan stopping at its first violator tests them
       * (the group's tests past it are not work the scan needs) */
      pairs += viol ? (unsigned)__builtin_ctz(viol) + 1u : n;
      if (viol) {
        stop = 1;
        return true;
      }
      while (ties) { /* own vertices (always ties) and exact ties */
        const uint32_t k = __builtin_ctz(ties);
        ties &= ties - 1;
        const uint32_t q = S.fid[p + k];
        if (q == t.x || q == t.y || q == t.z || q == t.w) continue;
        pending = true;
        push_q(q1, (uint32_t)id, q);
      }
      return false;
    };
    if constexpr (FINE)
      scan_sph_flat(g, S.fstart, box, group);
    else
      scan_box4(g, S.cstart, S.ball, box, visit);
    if (PHASE_B) {
      if (stop == 1)
        flag[id] = 0;
      else if (pending)
        flag[id] = 2;
    } else {
      flag[id] = stop == 1 ? 0 : (pending ? 2 : 1);
      if (stop == 2) rest_id = (uint32_t)id;
    }
  };
  if constexpr (PHASE_B) {
    for_work(S.bchunk, total, item); /* persistent launch */
  } else {
    if (gtid() < total) item(gtid());
  }
  if (!PHASE_B) {
    list_append(rest, nrest, S.lcap, rest_id != 0xffffffffu, rest_id);
  }
  (void)evals;
  block_flush(S.ctr, CTR_PAIRS + 0, pairs);
}

/* FP64 then double-double for queued pairs / uncertified tets */
template <class T>
__device__ __forceinline__ int tet_pair_code_T(const TetQ<T>& t, int sD,
                                               const Ball& e) {
  return tet_pair_code(t, sD, e);
}
__global__ __launch_bounds__(kBlock) void k_tet_ac2_esc(
    Stage S, QRef q, const uint4* tets, uint8_t* flag,
    Queue q2) {
  StageCnt sc;
  for_queue(q, [&](uint32_t t) {
    if (!live_entry(q.at(t), flag)) return;
    QEntry e = q.at(t);
    uint4 tt = tets[e.item];
    Ball v[4] = {S.ball[tt.x], S.ball[tt.y], S.ball[tt.z], S.ball[tt.w]};
    TetQ<FE<double>> t64;
    tetq_init(t64, v);
    int sD = fe_sign(t64.D);
    TetQ<FEDD> tdd;
    bool have_dd = false;
    if (!sD) {
      tetq_init(tdd, v);
      have_dd = true;
      sD = fe_sign(tdd.D);
    }
    auto pair = [&](uint32_t p) -> int { /* 1 violated, 0 clear, 2 undecided */
      Ball E = S.ball[p];
      int c = 2;
      if (fe_sign(t64.D) != 0) {
        sc.c[1]++;
        c = tet_pair_code(t64, sD, E);
      }
      if (c == 2) {
        if (!have_dd) {
          tetq_init(tdd, v);
          have_dd = true;
        }
        if (fe_sign(tdd.D) != 0) {
          sc.c[2]++;
          c = tet_pair_code(tdd, sD, E);
        }
      }
      return c;
    };
    if (e.aux != 0xffffffffu) {
      int c = sD ? pair(e.aux) : 2;
      if (c == 1) flag[e.item] = 0;
      else if (c == 2) push_q(q2, e.item, e.aux);
    } else if (!sD) {
      push_q(q2, e.item, 0xffffffffu);
    } else {
      /* full scan with the best certified centre */
      double cx, cy, cz, ex, ey, ez;
      if (fe_sign(t64.D) != 0) {
        double D2 = 2.0 * t64.D.v;
        cx = (double)v[0].x + t64.N.x.v / D2;
        cy = (double)v[0].y + t64.N.y.v / D2;
        cz = (double)v[0].z + t64.N.z.v / D2;
        ex = quot_err(t64.N.x, t64.D, 2.0);
        ey = quot_err(t64.N.y, t64.D, 2.0);
        ez = quot_err(t64.N.z, t64.D, 2.0);
      } else {
        double D2 = 2.0 * tdd.D.v.hi;
        cx = (double)v[0].x + tdd.N.x.v.hi / D2;
        cy = (double)v[0].y + tdd.N.y.v.hi / D2;
        cz = (double)v[0].z + tdd.N.z.v.hi / D2;
        ex = quot_err(tdd.N.x, tdd.D, 2.0);
        ey = quot_err(tdd.N.y, tdd.D, 2.0);
        ez = quot_err(tdd.N.z, tdd.D, 2.0);
      }
      CellBox box;
      const GridP g = grid_of(S, tt.x);
      if (!make_box(g, cx, cy, cz, ex, ey, ez, box)) {
        push_q(q2, e.item, 0xffffffffu);
      } else {
        bool dead = false;
        scan_box(g, S.cstart, box, [&](uint32_t p) {
          if (p == tt.x || p == tt.y || p == tt.z || p == tt.w) return false;
          int c = pair(p);
          if (c == 1) {
            dead = true;
            return true;
          }
          if (c == 2) push_q(q2, e.item, p);
          return false;
        });
        if (dead) flag[e.item] = 0;
      }
    }
  });
  sc.flush(S.ctr, FAM_TET_AC2);
}

/* exact resolution of one (tet, ball) pair, or a full exact scan when
 * aux == ~0 (tet whose orientation no float stage could certify); one
 * item per group of kTPI lanes */
__device__ int tet_pair_exact(const Stage& S, const Ball v[4],
                              const uint32_t vid[4], const uint32_t* perm,
                              const XI& D, const V3<XI>& N, const V3<XI>& pa,
                              const XI& wa, uint32_t p, unsigned* sos,
                              uint32_t* lm) {
  const LdCW<kTPI> L{S.ex.S};
  Ball E = S.ball[p];
  XI g = power_gap(pa, wa, L.p(E), L.w(E), D, N);
  int s = fe_sign(g), sD = fe_sign(D);
  if (s != 0) return (s * sD) < 0;
  /* exact tie: insphere_w = det[x y z l 1](a,b,c,d,e) with SoS; the tet
   * is violated iff SoS * orient3d(a,b,c,d) > 0, orient3d = -sign(D)
   * (AlphaComplexParallel.cu:653-683). */
  ++*sos;
  Ball bb[5] = {v[0], v[1], v[2], v[3], E};
  uint32_t ids[5] = {perm[vid[0]], perm[vid[1]], perm[vid[2]], perm[vid[3]],
                     perm[p]};
  int ss = sos_det5(bb, ids, L, lm);
  return (ss * -sD) > 0;
}

/* The exact tet test's per-tet part: its scaled vertices' D and N, the
 * first vertex's position and weight (what every pair test of the tet
 * reuses). */
struct TetExact {
  uint32_t item = 0xffffffffu;
  uint4 tt;
  uint32_t vid[4];
  Ball v[4];
  V3<XI> pa, N;
  XI wa, D;
};
__device__ __forceinline__ void tet_exact_setup(const Stage& S,
                                                const uint4* tets,
                                                uint32_t item, TetExact& T) {
  T.item = item;
  T.tt = tets[item];
  T.vid[0] = T.tt.x;
  T.vid[1] = T.tt.y;
  T.vid[2] = T.tt.z;
  T.vid[3] = T.tt.w;
#pragma unroll
  for (int i = 0; i < 4; ++i) T.v[i] = S.ball[T.vid[i]];
  const LdCW<kTPI> L{S.ex.S};
  T.pa = L.p(T.v[0]);
  T.wa = L.w(T.v[0]);
  tet_DN(T.pa, T.wa, L.p(T.v[1]), L.w(T.v[1]), L.p(T.v[2]), L.w(T.v[2]),
         L.p(T.v[3]), L.w(T.v[3]), &T.D, &T.N);
}
/* One queued entry of tet T: a (tet, ball) pair, or a full exact scan of
 * the tet's box (aux == ~0, a tet whose orientation no float stage could
 * certify).  Group-uniform; the group's first lane writes the flag. */
__device__ __forceinline__ void tet_exact_entry(const Stage& S,
                                                const TetExact& T, uint32_t aux,
                                                const uint32_t* perm,
                                                uint8_t* flag, bool lead,
                                                unsigned& cnt, unsigned& sos,
                                                uint32_t* lm) {
  if (aux != 0xffffffffu) {
    ++cnt;
    if (tet_pair_exact(S, T.v, T.vid, perm, T.D, T.N, T.pa, T.wa, aux, &sos,
                       lm) &&
        lead)
      flag[T.item] = 0;
  } else if (fe_sign(T.D) != 0) {
    double D2 = 2.0 * cw_to_double(T.D, 3 * S.ex.S);
    double cx = (double)T.v[0].x + cw_to_double(T.N.x, 4 * S.ex.S) / D2;
    double cy = (double)T.v[0].y + cw_to_double(T.N.y, 4 * S.ex.S) / D2;
    double cz = (double)T.v[0].z + cw_to_double(T.N.z, 4 * S.ex.S) / D2;
    double err = 1e-9 * (fabs(cx) + fabs(cy) + fabs(cz) + 1.0);
    CellBox box;
    bool viol = false;
    const GridP g = grid_of(S, T.tt.x);
    if (make_box(g, cx, cy, cz, err, err, err, box)) {
      for (int z = box.lo[2]; z <= box.hi[2] && !viol; ++z)
        for (int y = box.lo[1]; y <= box.hi[1] && !viol; ++y) {
          int64_t row = (int64_t)g.cbase + (int64_t)g.ex * (y + (int64_t)g.ey * z);
          uint32_t b = S.cstart[row + box.lo[0]],
                   en = S.cstart[row + box.hi[0] + 1];
          for (uint32_t p = b; p < en; ++p) {
            if (p == T.tt.x || p == T.tt.y || p == T.tt.z || p == T.tt.w)
              continue;
            ++cnt;
            if (tet_pair_exact(S, T.v, T.vid, perm, T.D, T.N, T.pa, T.wa, p,
                               &sos, lm)) {
              viol = true;
              break;
            }
          }
        }
    }
    if (lead) flag[T.item] = viol ? 0 : 2;
  }
}

__global__ __launch_bounds__(kBlock) void k_tet_ac2_exact(
    Stage S, QRef q, const uint4* tets,
    const uint32_t* perm, uint8_t* flag) {
  unsigned cnt = 0, sos = 0;
  const bool lead = Grp<kTPI>::idx() == 0;
  __shared__ uint32_t s_lift[kBlock / 64][20 * 64]; /* SoS lifted matrices */
  uint32_t* const lm = s_lift[threadIdx.x >> 6];
  for_queue_grp<kTPI>(q, [&](uint32_t t) {
    /* the group's first lane decides, so the group stays together */
    if (!Grp<kTPI>::from(live_entry(q.at(t), flag) ? 1u : 0u, 0)) return;
    QEntry e = q.at(t);
    TetExact T;
    tet_exact_setup(S, tets, e.item, T);
    tet_exact_entry(S, T, e.aux, perm, flag, lead, cnt, sos, lm);
  });
  /* every lane of a group counted the group's evaluations */
  flush_ctr(S.ctr, CTR_STAGE + FAM_TET_AC2 * 4 + 3, lead ? cnt : 0u);
  flush_ctr(S.ctr, CTR_SOS, lead ? sos : 0u);
}

/* The same over the entries sorted by tet (item[], aux[]): each group
 * takes kSortedRun consecutive entries at a time, so the pairs of one tet
 * (an exact tie usually has several co-spherical balls) share its D and N
 * instead of recomputing them per pair. */
constexpr uint32_t kSortedRun = 8;
__global__ __launch_bounds__(kBlock) void k_tet_ac2_exact_sorted(
    Stage S, const uint32_t* item, const uint32_t* aux, uint32_t n,
    const uint4* tets, const uint32_t* perm, uint8_t* flag) {
  unsigned cnt = 0, sos = 0;
  const bool lead = Grp<kTPI>::idx() == 0;
  __shared__ uint32_t s_lift[kBlock / 64][20 * 64];
  uint32_t* const lm = s_lift[threadIdx.x >> 6];
  const uint64_t groups = (uint64_t)gridDim.x * blockDim.x / kTPI;
  TetExact T;
  for (uint64_t c0 = (uint64_t)(gtid() / kTPI) * kSortedRun; c0 < n;
       c0 += groups * kSortedRun) {
    const uint32_t c1 = (uint32_t)min((uint64_t)n, c0 + kSortedRun);
    for (uint32_t k = (uint32_t)c0; k < c1; ++k) {
      const uint32_t it = item[k], ax = aux[k];
      const bool live = ax == 0xffffffffu || flag[it] != 0;
      if (!Grp<kTPI>::from(live ? 1u : 0u, 0)) continue;
      if (it != T.item) tet_exact_setup(S, tets, it, T);
      tet_exact_entry(S, T, ax, perm, flag, lead, cnt, sos, lm);
    }
  }
  flush_ctr(S.ctr, CTR_STAGE + FAM_TET_AC2 * 4 + 3, lead ? cnt : 0u);
  flush_ctr(S.ctr, CTR_SOS, lead ? sos : 0u);
}

/* entries of a host-built view, as (item, aux) arrays */
__global__ __launch_bounds__(kBlock) void k_view_split(QView v, uint32_t* item,
                                                       uint32_t* aux) {
  const int64_t i = gtid();
  if (i >= (int64_t)v.size()) return;
  const QEntry e = v.at((uint32_t)i);
  item[i] = e.item;
  aux[i] = e.aux;
}


/* ---------------- triangle / edge emptiness ---------------- */
template <class T>
struct LowQ {   /* triangle or edge: power gap with den, num */
  T den, wa;
  V3<T> pa, num;
};
template <class T>
__device__ __forceinline__ void triq_init(LowQ<T>& t, const Ball& a,
                                          const Ball& b, const Ball& c) {
  t.pa = Ld<T>::p(a);
  t.wa = Ld<T>::w(a);
  tri_X(t.pa, t.wa, Ld<T>::p(b), Ld<T>::w(b), Ld<T>::p(c), Ld<T>::w(c),
        &t.den, &t.num);
}
template <class T>
__device__ __forceinline__ void edgeq_init(LowQ<T>& t, const Ball& a,
                                           const Ball& b) {
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
                                             const Ball& e) {
  T g = power_gap(t.pa, t.wa, Ld<T>::p(e), Ld<T>::w(e), t.den, t.num);
  int s = fe_sign(g);
  return s == 0 ? 2 : (s < 0 ? 1 : 0);
}


template <class R, int KIND, class Box>
__device__ __forceinline__ bool low_setup(const GridP& g, const Ball* v,
                                          PairSS<R>& P, Box& box) {
  const R ax = v[0].x, ay = v[0].y, az = v[0].z;
  R ux = (R)v[1].x - ax, uy = (R)v[1].y - ay, uz = (R)v[1].z - az;
  R wa = (R)v[0].w * (R)v[0].w, w1 = (R)v[1].w * (R)v[1].w;
  R uu = ux * ux + uy * uy + uz * uz;
  SS<R> A{(uu + wa) - w1, (uu + wa) + w1};
  FE<R> den;
  V3<FE<R>> num;
  if (KIND == 0) {
    R vx = (R)v[2].x - ax, vy = (R)v[2].y - ay, vz = (R)v[2].z - az;
    R w2 = (R)v[2].w * (R)v[2].w;
    R vv = vx * vx + vy * vy + vz * vz;
    SS<R> B{(vv + wa) - w2, (vv + wa) + w2};
    /* permanents bounded by max norms: every component of n = u x v by
     * pn = 2 |u|_inf |v|_inf, of v x n by 2 |v|_inf pn, of u x n by
     * 2 |u|_inf pn */
    const R mu = fmax(fabs(ux), fmax(fabs(uy), fabs(uz)));
    const R mv = fmax(fabs(vx), fmax(fabs(vy), fabs(vz)));
    const R pn = ((R)2 * mu) * mv;
    const R nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
    SS<R> nn{nx * nx + ny * ny + nz * nz, ((R)3 * pn) * pn};
    /* X = A (v x n) - B (u x n) */
    const R vn[3] = {vy * nz - vz * ny, vz * nx - vx * nz, vx * ny - vy * nx};
    const R un[3] = {uy * nz - uz * ny, uz * nx - ux * nz, ux * ny - uy * nx};
    const R xp = ((R)2 * pn) * (A.p * mv + B.p * mu);
    SS<R> X[3];
#pragma unroll
    for (int q = 0; q < 3; ++q)
      X[q] = SS<R>{A.v * vn[q] - B.v * un[q], xp};
    den = ss_fe(nn, 8);
    num = V3<FE<R>>{ss_fe(X[0], 12), ss_fe(X[1], 12), ss_fe(X[2], 12)};
  } else {
    SS<R> uuS{uu, uu};
    den = ss_fe(uuS, 4);
    num = V3<FE<R>>{ss_fe(SS<R>{A.v * ux, A.p * fabs(ux)}, 8),
                    ss_fe(SS<R>{A.v * uy, A.p * fabs(uy)}, 8),
                    ss_fe(SS<R>{A.v * uz, A.p * fabs(uz)}, 8)};
  }
  V3<FE<R>> pa{FE<R>{ax, (R)0}, FE<R>{ay, (R)0}, FE<R>{az, (R)0}};
  FE<R> waf{wa, FEc<R>::u() * wa * FEc<R>::slack()};
  if (!pairss_init(P, pa, waf, den, num) || P.sden < 0) return false;
  R c[3], e[3], rho_up;
  if (!centre_err<R>(den, num, v[0], waf, c, e, &rho_up)) return false;
  return place_box<R>(g, simplex_reach<R>(g, rho_up), c, e, box);
}

/* KIND 0: triangles (items = alpha triangle ids), 1: edges (alpha edge ids).
 * Simplices already in the closure (in_flag) pass without a test.  Same
 * phase A / phase B split as the tet kernel. */
template <class R, int KIND, bool PHASE_B, bool FINE, bool MOL>
__global__ __launch_bounds__(kBlock) void k_low_ac2_hot(
    Stage S, Cnt c_nitems, const uint32_t* ecol, const uint32_t* eleft,
    const uint32_t* tk, const uint32_t* tedge, const uint8_t* in_flag,
    uint8_t* flag, Queue q1, uint32_t* rest, unsigned* nrest, int budget) {
  unsigned evals = 0, pairs = 0;
  __shared__ uint32_t s_pre[LSLOTS + 1];
  if (PHASE_B) list_table(nrest, S.lcap, s_pre);
  const int64_t total = PHASE_B ? (int64_t)s_pre[LSLOTS] : c_nitems.get();
  uint32_t rest_id = 0xffffffffu; /* phase A: at most one item per thread */
  auto item = [&](int64_t w) {
    const int64_t id =
        PHASE_B ? (int64_t)list_at(rest, S.lcap, s_pre, (uint32_t)w) : w;
    if (!PHASE_B && in_flag[id]) {
      flag[id] = 1;
      return;
    }
    uint32_t vid[3];
    if (KIND == 0) {
      uint32_t e = tedge[id];
      vid[0] = eleft[e];
      vid[1] = ecol[e];
      vid[2] = tk[id];
    } else {
      vid[0] = eleft[id];
      vid[1] = ecol[id];
      vid[2] = vid[1];
    }
    Ball v[3] = {S.ball[vid[0]], S.ball[vid[1]], S.ball[vid[2]]};
    const GridP g = FINE ? sgrid_at<MOL>(S, vid[0]) : grid_at<MOL>(S, vid[0]);
    PairSS<R> P;
    typename std::conditional<FINE, SphBox, CellBox>::type box;
    if (!low_setup<R, KIND>(g, v, P, box)) {
      if (!PHASE_B) {
        push_q(q1, (uint32_t)id, 0xffffffffu);
        flag[id] = 2;
      }
      return;
    }
    bool pending = false;
    int stop = 0; /* 1 a violator, 2 the budget (as in k_tet_ac2_hot) */
    int seen = 0;
    auto visit = [&](uint32_t j, const Ball& E) {
      if (!FINE && (j == vid[0] || j == vid[1] || j == vid[2])) return false;
      if (PHASE_B) {
        if (seen < budget) {
          ++seen;
          return false;
        }
      } else if (budget > 0 && seen == budget) {
        stop = 2;
        return true;
      } else {
        ++seen;
      }
      ++pairs;
      ++evals;
      int s = pairss_sign(P, E);
      if (s < 0) { /* power <= rho strictly: violator */
        stop = 1;
        return true;
      }
      if (s == 0) {
        const uint32_t p = FINE ? S.fid[j] : j;
        if (FINE && (p == vid[0] || p == vid[1] || p == vid[2])) return false;
        pending = true;
        push_q(q1, (uint32_t)id, p);
      }
      return false;
    };
    /* fine grid: four balls at a time, as in k_tet_ac2_hot */
    auto group = [&](uint32_t p, uint32_t n) -> bool {
      if (PHASE_B) {
        if (seen < budget) {
          seen += (int)n;
          return false;
        }
      } else if (budget > 0 && seen >= budget) {
        stop = 2;
        return true;
      }
      seen += (int)n;
      Ball E[4];
#pragma unroll
      for (uint32_t k = 0; k < 4; ++k) E[k] = S.fball[p + min(k, n - 1)];
      unsigned viol, ties; /* bits past n as in k_tet_ac2_hot */
      pairss_masks(P, E, viol, ties);
      ties &= (1u << n) - 1u;
      /* pairs counted as a scan stopping at its first violator tests them
       * (the group's tests past it are not work the scan needs) */
      pairs += viol ? (unsigned)__builtin_ctz(viol) + 1u : n;
      if (viol) { /* power <= rho strictly: violator */
        stop = 1;
        return true;
      }
      while (ties) {
        const uint32_t k = __builtin_ctz(ties);
        ties &= ties - 1;
        const uint32_t q = S.fid[p + k];
        if (q == vid[0] || q == vid[1] || q == vid[2]) continue;
        pending = true;
        push_q(q1, (uint32_t)id, q);
      }
      return false;
    };
    if constexpr (FINE)
      scan_sph_flat(g, S.fstart, box, group);
    else
      scan_box4(g, S.cstart, S.ball, box, visit);
    if (PHASE_B) {
      if (stop == 1)
        flag[id] = 0;
      else if (pending)
        flag[id] = 2;
    } else {
      flag[id] = stop == 1 ? 0 : (pending ? 2 : 1);
      if (stop == 2) rest_id = (uint32_t)id;
    }
  };
  if constexpr (PHASE_B) {
    for_work(S.bchunk, total, item); /* persistent launch */
  } else {
    if (gtid() < total) item(gtid());
  }
  if (!PHASE_B) {
    list_append(rest, nrest, S.lcap, rest_id != 0xffffffffu, rest_id);
  }
  (void)evals;
  block_flush(S.ctr, CTR_PAIRS + 1 + KIND, pairs);
}

template <int KIND>
__global__ __launch_bounds__(kBlock) void k_low_ac2_esc(
    Stage S, QRef q, const uint32_t* ecol,
    const uint32_t* eleft, const uint32_t* tk, const uint32_t* tedge,
    uint8_t* flag, Queue q2) {
  StageCnt sc;
  for_queue(q, [&](uint32_t t) {
    if (!live_entry(q.at(t), flag)) return;
    QEntry e = q.at(t);
    uint32_t vid[3];
    if (KIND == 0) {
      uint32_t ed = tedge[e.item];
      vid[0] = eleft[ed];
      vid[1] = ecol[ed];
      vid[2] = tk[e.item];
    } else {
      vid[0] = eleft[e.item];
      vid[1] = ecol[e.item];
      vid[2] = vid[1];
    }
    Ball v[3] = {S.ball[vid[0]], S.ball[vid[1]], S.ball[vid[2]]};
    LowQ<FE<double>> q64;
    LowQ<FEDD> qdd;
    bool have_dd = false;
    if (KIND == 0) triq_init(q64, v[0], v[1], v[2]); else edgeq_init(q64, v[0], v[1]);
    auto init_dd = [&]() {
      if (!have_dd) {
        if (KIND == 0) triq_init(qdd, v[0], v[1], v[2]); else edgeq_init(qdd, v[0], v[1]);
        have_dd = true;
      }
    };
    auto pair = [&](uint32_t p) -> int { /* 1 violated, 0 clear, 2 undecided */
      Ball E = S.ball[p];
      int c = 2;
      if (fe_sign(q64.den) > 0) {
        sc.c[1]++;
        c = low_pair_code(q64, E);
      }
      if (c == 2) {
        init_dd();
        if (fe_sign(qdd.den) > 0) {
          sc.c[2]++;
          c = low_pair_code(qdd, E);
        }
      }
      return c;
    };
    if (e.aux != 0xffffffffu) {
      int c = pair(e.aux);
      if (c == 1) flag[e.item] = 0;
      else if (c == 2) push_q(q2, e.item, e.aux);
    } else {
      double cx = 0, cy = 0, cz = 0, ex = INFINITY, ey = INFINITY, ez = INFINITY;
      bool ok = false;
      if (fe_sign(q64.den) > 0) {
        double d2 = 2.0 * q64.den.v;
        cx = (double)v[0].x + q64.num.x.v / d2;
        cy = (double)v[0].y + q64.num.y.v / d2;
        cz = (double)v[0].z + q64.num.z.v / d2;
        ex = quot_err(q64.num.x, q64.den, 2.0);
        ey = quot_err(q64.num.y, q64.den, 2.0);
        ez = quot_err(q64.num.z, q64.den, 2.0);
        ok = true;
      } else {
        init_dd();
        if (fe_sign(qdd.den) > 0) {
          double d2 = 2.0 * qdd.den.v.hi;
          cx = (double)v[0].x + qdd.num.x.v.hi / d2;
          cy = (double)v[0].y + qdd.num.y.v.hi / d2;
          cz = (double)v[0].z + qdd.num.z.v.hi / d2;
          ex = quot_err(qdd.num.x, qdd.den, 2.0);
          ey = quot_err(qdd.num.y, qdd.den, 2.0);
          ez = quot_err(qdd.num.z, qdd.den, 2.0);
          ok = true;
        }
      }
      CellBox box;
      const GridP g = grid_of(S, vid[0]);
      if (!ok || !make_box(g, cx, cy, cz, ex, ey, ez, box)) {
        push_q(q2, e.item, 0xffffffffu);
      } else {
        bool dead = false;
        scan_box(g, S.cstart, box, [&](uint32_t p) {
          if (p == vid[0] || p == vid[1] || p == vid[2]) return false;
          int c = pair(p);
          if (c == 1) {
            dead = true;
            return true;
          }
          if (c == 2) push_q(q2, e.item, p);
          return false;
        });
        if (dead) flag[e.item] = 0;
      }
    }
  });
  sc.flush(S.ctr, KIND == 0 ? FAM_TRI_AC2 : FAM_EDGE_AC2);
}

/* The exact triangle / edge test's per-simplex part: its vertices, the
 * first vertex's scaled position and weight, den and num (what every pair
 * test of the simplex reuses). */
struct LowExact {
  uint32_t item = 0xffffffffu;
  uint32_t vid[3];
  Ball a;
  V3<XI> pa, num;
  XI wa, den;
};
template <int KIND>
__device__ __forceinline__ void low_exact_setup(const Stage& S,
                                                const uint32_t* ecol,
                                                const uint32_t* eleft,
                                                const uint32_t* tk,
                                                const uint32_t* tedge,
                                                uint32_t item, LowExact& T) {
  T.item = item;
  if (KIND == 0) {
    const uint32_t ed = tedge[item];
    T.vid[0] = eleft[ed];
    T.vid[1] = ecol[ed];
    T.vid[2] = tk[item];
  } else {
    T.vid[0] = eleft[item];
    T.vid[1] = ecol[item];
    T.vid[2] = 0xffffffffu;
  }
  const LdCW<kTPI> L{S.ex.S};
  T.a = S.ball[T.vid[0]];
  const Ball b = S.ball[T.vid[1]];
  T.pa = L.p(T.a);
  T.wa = L.w(T.a);
  if (KIND == 0) {
    tri_X(T.pa, T.wa, L.p(b), L.w(b), L.p(S.ball[T.vid[2]]),
          L.w(S.ball[T.vid[2]]), &T.den, &T.num);
  } else {
    V3<XI> u = vsub(L.p(b), T.pa);
    T.den = vdot(u, u);
    XI A = T.den + T.wa - L.w(b);
    T.num = vscale(A, u);
  }
}
/* One queued entry of simplex T: a (simplex, ball) pair, or a full exact
 * scan of its box (aux == ~0, a simplex whose centre no float stage could
 * certify).  Group-uniform; the group's first lane writes the flag. */
template <int KIND>
__device__ __forceinline__ void low_exact_entry(const Stage& S,
                                                const LowExact& T,
                                                uint32_t aux, uint8_t* flag,
                                                bool lead, unsigned& cnt) {
  const LdCW<kTPI> L{S.ex.S};
  if (aux != 0xffffffffu) {
    ++cnt;
    const Ball E = S.ball[aux];
    if (fe_sign(power_gap(T.pa, T.wa, L.p(E), L.w(E), T.den, T.num)) <= 0 &&
        lead)
      flag[T.item] = 0;
  } else if (fe_sign(T.den) > 0) {
    double d2 = 2.0 * cw_to_double(T.den, (KIND == 0 ? 4 : 2) * S.ex.S);
    int sh = (KIND == 0 ? 5 : 3) * S.ex.S;
    double cx = (double)T.a.x + cw_to_double(T.num.x, sh) / d2;
    double cy = (double)T.a.y + cw_to_double(T.num.y, sh) / d2;
    double cz = (double)T.a.z + cw_to_double(T.num.z, sh) / d2;
    double err = 1e-9 * (fabs(cx) + fabs(cy) + fabs(cz) + 1.0);
    CellBox box;
    bool viol = false;
    const GridP g = grid_of(S, T.vid[0]);
    if (make_box(g, cx, cy, cz, err, err, err, box)) {
      for (int z = box.lo[2]; z <= box.hi[2] && !viol; ++z)
        for (int y = box.lo[1]; y <= box.hi[1] && !viol; ++y) {
          int64_t row = (int64_t)g.cbase + (int64_t)g.ex * (y + (int64_t)g.ey * z);
          uint32_t bb = S.cstart[row + box.lo[0]],
                   en = S.cstart[row + box.hi[0] + 1];
          for (uint32_t p = bb; p < en; ++p) {
            if (p == T.vid[0] || p == T.vid[1] || (KIND == 0 && p == T.vid[2]))
              continue;
            ++cnt;
            const Ball E = S.ball[p];
            if (fe_sign(power_gap(T.pa, T.wa, L.p(E), L.w(E), T.den, T.num)) <=
                0) {
              viol = true;
              break;
            }
          }
        }
    }
    if (lead) flag[T.item] = viol ? 0 : 2;
  } else if (lead) {
    flag[T.item] = 0; /* degenerate simplex: cannot be an alpha simplex */
  }
}

/* exact triangle / edge pair (or full scan when aux == ~0); one item per
 * group of kTPI lanes */
template <int KIND>
__global__ __launch_bounds__(kBlock) void k_low_ac2_exact(
    Stage S, QRef q, const uint32_t* ecol,
    const uint32_t* eleft, const uint32_t* tk, const uint32_t* tedge,
    uint8_t* flag) {
  unsigned cnt = 0;
  const bool lead = Grp<kTPI>::idx() == 0;
  for_queue_grp<kTPI>(q, [&](uint32_t t) {
    /* the group's first lane decides, so the group stays together */
    if (!Grp<kTPI>::from(live_entry(q.at(t), flag) ? 1u : 0u, 0)) return;
    const QEntry e = q.at(t);
    LowExact T;
    low_exact_setup<KIND>(S, ecol, eleft, tk, tedge, e.item, T);
    low_exact_entry<KIND>(S, T, e.aux, flag, lead, cnt);
  });
  flush_ctr(S.ctr, CTR_STAGE + (KIND == 0 ? FAM_TRI_AC2 : FAM_EDGE_AC2) * 4 + 3,
            lead ? cnt : 0u);
}

/* The same over the entries sorted by simplex (item[], aux[]), kSortedRun
 * consecutive entries per group at a time: the pairs of one simplex (an
 * exact tie usually has several co-circular balls) share its den and num. */
template <int KIND>
__global__ __launch_bounds__(kBlock) void k_low_ac2_exact_sorted(
    Stage S, const uint32_t* item, const uint32_t* aux, uint32_t n,
    const uint32_t* ecol, const uint32_t* eleft, const uint32_t* tk,
    const uint32_t* tedge, uint8_t* flag) {
  unsigned cnt = 0;
  const bool lead = Grp<kTPI>::idx() == 0;
  const uint64_t groups = (uint64_t)gridDim.x * blockDim.x / kTPI;
  LowExact T;
  for (uint64_t c0 = (uint64_t)(gtid() / kTPI) * kSortedRun; c0 < n;
       c0 += groups * kSortedRun) {
    const uint32_t c1 = (uint32_t)min((uint64_t)n, c0 + kSortedRun);
    for (uint32_t k = (uint32_t)c0; k < c1; ++k) {
      const uint32_t it = item[k], ax = aux[k];
      const bool live = ax == 0xffffffffu || flag[it] != 0;
      if (!Grp<kTPI>::from(live ? 1u : 0u, 0)) continue;
      if (it != T.item)
        low_exact_setup<KIND>(S, ecol, eleft, tk, tedge, it, T);
      low_exact_entry<KIND>(S, T, ax, flag, lead, cnt);
    }
  }
  flush_ctr(S.ctr, CTR_STAGE + (KIND == 0 ? FAM_TRI_AC2 : FAM_EDGE_AC2) * 4 + 3,
            lead ? cnt : 0u);
}

/* ---------------- launchers ---------------- */
template <class R, bool FINE, bool MOL>
static void tet_hot3(const Stage& S, Cnt nk, const uint4* tets, uint8_t* flag,
                     Queue q1, uint32_t* rest, unsigned* nrest, uint32_t* next,
                     unsigned* nnext, int budget, PipeLaunch L) {
  dim3 gr((unsigned)((nk.cap + kBlock - 1) / kBlock)), gb((unsigned)S.bgrid);
  const uint32_t* items = nullptr;
  if (FINE && S.quick) { /* quick pass, then the certified scan of the rest */
    zero_u32(nnext, LSLOTS, L.s);
    hipLaunchKernelGGL((k_tet_ac2_quick<R, MOL>), gr, dim3(kBlock), 0, L.s, S,
                       nk, tets, flag, next, nnext);
    items = next;
  } else if (S.chunk) {
    zero_u32(nnext, LSLOTS, L.s);
    hipLaunchKernelGGL(k_tet_need, gr, dim3(kBlock), 0, L.s, S, nk, tets, flag,
                       next, nnext);
    items = next;
  }
  hipLaunchKernelGGL((k_tet_ac2_hot<R, false, FINE, MOL>), gr, dim3(kBlock), 0,
                     L.s, S, nk, items, nnext, tets, flag, q1, rest, nrest,
                     budget);
  if (budget > 0) {
    hipLaunchKernelGGL((k_tet_ac2_hot<R, true, FINE, MOL>), gb,
                       dim3((unsigned)S.bblock), 0, L.s, S, nk, items, nnext,
                       tets, flag, q1, rest, nrest, budget);
  }
}
template <class R, bool FINE>
static void tet_hot2(const Stage& S, Cnt nk, const uint4* tets, uint8_t* flag,
                     Queue q1, uint32_t* rest, unsigned* nrest, uint32_t* next,
                     unsigned* nnext, int budget, PipeLaunch L) {
  if (S.mol)
    tet_hot3<R, FINE, true>(S, nk, tets, flag, q1, rest, nrest, next, nnext,
                            budget, L);
  else
    tet_hot3<R, FINE, false>(S, nk, tets, flag, q1, rest, nrest, next, nnext,
                             budget, L);
}
__global__ __launch_bounds__(kBlock) void k_own_mark(
    Stage S, int kind, Cnt c_n, const uint32_t* ecol, const uint32_t* eleft,
    const uint32_t* tk, const uint32_t* tedge, uint8_t* in_flag) {
  const int64_t id = gtid();
  if (id >= c_n.get()) return;
  bool need;
  if (kind == 0) {
    const uint32_t e = tedge[id];
    need = owned(S, eleft[e]) || owned(S, ecol[e]);
  } else {
    need = owned(S, eleft[id]);
  }
  if (!need) in_flag[id] = 1;
}
void launch_own_mark(const Stage& S, int kind, Cnt n, const uint32_t* ecol,
                     const uint32_t* eleft, const uint32_t* tk,
                     const uint32_t* tedge, uint8_t* in_flag, PipeLaunch L) {
  if (!S.chunk || n.cap <= 0) return;
  hipLaunchKernelGGL(k_own_mark, dim3((unsigned)((n.cap + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, L.s, S, kind, n, ecol, eleft, tk, tedge,
                     in_flag);
}
void launch_tet_ac2_hot(const Stage& S, Cnt nk, const uint4* tets,
                        uint8_t* flag, Queue q1, uint32_t* rest,
                        unsigned* nrest, uint32_t* next, unsigned* nnext,
                        int budget, PipeLaunch L) {
  if (nk.cap <= 0) return;
  if constexpr (!kBallF64) {
    if (S.staged) {
      if (S.fine)
        tet_hot2<float, true>(S, nk, tets, flag, q1, rest, nrest, next, nnext,
                              budget, L);
      else
        tet_hot2<float, false>(S, nk, tets, flag, q1, rest, nrest, next, nnext,
                               budget, L);
      return;
    }
  }
  if (S.fine)
    tet_hot2<double, true>(S, nk, tets, flag, q1, rest, nrest, next, nnext,
                           budget, L);
  else
    tet_hot2<double, false>(S, nk, tets, flag, q1, rest, nrest, next, nnext,
                            budget, L);
}
void launch_tet_ac2_esc(const Stage& S, QRef q,
                        const uint4* tets, uint8_t* flag, Queue q2,
                        PipeLaunch L) {
  if (!q.launch_items()) return;
  hipLaunchKernelGGL(k_tet_ac2_esc, qgrid(q, kEscBlocks),
                     dim3(kBlock), 0, L.s, S, q, tets, flag, q2);
}
void launch_tet_ac2_exact(const Stage& S, QRef q,
                          const uint4* tets, const uint32_t* perm,
                          uint8_t* flag, PipeLaunch L) {
  if (!q.launch_items()) return;
  hipLaunchKernelGGL(k_tet_ac2_exact, qgrid_grp(q, kExactBlocks, kTPI),
                     dim3(kBlock), 0, L.s, S, q, tets, perm, flag);
}
void launch_view_split(const QView& v, uint32_t* item, uint32_t* aux,
                       PipeLaunch L) {
  const uint32_t n = v.size();
  if (!n) return;
  hipLaunchKernelGGL(k_view_split, dim3((n + kBlock - 1) / kBlock),
                     dim3(kBlock), 0, L.s, v, item, aux);
}
void launch_tet_ac2_exact_sorted(const Stage& S, const uint32_t* item,
                                 const uint32_t* aux, uint32_t n,
                                 const uint4* tets, const uint32_t* perm,
                                 uint8_t* flag, PipeLaunch L) {
  if (!n) return;
  const uint64_t groups = (n + kSortedRun - 1) / kSortedRun;
  const uint64_t b = (groups * kTPI + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_tet_ac2_exact_sorted,
                     dim3((unsigned)(b < 65536 ? b : 65536)), dim3(kBlock), 0,
                     L.s, S, item, aux, n, tets, perm, flag);
}
template <class R, int KIND, bool FINE, bool MOL>
static void low_hot4(const Stage& S, Cnt nitems, const uint32_t* ecol,
                     const uint32_t* eleft, const uint32_t* tk,
                     const uint32_t* tedge, const uint8_t* in_flag,
                     uint8_t* flag, Queue q1, uint32_t* rest, unsigned* nrest,
                     int budget, PipeLaunch L) {
  dim3 gr((unsigned)((nitems.cap + kBlock - 1) / kBlock)), gb((unsigned)S.bgrid);
  hipLaunchKernelGGL((k_low_ac2_hot<R, KIND, false, FINE, MOL>), gr,
                     dim3(kBlock), 0, L.s, S, nitems, ecol, eleft, tk, tedge,
                     in_flag, flag, q1, rest, nrest, budget);
  if (budget > 0) {
    hipLaunchKernelGGL((k_low_ac2_hot<R, KIND, true, FINE, MOL>), gb,
                       dim3((unsigned)S.bblock), 0, L.s, S, nitems, ecol, eleft,
                       tk, tedge, in_flag, flag, q1, rest, nrest, budget);
  }
}
template <class R, int KIND, bool FINE>
static void low_hot3(const Stage& S, Cnt nitems, const uint32_t* ecol,
                     const uint32_t* eleft, const uint32_t* tk,
                     const uint32_t* tedge, const uint8_t* in_flag,
                     uint8_t* flag, Queue q1, uint32_t* rest, unsigned* nrest,
                     int budget, PipeLaunch L) {
  if (S.mol)
    low_hot4<R, KIND, FINE, true>(S, nitems, ecol, eleft, tk, tedge, in_flag,
                                  flag, q1, rest, nrest, budget, L);
  else
    low_hot4<R, KIND, FINE, false>(S, nitems, ecol, eleft, tk, tedge, in_flag,
                                   flag, q1, rest, nrest, budget, L);
}
template <class R, int KIND>
static void low_hot2(const Stage& S, Cnt nitems, const uint32_t* ecol,
                     const uint32_t* eleft, const uint32_t* tk,
                     const uint32_t* tedge, const uint8_t* in_flag,
                     uint8_t* flag, Queue q1, uint32_t* rest, unsigned* nrest,
                     int budget, PipeLaunch L) {
  if (S.fine)
    low_hot3<R, KIND, true>(S, nitems, ecol, eleft, tk, tedge, in_flag, flag,
                            q1, rest, nrest, budget, L);
  else
    low_hot3<R, KIND, false>(S, nitems, ecol, eleft, tk, tedge, in_flag, flag,
                             q1, rest, nrest, budget, L);
}
void launch_low_ac2_hot(const Stage& S, int kind, Cnt nitems,
                        const uint32_t* ecol, const uint32_t* eleft,
                        const uint32_t* tk, const uint32_t* tedge,
                        const uint8_t* in_flag, uint8_t* flag, Queue q1,
                        uint32_t* rest, unsigned* nrest, int budget,
                        PipeLaunch L) {
  if (nitems.cap <= 0) return;
  if constexpr (!kBallF64) {
    if (S.staged) {
      if (kind == 0)
        low_hot2<float, 0>(S, nitems, ecol, eleft, tk, tedge, in_flag, flag, q1,
                           rest, nrest, budget, L);
      else
        low_hot2<float, 1>(S, nitems, ecol, eleft, tk, tedge, in_flag, flag, q1,
                           rest, nrest, budget, L);
      return;
    }
  }
  if (kind == 0)
    low_hot2<double, 0>(S, nitems, ecol, eleft, tk, tedge, in_flag, flag, q1,
                        rest, nrest, budget, L);
  else
    low_hot2<double, 1>(S, nitems, ecol, eleft, tk, tedge, in_flag, flag, q1,
                        rest, nrest, budget, L);
}
void launch_low_ac2_esc(const Stage& S, int kind, QRef q,
                        const uint32_t* ecol, const uint32_t* eleft,
                        const uint32_t* tk, const uint32_t* tedge,
                        uint8_t* flag, Queue q2, PipeLaunch L) {
  if (!q.launch_items()) return;
  dim3 gr = qgrid(q, kEscBlocks);
  if (kind == 0)
    hipLaunchKernelGGL(k_low_ac2_esc<0>, gr, dim3(kBlock), 0, L.s, S, q,
                       ecol, eleft, tk, tedge, flag, q2);
  else
    hipLaunchKernelGGL(k_low_ac2_esc<1>, gr, dim3(kBlock), 0, L.s, S, q,
                       ecol, eleft, tk, tedge, flag, q2);
}
void launch_low_ac2_exact(const Stage& S, int kind, QRef q, const uint32_t* ecol,
                          const uint32_t* eleft, const uint32_t* tk,
                          const uint32_t* tedge, uint8_t* flag, PipeLaunch L) {
  if (!q.launch_items()) return;
  dim3 gr = qgrid_grp(q, kExactBlocks, kTPI);
  if (kind == 0)
    hipLaunchKernelGGL(k_low_ac2_exact<0>, gr, dim3(kBlock), 0, L.s, S, q,
                       ecol, eleft, tk, tedge, flag);
  else
    hipLaunchKernelGGL(k_low_ac2_exact<1>, gr, dim3(kBlock), 0, L.s, S, q,
                       ecol, eleft, tk, tedge, flag);
}
void launch_low_ac2_exact_sorted(const Stage& S, int kind, const uint32_t* item,
                                 const uint32_t* aux, uint32_t n,
                                 const uint32_t* ecol, const uint32_t* eleft,
                                 const uint32_t* tk, const uint32_t* tedge,
                                 uint8_t* flag, PipeLaunch L) {
  if (!n) return;
  const uint64_t groups = (n + kSortedRun - 1) / kSortedRun;
  const uint64_t b = (groups * kTPI + kBlock - 1) / kBlock;
  const dim3 gr((unsigned)(b < 65536 ? b : 65536));
  if (kind == 0)
    hipLaunchKernelGGL(k_low_ac2_exact_sorted<0>, gr, dim3(kBlock), 0, L.s, S,
                       item, aux, n, ecol, eleft, tk, tedge, flag);
  else
    hipLaunchKernelGGL(k_low_ac2_exact_sorted<1>, gr, dim3(kBlock), 0, L.s, S,
                       item, aux, n, ecol, eleft, tk, tedge, flag);
}

int sos_table_host(int32_t* out3, int cap) {
  for (int t = 0; t < PAC_SOS_TERMS && t < cap; ++t)
    for (int c = 0; c < 3; ++c) out3[3 * t + c] = pac_sos_table[t][c];
  return PAC_SOS_TERMS;
}

}  // namespace PAC_V
}  // namespace pac
