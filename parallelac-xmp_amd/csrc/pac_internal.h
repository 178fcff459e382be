/*
 * pac_internal.h -- shared between the host engine and the kernel
 * translation units.  Not part of the C ABI.  Types and launchers that do
 * not depend on the ball type live in namespace pac; the rest in the
 * compilation's ball-type namespace (pac_variant.h).
 */
#ifndef PAC_INTERNAL_H
#define PAC_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pac_arith.h"

namespace pac {

constexpr int kBlock = 256;      /* threads per block, 4 wave64s */
constexpr int kScanItems = 8;    /* items per thread in scans / radix */
constexpr int kTile = kBlock * kScanItems;
constexpr int kMaxFine = 8;      /* largest x subdivision of a scan grid */

/* Reference grid (AlphaComplexParallel.cu:1120-1124) of one molecule plus
 * the stencil half-widths used for enumeration and emptiness scans.  In a
 * batch every molecule has its own grid; its cells occupy the global cell
 * ids [cbase, cbase + ex*ey*ez), so one stable sort by global cell id keeps
 * the molecules contiguous and in input order. */
struct GridP {
  float minx, miny, minz, step;
  int ex, ey, ez;
  int k_enum, k_ac2;
  uint32_t cbase;
  double reach;   /* sqrt(max w + alpha): violators lie this close to the centre */
  double margin;  /* FP32 cell-key rounding, in length units */
  double inv_step; /* 1 / cell size along x */
  double inv_yz;   /* 1 / cell size along y and z (a fine scan grid splits
                      cells along x only; the reference grid is cubic) */
  float xsplit;    /* x cells per y-z cell width (inv_step / inv_yz) */
  double maxw;    /* max r^2 of the molecule; < 0: keep the global reach */
};

/* raw-input bounds reduced on the device (values of the input type) */
struct Bounds {
  double minx, miny, minz, maxx, maxy, maxz;
  double max_raw_r, max_r; /* raw radius, and r_raw + solvent */
  double max_abs;          /* max |coordinate| */
  int shift;               /* max over inputs of -lsb exponent (>= 0) */
  int nonfinite;           /* count of non-finite inputs */
};

/* Item count of a launch.  The host sizes the grid for `cap` items; when
 * `d` is set the actual count is *d, read on the device, which lets a whole
 * pipeline run (and be captured as one graph) without reading counts back.
 * Converts from a plain count. */
struct Cnt {
  int64_t cap;
  const uint64_t* d;
  __host__ __device__ Cnt(int64_t n = 0) : cap(n), d(nullptr) {}
  __host__ __device__ Cnt(int64_t c, const uint64_t* p) : cap(c), d(p) {}
  __device__ int64_t get() const {
    if (!d) return cap;
    const int64_t v = (int64_t)*d;
    return v < cap ? v : cap;
  }
};

struct QEntry {
  uint32_t item;  /* candidate / simplex index */
  uint32_t aux;   /* segment (ball / edge / triangle) or ball index */
};

/* counters layout (unsigned long long) */
enum {
  CTR_STAGE = 0,        /* [6][4] stage evaluations */
  CTR_SOS = 24,
  CTR_PAIRS = 25,       /* [3] tet, tri, edge pairs */
  CTR_CAND = 28,        /* [3] candidates: (unused), triangles, tets */
  CTR_COUNT = 31
};
constexpr int CTR_SLOTS = 32;   /* copies of each counter (see flush_ctr) */
enum { FAM_EDGE = 0, FAM_TRI = 1, FAM_TET = 2, FAM_TET_AC2 = 3,
       FAM_TRI_AC2 = 4, FAM_EDGE_AC2 = 5 };

/* ---- primitives (k_prims.hip) ---- */
/* exclusive scan; out has n+1 entries, out[n] = total */
/* total64 (optional, device): also receives the total */
void scan_u32_to_u64(const uint32_t* in, uint64_t* out, Cnt n, void* tmp,
                     size_t tmp_bytes, hipStream_t s,
                     uint64_t* total64 = nullptr);
void scan_u64_to_u64(const uint64_t* in, uint64_t* out, Cnt n, void* tmp,
                     size_t tmp_bytes, hipStream_t s,
                     uint64_t* total64 = nullptr);
/* flags (nonzero = 1) -> u32 positions; n < 2^32 */
void scan_u8flag_to_u32(const uint8_t* in, uint32_t* out, Cnt n, void* tmp,
                        size_t tmp_bytes, hipStream_t s,
                        uint64_t* total64 = nullptr);
size_t scan_tmp_bytes(int64_t n);
/* tile offsets of flags (nonzero = 1) for the compactions below, inside
 * tmp (scan_tmp_bytes(n)); the count of nonzero flags at *total64 */
const uint64_t* flag_tile_offsets(const uint8_t* in, Cnt n, void* tmp,
                                  hipStream_t s, uint64_t* total64);
/* stable LSD radix sort of (key, val) by the low `bits` bits of key */
void radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt,
                      uint32_t* vals_alt, int64_t n, int bits, void* tmp,
                      size_t tmp_bytes, hipStream_t s, bool* result_in_alt);
size_t radix_tmp_bytes(int64_t n);

/* small fills on a stream (kernels, not hipMemsetAsync: they may be
 * captured into graphs) */
void fill_u8(uint8_t* p, int64_t n, uint8_t v, hipStream_t s);
void zero_u32(unsigned* p, int n, hipStream_t s);
void zero_u64(unsigned long long* p, int64_t n, hipStream_t s);

/* ---- chunked driver (k_chunk.hip) ---- */
/* rank[sidx[p]] = p */
void launch_rank_of(const uint32_t* sidx, int64_t n, uint32_t* rank,
                    hipStream_t s);
/* flag[i] = lo <= rank[i] < hi */
void launch_rank_flag(const uint32_t* rank, int64_t n, uint32_t lo,
                      uint32_t hi, uint8_t* flag, hipStream_t s);
/* flagged balls in input order (pos: exclusive scan of flag, m = total):
 * out[pos[i]] = raw[i], sid[pos[i]] = i */
void launch_select_balls(const float4* raw, const uint8_t* flag,
                         const uint32_t* pos, int64_t n, int64_t m,
                         float4* out, uint32_t* sid, hipStream_t s);
void launch_select_balls(const double4* raw, const uint8_t* flag,
                         const uint32_t* pos, int64_t n, int64_t m,
                         double4* out, uint32_t* sid, hipStream_t s);
/* out[0..1] = the run of rows of a reference-order list (width w, sub-input
 * ids) whose first vertex has rank[sid[.]] in [lo, hi) */
void launch_row_bounds(const int32_t* list, int64_t rows, int w,
                       const uint32_t* sid, const uint32_t* rank, uint32_t lo,
                       uint32_t hi, uint64_t* out, hipStream_t s);
/* out[j] = sid[in[j]] */
void launch_map_ids(const int32_t* in, int64_t cnt, const uint32_t* sid,
                    int32_t* out, hipStream_t s);

struct PipeLaunch {
  hipStream_t s;
  int block;
};

/* Queues of undecided items between the stages of the precision ladder.
 * A producer appends to one of QSLOTS sub-queues (picked by its block) so
 * that many simultaneous appends do not serialize on one counter; sub-queue
 * k holds entries [k*cap, k*cap + n[k]). */
constexpr int QSLOTS = 32;
/* Item lists appended to a block at a time (block_append) by
 * one-item-per-thread kernels -- the simplices an emptiness scan's first
 * phase leaves, the tets a quick rejection pass leaves -- are spread the
 * same way over LSLOTS sub-lists: block b appends to sub-list b % LSLOTS,
 * entries [k * lcap, k * lcap + n[k]) with n[k] <= lcap, lcap =
 * list_slot_cap(the producer's blocks).  One counter for the ~10^5 blocks
 * of a launch serialized their device-scope atomics (on the 100k globule,
 * 0.45 ms of a 0.88 ms tet emptiness first phase). */
constexpr int LSLOTS = 32;
inline uint32_t list_slot_cap(int64_t blocks) {
  return (uint32_t)((blocks + LSLOTS - 1) / LSLOTS * kBlock);
}
/* queue counter words of a run: the hot and escalation sub-queue counters
 * (2 QSLOTS), then the counters of the list of simplices left for the
 * second emptiness phase and of the list of tets a quick rejection pass
 * left (2 LSLOTS) */
constexpr int QN_WORDS = 2 * QSLOTS + 2 * LSLOTS;
struct Queue {
  QEntry* q;
  unsigned* n;   /* QSLOTS counters */
  uint32_t cap;  /* per sub-queue */
};
/* consumer view: entry i of the concatenated sub-queues */
struct QView {
  const QEntry* q;
  uint32_t cap;
  uint32_t off[QSLOTS + 1]; /* exclusive prefix of the sub-queue counts */
  __host__ __device__ uint32_t size() const { return off[QSLOTS]; }
  __device__ QEntry at(uint32_t i) const {
    int lo = 0, hi = QSLOTS - 1; /* last k with off[k] <= i */
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (off[mid] <= i) lo = mid;
      else hi = mid - 1;
    }
    return q[(uint64_t)lo * cap + (i - off[lo])];
  }
};
/* what a consumer kernel gets: the view by value (built on the host), or a
 * pointer to one built on the device (graph-captured pipelines) and the
 * number of items its grid covers */
struct QRef {
  QView v;
  const QView* dv;
  uint32_t bound;
  __device__ uint32_t size() const { return dv ? dv->size() : v.size(); }
  __device__ QEntry at(uint32_t i) const { return dv ? dv->at(i) : v.at(i); }
  uint32_t launch_items() const { return dv ? bound : v.size(); }
};
/* grid of a queue consumer: host views (size known) one entry per
 * thread; device views (size known only on the device) at most maxb
 * blocks striding over the entries, so that a queue sized for many
 * entries but holding few costs few idle blocks */
constexpr unsigned kEscBlocks = 1024, kExactBlocks = 512;
inline dim3 qgrid(const QRef& q, unsigned maxb) {
  const unsigned b = (q.launch_items() + kBlock - 1) / kBlock;
  if (!q.dv) return dim3(b ? b : 1); /* host view: one entry per thread */
  return dim3(b < maxb ? (b ? b : 1) : maxb);
}
/* the same for consumers that take one entry per group of tpi lanes */
inline dim3 qgrid_grp(const QRef& q, unsigned maxb, unsigned tpi) {
  const uint64_t b = ((uint64_t)q.launch_items() * tpi + kBlock - 1) / kBlock;
  if (!q.dv) return dim3(b ? (unsigned)b : 1u);
  return dim3(b < maxb ? (b ? (unsigned)b : 1u) : maxb);
}

/* ---- recorded pipelines (k_prims.hip) ----
 * Fixed-capacity runs keep every count on the device; overflows and a
 * stale plan are reported through status bits. */
enum { REC_CAP = 0,    /* bits 0..8: a count exceeded its capacity */
       REC_Q1 = 9,     /* hot queue overflow */
       REC_Q2 = 10,    /* escalation queue overflow */
       REC_STALE = 11  /* the input no longer matches the recorded plan */ };
/* view of QSLOTS sub-queues (counters cleared after reading, and the
 * LSLOTS list counters at also);
 * fresh (an eager ladder's first view): status / fill overwritten, and
 * the view left empty when a sub-queue overflowed (the ladder reruns);
 * else status / fill merged and the view clamped to the capacity */
void launch_make_view(const QEntry* q, uint32_t cap, unsigned* cnt, QView* v,
                      unsigned* status, unsigned bit, unsigned* fill,
                      unsigned* also, bool fresh, hipStream_t s);
void launch_cap_check(const uint64_t* v, uint64_t cap, unsigned* status,
                      unsigned bit, hipStream_t s);
/* eager ladder state (u32 words): [kLadStatus] status, [kLadFill] fill of
 * the first view, [kLadCount] the count (u64), [kLadCnt, + QSLOTS) the
 * last queue's sub-queue counts */
constexpr int kLadStatus = 0, kLadFill = 1, kLadCount = 4, kLadCnt = 6;
constexpr int kLadWords = kLadCnt + QSLOTS;
void launch_ladder_state(unsigned* cnt, const uint64_t* count, unsigned* lad,
                         hipStream_t s);
void launch_plan_check(const Bounds* now, const Bounds* ref, int64_t m,
                       int scale, unsigned* status, unsigned bit,
                       hipStream_t s);

struct CandCtx {
  const uint32_t *eleft, *ecol, *tk, *tedge;
  const uint4* ctet; /* tet candidates (i, j, k, l); x = ~0: unused slot */
};

/* ---- compactions by flags over tile offsets (k_prims.hip) ----
 * offs from flag_tile_offsets of the same flags; ocap: capacity of the
 * output (entries past it are dropped) */
void launch_compact_tets(Cnt n, const uint4* in, const uint8_t* flag,
                         const uint64_t* offs, uint4* out, PipeLaunch L,
                         uint64_t ocap = ~0ull);
void launch_compact_ids(Cnt n, const uint8_t* flag, const uint64_t* offs,
                        uint32_t* out, PipeLaunch L, uint64_t ocap = ~0ull);
void launch_compact_u8(Cnt n, const uint8_t* flag, const uint64_t* offs,
                       const uint8_t* val, uint8_t* out, PipeLaunch L,
                       uint64_t ocap = ~0ull);
/* kept alpha edges as (perm[left], perm[col]) pairs */
void launch_emit_edges(Cnt ne, const uint32_t* ecol, const uint32_t* eleft,
                       const uint8_t* flag, const uint64_t* offs,
                       const uint32_t* perm, int32_t* out, PipeLaunch L,
                       uint64_t ocap = ~0ull);

/* ---- closure + emission (k_closure.hip) ---- */
void launch_or_flags(Cnt n, const uint8_t* a, const uint8_t* b, uint8_t* out,
                     PipeLaunch L);
void launch_mark_faces(Cnt nk, const uint4* ftets, const uint64_t* erow,
                       const uint32_t* ecol, const uint64_t* trow,
                       const uint32_t* tk, uint8_t* tri_in, PipeLaunch L);
void launch_mark_edges(Cnt nt, const uint32_t* ftri, const uint32_t* tedge,
                       const uint32_t* teik, const uint32_t* tejk,
                       uint8_t* edge_in, PipeLaunch L);
void launch_emit_tets(Cnt nk, const uint4* ftets, const uint32_t* perm,
                      int32_t* out, PipeLaunch L);
void launch_emit_tris(Cnt nt, const uint32_t* ftri, const uint32_t* ecol,
                      const uint32_t* eleft, const uint32_t* tk,
                      const uint32_t* tedge, const uint32_t* perm,
                      int32_t* out, PipeLaunch L);

/* ---- outputs (k_output.hip) ---- */
/* Filtration keys of all simplices, key[0, n + E + T + K) (balls, edges,
 * triangles, tets; order-preserving 64-bit images of the values): lists in
 * reference order, gab[0..1] the own-value flags of edges / triangles,
 * rank[original id] = cell-sorted position; *missing counts faces not
 * found among the lists */
void birth_keys(const int32_t* const lists[3], const int64_t counts[4],
                const double4* balls, const uint8_t* const gab[2],
                const uint32_t* rank, uint64_t* key, unsigned* missing,
                hipStream_t s);
/* values of m keys; *missing counts keys that never got a value */
void birth_values(int64_t m, const uint64_t* key, double* out,
                  unsigned* missing, hipStream_t s);
/* all m keys sorted by value (stable): dim / index in its list / value;
 * a, b, c, e: m u32 scratch each; rtmp: radix_tmp_bytes(m) */
void filtration_order(int64_t m, const uint64_t* key, const int64_t counts[4],
                      int32_t* dim, int64_t* index, double* value,
                      uint32_t* a, uint32_t* b, uint32_t* c, uint32_t* e,
                      void* rtmp, size_t rtb, hipStream_t s);
void launch_rho(int dim, int64_t m, const int32_t* rows, const double4* balls,
                double* rho, hipStream_t s);
/* adjoint: per-(simplex, vertex) contributions at slots slot0 + (dim+1)*id + q
 * with key = vertex, val = slot; then per-ball sums over vstart ranges of
 * the slots sorted by vertex */
void launch_adj_contrib(int dim, int64_t m, const int32_t* rows,
                        const double4* balls, const double* g, int64_t slot0,
                        uint32_t* key, uint32_t* val, double4* contrib,
                        hipStream_t s);
void launch_adj_gather(int64_t n, const uint32_t* vstart, const uint32_t* sval,
                       const double4* contrib, double* dxyz, double* dr,
                       hipStream_t s);
void canonical_rows(int K, int64_t m, int32_t* rows, int idbits,
                    uint32_t* idx, uint32_t* key, uint32_t* idx2,
                    uint32_t* key2, void* rtmp, size_t rtb, int32_t* copy,
                    hipStream_t s);

/* batch: first row of each molecule in a list of m rows of K original ids
 * grouped by molecule (moff: first original id of each molecule);
 * out[nmol+1] */
void launch_mol_rows(const int32_t* rows, Cnt m, int K,
                     const uint32_t* moff, int64_t nmol, uint64_t* out,
                     hipStream_t s);



}  // namespace pac

namespace pac {
namespace PAC_V {
/* ---- binning (k_enum.hip) ---- */

void launch_bounds(const Ball* raw, int64_t n, Real solvent, Bounds* part,
                   Bounds* out, PipeLaunch L);
/* per-molecule bounds of a batch: one block per molecule, moff[nmol+1] */
void launch_bounds_mol(const Ball* raw, const uint32_t* moff, int64_t nmol,
                       Real solvent, Bounds* out, PipeLaunch L);
/* mol[i] = molecule of ball i (same for input and cell-sorted order) */
void launch_mol_ids(const uint32_t* moff, int64_t nmol, int64_t n,
                    uint32_t* mol, PipeLaunch L);
/* global cell key; grids/mol == nullptr: the single grid g */
void launch_cell_key(const Ball* raw, int64_t n, GridP g,
                     const GridP* grids, const uint32_t* mol, uint32_t* key,
                     uint32_t* idx, PipeLaunch L);
void launch_gather(const Ball* raw, const uint32_t* sidx,
                   const uint32_t* skey, int64_t n, Real solvent,
                   Ball* ball, uint32_t* perm, uint32_t* cellk,
                   PipeLaunch L);
void launch_cell_starts(const uint32_t* cellk, int64_t n, int64_t G,
                        uint32_t* cstart, PipeLaunch L);
/* fine scan grid: reference cells split into kx (<= kMaxFine) cells
 * along x; fball / fid the balls and their ids in fine-cell order,
 * fstart[G kx + 1] the fine cell starts; g / grids: the fine grids */
void launch_fine_split(const Ball* ball, const uint32_t* cstart, int64_t G,
                       int kx, GridP g, const GridP* grids,
                       const uint32_t* mol, Ball* fball, uint32_t* fid,
                       uint32_t* fstart, PipeLaunch L);

struct Stage {
  const Ball* ball;
  int64_t n;
  GridP g;              /* the grid of a single molecule */
  const GridP* grids;   /* batch: grid of each molecule (else nullptr) */
  const uint32_t* mol;  /* batch: molecule of each ball (else nullptr) */
  const uint32_t* cellk;
  const uint32_t* cstart;
  Real alpha;
  ExactScale ex;
  unsigned long long* ctr;
  int staged;
  /* second emptiness phase: a persistent launch of bgrid blocks of bblock
   * threads whose waves take bchunk entries of the left-over list at a
   * time (for_work) */
  int bgrid;
  int bblock;
  unsigned bchunk;
  /* entries per sub-list of the left-over and quick-pass lists */
  uint32_t lcap;
  /* exact cell policy: the emptiness scans run over a finer grid of their
   * own (reference cells split along x), its balls in fine-cell order */
  int fine;
  int quick;             /* fine grid: quick rejection pass over the tets */
  GridP sg;              /* fine grid of a single molecule */
  const GridP* sgrids;   /* batch: fine grid of each molecule */
  const Ball* fball;     /* balls in fine-cell order */
  const uint32_t* fid;   /* their ball (cell-sorted) ids */
  const uint32_t* fstart;
  /* chunk != 0: a chunk's sub-input, whose output keeps only the simplices
   * with lowest vertex in [own_lo, own_hi) (cell-sorted ids).  Emptiness
   * tests that cannot change one of those are skipped: tets with none of
   * their three lowest vertices owned are rejected, triangles with neither
   * of their two lowest owned and edges with the lower one not owned pass
   * untested (none of them has an owned face or edge). */
  int chunk;
  uint32_t own_lo, own_hi;
};



/* ---- enumeration + filters (k_enum.hip) ----
 * Candidates of each dimension are (segment, new vertex) pairs: segment =
 * ball (edges), alpha edge (triangles), alpha triangle (tets). */

void launch_edge_count(const Stage& S, uint32_t* cnt, PipeLaunch L);
/* edge candidates of every ball, each filtered by the first precision
 * stage in the same kernel: ball i's kept and undecided candidates (j, i)
 * with their flags, in order, compacted from its first slot off[i] on and
 * their number to bcnt[i]; undecided ones queued on q1 by position (aux =
 * i) */
void launch_edge_enum_filter(const Stage& S, const uint64_t* off,
                             uint32_t* cval, uint32_t* cseg, uint8_t* flag,
                             uint32_t* bcnt, Queue q1, PipeLaunch L,
                             uint64_t lim = ~0ull);
/* after the ladder: keep[i] = kept candidates of ball i; then, with erow
 * their exclusive scan (the alpha-edge CSR rows), the kept edges to
 * (ecol, eleft) in order */
void launch_seg_keep(int64_t n, const uint64_t* off, const uint32_t* bcnt,
                     const uint8_t* flag, uint32_t* keep, PipeLaunch L);
void launch_edge_copy(int64_t n, const uint64_t* off, const uint32_t* bcnt,
                      const uint32_t* cval, const uint8_t* flag,
                      const uint64_t* erow, uint32_t* ecol, uint32_t* eleft,
                      PipeLaunch L, uint64_t ocap = ~0ull);
/* slot capacity of item q = row[owner[q] + 1] - q - 1 */
void launch_tail_caps(Cnt m, const uint64_t* row, const uint32_t* owner,
                      uint32_t* cap, PipeLaunch L);
/* tet slots of each alpha triangle: launch_tail_caps over its edge's row,
 * or none for the triangles a chunk sub-input does not need */
void launch_tet_caps(const Stage& S, Cnt m, const uint64_t* trow,
                     const uint32_t* tedge, const uint32_t* eleft,
                     const uint32_t* ecol, const uint32_t* tk, uint32_t* cap,
                     PipeLaunch L);
/* exact triangle count per alpha edge */
void launch_tri_count(Cnt ne, const uint64_t* erow, const uint32_t* ecol,
                      const uint32_t* eleft, uint32_t* cnt, PipeLaunch L);
/* triangle candidates of the alpha edges, each filtered by the first
 * precision stage in the same kernel: enumeration block b (256 alpha
 * edges) writes its kept and undecided candidates (k, edge, edge(i,k),
 * edge(j,k)) with their flags, in order, compacted from its first slot
 * off[256 b] on and their number to bcnt[b]; undecided ones are queued on
 * q1 by position (aux = their edge) */
void launch_tri_enum_filter(const Stage& S, Cnt ne, const uint64_t* erow,
                            const uint32_t* ecol, const uint32_t* eleft,
                            const uint64_t* off, uint32_t* cval,
                            uint32_t* cseg, uint32_t* cea, uint32_t* ceb,
                            uint8_t* flag, uint32_t* bcnt, uint32_t* bkeep,
                            Queue q1, PipeLaunch L, uint64_t lim = ~0ull);
/* after the ladder (bkeep by launch_tet_block_keep over the ne alpha
 * edges, bout its exclusive scan): the kept triangles of each block to
 * (tk, tedge, teik, tejk) in order, the alpha-triangle row start of every
 * alpha edge, trow[0 .. ne], and (tpair, if given) each edge's tet slots,
 * L (L - 1) / 2 for its L triangles (each one's row tail) */
void launch_tri_block_copy(Cnt ne, const uint64_t* off, const uint32_t* bcnt,
                           const uint32_t* cval, const uint32_t* cseg,
                           const uint32_t* cea, const uint32_t* ceb,
                           const uint8_t* flag, const uint64_t* bout,
                           uint32_t* tk, uint32_t* tedge, uint32_t* teik,
                           uint32_t* tejk, uint64_t* trow, uint64_t* tpair,
                           PipeLaunch L, uint64_t ocap = ~0ull);
/* tet candidates of the alpha triangles, each filtered by the first
 * precision stage in the same kernel; enumeration block b (256 alpha
 * triangles) writes its kept and undecided tets (i, j, k, l) with their
 * flags, in order, compacted from its first slot on (bbase[b]) and their
 * number to bcnt[b]; undecided ones are queued on q1 by position.  Slot
 * offsets per triangle from off[], or (off null) from roff, the exclusive
 * scan of launch_tri_block_copy's tpair (each triangle's slots its row's
 * tail: not for chunk sub-inputs, whose unowned triangles get none) */
void launch_tet_enum_filter(const Stage& S, Cnt nt, const uint64_t* trow,
                            const uint32_t* tk, const uint32_t* tedge,
                            const uint32_t* teik, const uint32_t* tejk,
                            const uint32_t* eleft, const uint32_t* ecol,
                            const uint64_t* off, const uint64_t* roff,
                            uint64_t* bbase, uint4* ctet, uint8_t* flag,
                            uint32_t* bcnt, uint32_t* bkeep, Queue q1,
                            PipeLaunch L, uint64_t lim = ~0ull);
/* The triangle and tet enumerations also write bkeep[b]: bcnt[b] when the
 * block queued no undecided candidate (the ladder cannot drop any of its
 * entries), else kRecount. */
constexpr uint32_t kRecount = 0xffffffffu;
/* after the ladder: kept tets per enumeration block (ceil(nt.cap / 256)
 * entries; blocks marked kRecount recounted from their flags), then, with
 * bout their exclusive scan, the kept tets in order; block b's entries
 * start at off[b * stride] (per-item offsets: stride 256; per-block
 * bases: 1) */
void launch_tet_block_keep(Cnt nt, const uint64_t* off, int stride,
                           const uint32_t* bcnt, const uint8_t* flag,
                           uint32_t* bkeep, PipeLaunch L);
void launch_tet_block_copy(Cnt nt, const uint64_t* off, int stride,
                           const uint32_t* bcnt, const uint4* ctet,
                           const uint8_t* flag, const uint64_t* bout,
                           uint4* tets, PipeLaunch L, uint64_t ocap = ~0ull);
/* escalation and exact stages of the filters (the first stages are fused
 * with the enumerations): kind 1 edges, 2 triangles, 3 tets */
void launch_filter_esc(const Stage& S, int kind, CandCtx C, QRef q, const uint32_t* cval, uint8_t* flag,
                       Queue q2, PipeLaunch L);
void launch_filter_exact(const Stage& S, int kind, CandCtx C, QRef q, const uint32_t* cval, uint8_t* flag,
                         PipeLaunch L);

/* ---- emptiness (k_ac2.hip) ---- */
/* rest / nrest: scratch list of simplices left for phase B (capacity =
 * number of items); budget = phase A pair budget (0: single phase) */
/* next / nnext: scratch list (capacity = number of tets) of the tets a
 * quick rejection pass leaves for the certified scan (fine scan grid) */
/* chunk sub-input: in_flag[id] = 1 (pass untested) for the triangles
 * (kind 0) / edges (kind 1) no owned simplex depends on */
void launch_own_mark(const Stage& S, int kind, Cnt n, const uint32_t* ecol,
                     const uint32_t* eleft, const uint32_t* tk,
                     const uint32_t* tedge, uint8_t* in_flag, PipeLaunch L);
void launch_tet_ac2_hot(const Stage& S, Cnt nk, const uint4* tets,
                        uint8_t* flag, Queue q1, uint32_t* rest,
                        unsigned* nrest, uint32_t* next, unsigned* nnext,
                        int budget, PipeLaunch L);
void launch_tet_ac2_esc(const Stage& S, QRef q,
                        const uint4* tets, uint8_t* flag, Queue q2,
                        PipeLaunch L);
void launch_tet_ac2_exact(const Stage& S, QRef q,
                          const uint4* tets, const uint32_t* perm,
                          uint8_t* flag, PipeLaunch L);
/* the entries of a host-built view as (item, aux) arrays, and the exact
 * tet test over entries sorted by tet (pairs of one tet share its D, N) */
void launch_view_split(const QView& v, uint32_t* item, uint32_t* aux,
                       PipeLaunch L);
void launch_tet_ac2_exact_sorted(const Stage& S, const uint32_t* item,
                                 const uint32_t* aux, uint32_t n,
                                 const uint4* tets, const uint32_t* perm,
                                 uint8_t* flag, PipeLaunch L);
/* kind 0 triangles, 1 edges */
void launch_low_ac2_hot(const Stage& S, int kind, Cnt nitems,
                        const uint32_t* ecol, const uint32_t* eleft,
                        const uint32_t* tk, const uint32_t* tedge,
                        const uint8_t* in_flag, uint8_t* flag, Queue q1,
                        uint32_t* rest, unsigned* nrest, int budget,
                        PipeLaunch L);
void launch_low_ac2_esc(const Stage& S, int kind, QRef q,
                        const uint32_t* ecol, const uint32_t* eleft,
                        const uint32_t* tk, const uint32_t* tedge,
                        uint8_t* flag, Queue q2, PipeLaunch L);
void launch_low_ac2_exact(const Stage& S, int kind, QRef q, const uint32_t* ecol,
                          const uint32_t* eleft, const uint32_t* tk,
                          const uint32_t* tedge, uint8_t* flag, PipeLaunch L);
/* the same over entries sorted by simplex (the pairs of a simplex share
 * its exact den and num) */
void launch_low_ac2_exact_sorted(const Stage& S, int kind, const uint32_t* item,
                                 const uint32_t* aux, uint32_t n,
                                 const uint32_t* ecol, const uint32_t* eleft,
                                 const uint32_t* tk, const uint32_t* tedge,
                                 uint8_t* flag, PipeLaunch L);


/* SoS table as compiled into the device code (for tests) */
int sos_table_host(int32_t* out3, int cap);

}  // namespace PAC_V
}  // namespace pac
#endif
