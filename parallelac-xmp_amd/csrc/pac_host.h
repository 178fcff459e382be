/*
 * pac_host.h -- host-side runtime shared by the C ABI (pac_engine.cpp) and
 * the pipeline driver (pac_pipeline.cpp, compiled once per ball type):
 * status errors, the per-device engine with its caching pool, device
 * buffers, results, event timers.
 */
#ifndef PAC_HOST_H
#define PAC_HOST_H

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/pac.h"
#include "pac_internal.h"

namespace pac {

struct PacError : std::runtime_error {
  int code;
  PacError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] inline void fail(int code, const std::string& m) {
  throw PacError(code, m);
}
#define HIPCHK(x)                                                        \
  do {                                                                   \
    hipError_t e_ = (x);                                                 \
    if (e_ != hipSuccess)                                                \
      ::pac::fail(e_ == hipErrorOutOfMemory ? PAC_ERR_NOMEM : PAC_ERR_DEVICE, \
                  std::string(#x) + ": " + hipGetErrorString(e_));       \
  } while (0)
#define KCHK() HIPCHK(hipGetLastError())

/* ------------------------------------------------------------------ */
/* caching device pool                                                 */
/* ------------------------------------------------------------------ */
struct Pool {
  std::mutex m;
  std::multimap<size_t, void*> free_;
  std::map<void*, size_t> cls_;
  /* bytes in live buffers, their high-water mark, bytes held (live +
   * cached) -- by size class */
  int64_t in_use = 0, peak = 0, held = 0;
  /* size classes: powers of two up to 1 MiB, then eight per octave over
   * the request plus 4 KiB (at most 12.5% over: the candidate arrays of a
   * large input run to gigabytes) */
  static size_t klass(size_t b) {
    size_t c = 4096;
    while (c < b && c < ((size_t)1 << 20)) c <<= 1;
    if (c >= b) return c;
    b += 4096;
    size_t oct = (size_t)1 << 20;
    while (oct * 2 <= b) oct <<= 1;
    const size_t step = oct >> 3;
    return (b + step - 1) / step * step;
  }
  /* Arena mode (set_arena): blocks are carved from one fixed device range
   * instead of hipMalloc, 256-byte aligned, top of stack first; a block
   * freed while it is the topmost is popped with any freed blocks below
   * it.  Exhausting the range fails with PAC_ERR_NOMEM. */
  char* abase = nullptr;
  size_t acap = 0, atop = 0, ahigh = 0;
  std::vector<std::pair<size_t, size_t>> ablk; /* (offset, 1 = freed) */
  void set_arena(void* base, size_t cap) {
    std::lock_guard<std::mutex> g(m);
    abase = (char*)base;
    acap = cap;
    atop = ahigh = 0;
    ablk.clear();
  }
  bool in_arena(const void* p) const {
    return abase && (const char*)p >= abase && (const char*)p < abase + acap;
  }
  void* get(size_t bytes) {
    if (abase) {
      std::lock_guard<std::mutex> g(m);
      const size_t b = ((bytes ? bytes : 1) + 255) & ~(size_t)255;
      if (b > acap - atop)
        fail(PAC_ERR_NOMEM, "device scratch of " + std::to_string(acap) +
                                " bytes exhausted (" + std::to_string(atop) +
                                " in use, " + std::to_string(b) +
                                " more requested): input too large for the "
                                "allocation-free entry point");
      ablk.push_back({atop, 0});
      void* p = abase + atop;
      atop += b;
      ahigh = std::max(ahigh, atop);
      return p;
    }
    size_t c = klass(bytes ? bytes : 1);
    {
      std::lock_guard<std::mutex> g(m);
      /* the smallest cached block that fits, if within a class or two */
      auto it = free_.lower_bound(c);
      if (it != free_.end() && it->first <= c + c / 4) {
        void* p = it->second;
        in_use += (int64_t)it->first;
        peak = std::max(peak, in_use);
        free_.erase(it);
        return p;
      }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, c);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      trim();
      e = hipMalloc(&p, c);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        fail(PAC_ERR_NOMEM, "device allocation of " + std::to_string(c) +
                                " bytes failed");
      }
    }
    std::lock_guard<std::mutex> g(m);
    cls_[p] = c;
    held += (int64_t)c;
    in_use += (int64_t)c;
    peak = std::max(peak, in_use);
    return p;
  }
  void put(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> g(m);
    if (in_arena(p)) {
      const size_t off = (size_t)((char*)p - abase);
      for (size_t k = ablk.size(); k-- > 0;)
        if (ablk[k].first == off && !ablk[k].second) {
          ablk[k].second = 1;
          break;
        }
      while (!ablk.empty() && ablk.back().second) {
        atop = ablk.back().first;
        ablk.pop_back();
      }
      return;
    }
    const size_t c = cls_[p];
    in_use -= (int64_t)c;
    free_.insert({c, p});
  }
  void trim() {
    std::lock_guard<std::mutex> g(m);
    for (auto& kv : free_) {
      (void)hipFree(kv.second);
      cls_.erase(kv.second);
      held -= (int64_t)kv.first;
    }
    free_.clear();
  }
};

struct Engine {
  int device = 0;
  hipStream_t stream = nullptr;
  Pool pool;
  std::mutex run_mu;
  void* pinned = nullptr; /* small pinned staging for scalar read-backs */
  /* the allocation-free entry point: the device's scratch array (its
   * address on this device) as an arena, and the pool a run takes its
   * buffers from while it holds run_mu (NULL: `pool`) */
  void* scratch = nullptr;
  Pool arena;
  Pool* run_pool = nullptr;
};
/* the pool the current run allocates from */
inline Pool* pool_of(Engine* E) { return E->run_pool ? E->run_pool : &E->pool; }
/* the per-device engine (created on first use) */
Engine* engine_for(int dev);

template <class T>
struct DBuf {
  Pool* pool = nullptr;
  T* p = nullptr;
  size_t n = 0;
  DBuf() {}
  DBuf(Pool* pl, size_t count) : pool(pl), n(count) {
    p = (T*)pl->get(std::max<size_t>(count, 1) * sizeof(T));
  }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : pool(o.pool), p(o.p), n(o.n) { o.p = nullptr; }
  DBuf& operator=(DBuf&& o) noexcept {
    release();
    pool = o.pool;
    p = o.p;
    n = o.n;
    o.p = nullptr;
    return *this;
  }
  ~DBuf() { release(); }
  void release() {
    if (p && pool) pool->put(p);
    p = nullptr;
  }
  T* get() const { return p; }
};

}  // namespace pac

/* ------------------------------------------------------------------ */
/* result                                                              */
/* ------------------------------------------------------------------ */
struct pac_result {
  pac::Engine* eng = nullptr;
  int64_t counts[4] = {0, 0, 0, 0};
  pac::DBuf<int32_t> lists[3];     /* edges, tris, tets (original ids) */
  pac::DBuf<double4> balls;        /* original order, r = r_raw + s (as computed) */
  double alpha = 0;
  pac_stats st;
  /* phase times (ms) measured outside the run's own timer (chunked runs) */
  double phase_add[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  /* filtration (pac_options.filtration): per output edge / triangle, 1 if
   * its smallest orthosphere is empty (it enters at its own rho); rank of
   * each ball in cell-sorted order; filtration keys, made on first use */
  bool filt = false;
  pac::DBuf<uint8_t> gab[2];
  pac::DBuf<uint32_t> rank;
  mutable pac::DBuf<uint64_t> bkey;
  mutable bool bkey_done = false;
  /* sizes seen by the run (sizing a recorded graph): candidate / survivor
   * counts of each stage, final counts, largest sub-queue fills */
  int64_t sizes[12] = {0};
};

/* a batch: one pac_result per pipeline pass, molecules in order */
struct pac_batch_result {
  pac::Engine* eng = nullptr;
  int64_t nmol = 0;
  std::vector<int64_t> off;                        /* ball offsets, nmol+1 */
  std::vector<std::unique_ptr<pac_result>> groups; /* one per pipeline pass */
  std::vector<int64_t> group_first;                /* first molecule of each group */
  std::vector<int32_t> mol_group;
  std::vector<int64_t> counts;                     /* nmol x 4 */
  std::vector<uint64_t> row0[3];                   /* first row of each molecule */
  pac_stats st;
};

/* thread's last error message (pac_last_error) */
void pac_set_error(const char* m);

namespace pac {
/* runs f, turning exceptions into status codes and the error message */
template <class F>
inline int guarded(F f) {
  try {
    f();
    return PAC_OK;
  } catch (const PacError& e) {
    pac_set_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    pac_set_error("host allocation failed");
    return PAC_ERR_NOMEM;
  } catch (const std::exception& e) {
    pac_set_error(e.what());
    return PAC_ERR_INVALID;
  }
}

/* batches (pac_engine.cpp) */
void add_stats(pac_stats& a, const pac_stats& b);
void compute_batch(const float* xyzr, const int64_t* offsets, int64_t nmol,
                   float solvent, float alpha, const pac_options* o,
                   pac_batch_result** out);

/* ------------------------------------------------------------------ */
/* helpers                                                             */
/* ------------------------------------------------------------------ */
inline int neg_lsb_exp_host(double v) {
  if (v == 0) return 0;
  int e;
  double m = std::frexp(v, &e);
  long long mi = (long long)std::ldexp(std::fabs(m), 53);
  int lsb = e - 53 + __builtin_ctzll((unsigned long long)mi);
  return lsb < 0 ? -lsb : 0;
}

struct Timer {
  hipEvent_t ev[10];
  int used = 0;
  hipStream_t s;
  bool on = true; /* off: marks are skipped (recorded pipelines) */
  bool made = true;
  /* events = false: no events are created, marks are skipped and every
   * phase reads 0 (the allocation-free entry point) */
  explicit Timer(hipStream_t st, bool events = true)
      : s(st), on(events), made(events) {
    if (on)
      for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  }
  ~Timer() {
    if (made)
      for (auto& e : ev) (void)hipEventDestroy(e);
  }
  void mark(int i) {
    if (on) HIPCHK(hipEventRecord(ev[i], s));
  }
  double ms(int a, int b) {
    if (!made) return 0.0;
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, ev[a], ev[b]));
    return t;
  }
};

/* optional per-kernel hipEvent timing (pac_options.profile) */
struct KProf {
  bool on = false;
  hipStream_t s = nullptr;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev;
  void begin(int k) {
    if (!on) return;
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    HIPCHK(hipEventRecord(a, s));
    ev.push_back({k, {a, b}});
  }
  void end() {
    if (!on) return;
    HIPCHK(hipEventRecord(ev.back().second.second, s));
  }
  void collect(double* out) {
    for (auto& e : ev) {
      float t = 0;
      if (hipEventElapsedTime(&t, e.second.first, e.second.second) == hipSuccess)
        out[e.first] += t;
      (void)hipEventDestroy(e.second.first);
      (void)hipEventDestroy(e.second.second);
    }
    ev.clear();
  }
  ~KProf() {
    for (auto& e : ev) {
      (void)hipEventDestroy(e.second.first);
      (void)hipEventDestroy(e.second.second);
    }
  }
};
enum { KP_EDGE_FILL, KP_TRI_FILL, KP_TET_FILL, KP_TET_AC2, KP_TRI_AC2,
       KP_EDGE_AC2, KP_SORT, KP_TET_AC2_EXACT };

template <class T>
inline T readback(Engine* E, hipStream_t s, const T* dptr) {
  HIPCHK(hipMemcpyAsync(E->pinned, dptr, sizeof(T), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  T v;
  std::memcpy(&v, E->pinned, sizeof(T));
  return v;
}

/* Statistics from the device counters (CTR_*, CTR_SLOTS copies each):
 * evaluations per precision stage of each family, SoS calls, emptiness
 * pairs.  Hot-stage evaluations of the filters are one per candidate (the
 * edge candidates are counted on the host, the others on the device);
 * those of the emptiness scans one per (simplex, ball) pair.  Adds to st. */
inline void stats_from_counters(const unsigned long long* hs, bool staged,
                                int64_t edge_candidates, pac_stats& st) {
  unsigned long long h[CTR_COUNT];
  for (int c = 0; c < CTR_COUNT; ++c) {
    h[c] = 0;
    for (int k = 0; k < CTR_SLOTS; ++k) h[c] += hs[c * CTR_SLOTS + k];
  }
  for (int f = 0; f < 6; ++f)
    for (int q = 0; q < 4; ++q)
      st.stage_evals[f][q] += (int64_t)h[CTR_STAGE + 4 * f + q];
  const int hot = staged ? 0 : 1;
  st.tri_candidates += (int64_t)h[CTR_CAND + 1];
  st.tet_candidates += (int64_t)h[CTR_CAND + 2];
  st.stage_evals[0][hot] += edge_candidates;
  st.stage_evals[1][hot] += (int64_t)h[CTR_CAND + 1];
  st.stage_evals[2][hot] += (int64_t)h[CTR_CAND + 2];
  for (int q = 0; q < 3; ++q) {
    st.stage_evals[3 + q][hot] += (int64_t)h[CTR_PAIRS + q];
    st.ac2_pairs[q] += (int64_t)h[CTR_PAIRS + q];
  }
  st.sos_calls += (int64_t)h[CTR_SOS];
}


/* One pipeline pass recorded with fixed capacities (for capture into a
 * graph): nothing is read back, every count stays on the device, and
 * overflows or a stale plan set status bits (REC_*). */
enum { RC_ECAND, RC_EALPHA, RC_TSLOTS, RC_TALPHA, RC_KSLOTS, RC_KALPHA,
       RC_KF, RC_TF, RC_EF, RC_COUNT };
struct PassRec {
  Pool* pool;              /* the pass's buffers */
  const void* plan;        /* PassPlan of the pipeline's ball type */
  const Bounds* bref;      /* device: the bounds the plan was made from */
  Bounds* bnow;            /* device: bounds of the current input */
  int64_t cap[RC_COUNT];   /* capacity of each count */
  uint64_t* dcnt;          /* device [RC_COUNT]: the counts */
  unsigned* status;        /* device: REC_* bits */
  unsigned* fill;          /* device [2]: largest sub-queue fills */
  unsigned long long* ctr; /* device counters (shared by the passes) */
  unsigned* qn;            /* device [QN_WORDS] queue counters */
  QView* views;            /* device [2] */
  QEntry *q1, *q2;         /* hot / escalation queues */
  uint32_t c1, c2;         /* their sub-queue capacities */
};

/* A chunk of the chunked driver run as a sub-input: planned on the grid
 * of the whole input, its output kept only for the simplices whose lowest
 * vertex has a cell-sorted id (of the sub-input) in [own_lo, own_hi). */
struct SubInput {
  Bounds grid;
  uint32_t own_lo, own_hi;
};

namespace PAC_V {
/* grids, exact scale and device molecule tables of one pipeline pass */
struct PassPlan {
  std::vector<Bounds> bounds; /* per molecule, as reduced on the device */
  std::vector<GridP> grids;
  int64_t G = 0;
  int kf = 0;                  /* fine scan grid subdivision (0: none) */
  std::vector<GridP> sgrids;   /* fine scan grids */
  int64_t Gf = 0;
  DBuf<GridP> dsgrids;
  int S = 0;
  WI alpha_wi;
  bool no_edges = true;
  int k_enum = 0, k_ac2 = 0;
  DBuf<uint32_t> dmoff, dmol;
  DBuf<GridP> dgrids;
};
void plan_pass(Engine* E, const Ball* raw, int64_t n, Real solvent,
               Real alpha, const pac_options& opt, hipStream_t s,
               const int64_t* moff, int64_t nmol, PassPlan& pl,
               const Bounds* gridb = nullptr);
}  // namespace PAC_V
namespace vf32 {
void run_pipeline(Engine* E, const float4* raw, int64_t n, float solvent,
                  float alpha, const pac_options& opt, pac_result* R,
                  hipStream_t s, Timer& T, KProf& KP,
                  const int64_t* moff = nullptr, int64_t nmol = 1,
                  const PassRec* rec = nullptr,
                  const SubInput* sub = nullptr);
/* the chunked (LD) driver over opt.chunk_balls */
void run_chunked(Engine* E, const float4* raw, int64_t n, float solvent,
                 float alpha, const pac_options& opt, pac_result* R,
                 hipStream_t s, Timer& T, KProf& KP);
}
namespace vf64 {
void run_pipeline(Engine* E, const double4* raw, int64_t n, double solvent,
                  double alpha, const pac_options& opt, pac_result* R,
                  hipStream_t s, Timer& T, KProf& KP,
                  const int64_t* moff = nullptr, int64_t nmol = 1,
                  const PassRec* rec = nullptr,
                  const SubInput* sub = nullptr);
/* the chunked (LD) driver over opt.chunk_balls */
void run_chunked(Engine* E, const double4* raw, int64_t n, double solvent,
                 double alpha, const pac_options& opt, pac_result* R,
                 hipStream_t s, Timer& T, KProf& KP);
}

}  // namespace pac
#endif
