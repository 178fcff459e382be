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
  static size_t klass(size_t b) {
    size_t c = 4096;
    while (c < b) c <<= 1;
    return c;
  }
  void* get(size_t bytes) {
    size_t c = klass(bytes ? bytes : 1);
    {
      std::lock_guard<std::mutex> g(m);
      auto it = free_.find(c);
      if (it != free_.end()) {
        void* p = it->second;
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
    return p;
  }
  void put(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> g(m);
    free_.insert({cls_[p], p});
  }
  void trim() {
    std::lock_guard<std::mutex> g(m);
    for (auto& kv : free_) {
      (void)hipFree(kv.second);
      cls_.erase(kv.second);
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
};
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
  /* sizes seen by the run (sizing a recorded graph): candidate / survivor
   * counts of each stage, final counts, largest sub-queue fills */
  int64_t sizes[12] = {0};
};

namespace pac {
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
  explicit Timer(hipStream_t st) : s(st) {
    for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  }
  ~Timer() {
    for (auto& e : ev) (void)hipEventDestroy(e);
  }
  void mark(int i) { HIPCHK(hipEventRecord(ev[i], s)); }
  double ms(int a, int b) {
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

/* small fills on a stream (kernels, not hipMemsetAsync) */
void fill_u8(uint8_t* p, int64_t n, uint8_t v, hipStream_t s);
void zero_u32(unsigned* p, int n, hipStream_t s);
void zero_u64(unsigned long long* p, int64_t n, hipStream_t s);

namespace PAC_V {
/* grids, exact scale and device molecule tables of one pipeline pass */
struct PassPlan {
  std::vector<GridP> grids;
  int64_t G = 0;
  int S = 0;
  WI alpha_wi;
  bool no_edges = true;
  int k_enum = 0, k_ac2 = 0;
  DBuf<uint32_t> dmoff, dmol;
  DBuf<GridP> dgrids;
};
void plan_pass(Engine* E, const Ball* raw, int64_t n, Real solvent,
               Real alpha, const pac_options& opt, hipStream_t s,
               const int64_t* moff, int64_t nmol, PassPlan& pl);
}  // namespace PAC_V
namespace vf32 {
void run_pipeline(Engine* E, const float4* raw, int64_t n, float solvent,
                  float alpha, const pac_options& opt, pac_result* R,
                  hipStream_t s, Timer& T, KProf& KP,
                  const int64_t* moff = nullptr, int64_t nmol = 1);
}
namespace vf64 {
void run_pipeline(Engine* E, const double4* raw, int64_t n, double solvent,
                  double alpha, const pac_options& opt, pac_result* R,
                  hipStream_t s, Timer& T, KProf& KP,
                  const int64_t* moff = nullptr, int64_t nmol = 1);
}

}  // namespace pac
#endif
