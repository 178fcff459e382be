/*
 * pac_engine.cpp -- host orchestration of the MI355X pipeline and the C ABI
 * declared in include/pac.h.
 *
 * Replaces the reference's main() phase sequence (AlphaComplexParallel.cu:
 * 1126-1806): one HIP stream per call, device-resident intermediates taken
 * from a caching pool (no hipMalloc in steady state), counts read back
 * between stages only where a buffer must be sized, hipEvents around the
 * reference's eight phases.  Errors are returned as status codes, never
 * printed (the reference prints and continues: :55-61).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pac.h"
#include "pac_internal.h"

using namespace pac;

/* ------------------------------------------------------------------ */
/* errors                                                              */
/* ------------------------------------------------------------------ */
static thread_local std::string g_err;
struct PacError : std::runtime_error {
  int code;
  PacError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
static void fail(int code, const std::string& m) { throw PacError(code, m); }
void pac_set_error(const char* m) { g_err = m ? m : ""; }
#define HIPCHK(x)                                                        \
  do {                                                                   \
    hipError_t e_ = (x);                                                 \
    if (e_ != hipSuccess)                                                \
      fail(e_ == hipErrorOutOfMemory ? PAC_ERR_NOMEM : PAC_ERR_DEVICE,   \
           std::string(#x) + ": " + hipGetErrorString(e_));              \
  } while (0)
#define KCHK() HIPCHK(hipGetLastError())

template <class F>
static int guarded(F f) {
  try {
    f();
    return PAC_OK;
  } catch (const PacError& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return PAC_ERR_NOMEM;
  } catch (const std::exception& e) {
    g_err = e.what();
    return PAC_ERR_INVALID;
  }
}

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
static std::mutex g_eng_mu;
static std::map<int, Engine*> g_engines;
static Engine* engine_for(int dev) {
  std::lock_guard<std::mutex> g(g_eng_mu);
  auto it = g_engines.find(dev);
  if (it != g_engines.end()) return it->second;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    fail(PAC_ERR_DEVICE, "no HIP device available");
  }
  if (dev < 0 || dev >= ndev) fail(PAC_ERR_INVALID, "bad device ordinal");
  HIPCHK(hipSetDevice(dev));
  Engine* e = new Engine();
  e->device = dev;
  HIPCHK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
  HIPCHK(hipHostMalloc(&e->pinned, 4096, hipHostMallocDefault));
  g_engines[dev] = e;
  return e;
}

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

/* ------------------------------------------------------------------ */
/* result                                                              */
/* ------------------------------------------------------------------ */
struct pac_result {
  Engine* eng = nullptr;
  int64_t counts[4] = {0, 0, 0, 0};
  DBuf<int32_t> lists[3];     /* edges, tris, tets (original ids) */
  DBuf<double4> balls;        /* original order, r = r_raw + s (as computed) */
  double alpha = 0;
  pac_stats st;
};

/* ------------------------------------------------------------------ */
/* helpers                                                             */
/* ------------------------------------------------------------------ */
static int neg_lsb_exp_host(double v) {
  if (v == 0) return 0;
  int e;
  double m = std::frexp(v, &e);
  long long mi = (long long)std::ldexp(std::fabs(m), 53);
  int lsb = e - 53 + __builtin_ctzll((unsigned long long)mi);
  return lsb < 0 ? -lsb : 0;
}

/* v * 2^shift as a 512-bit two's-complement integer (v*2^shift integral) */
static WI wi_from_scaled_host(double v, int shift) {
  WI r;
  std::memset(&r, 0, sizeof(r));
  if (v == 0) return r;
  int e;
  double m = std::frexp(v, &e);
  long long mi = (long long)std::ldexp(m, 53);
  bool neg = mi < 0;
  unsigned long long mag = (unsigned long long)(neg ? -mi : mi);
  int ex = e - 53 + shift;
  while (ex < 0 && (mag & 1ull) == 0) {
    mag >>= 1;
    ++ex;
  }
  if (ex < 0) fail(PAC_ERR_RANGE, "alpha not representable at the scale");
  /* place mag << ex into limbs */
  for (int b = 0; b < 64; ++b)
    if (mag >> b & 1ull) {
      int pos = b + ex;
      if (pos >= 32 * WL - 1) fail(PAC_ERR_RANGE, "alpha too large");
      r.l[pos / 32] |= 1u << (pos % 32);
    }
  if (neg) {
    /* two's complement negate */
    uint64_t carry = 1;
    for (int i = 0; i < WL; ++i) {
      uint64_t t = (uint64_t)(uint32_t)~r.l[i] + carry;
      r.l[i] = (uint32_t)t;
      carry = t >> 32;
    }
  }
  return r;
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
static T readback(Engine* E, hipStream_t s, const T* dptr) {
  HIPCHK(hipMemcpyAsync(E->pinned, dptr, sizeof(T), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  T v;
  std::memcpy(&v, E->pinned, sizeof(T));
  return v;
}

__global__ void k_fill_u8(uint8_t* p, int64_t n, uint8_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_zero_u32(unsigned* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}
__global__ void k_zero_u64(unsigned long long* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}
static void fill_u8(uint8_t* p, int64_t n, uint8_t v, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_u8, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     s, p, n, v);
}
static void zero_u32(unsigned* p, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_u32, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
}

/* ------------------------------------------------------------------ */
/* the pipeline                                                        */
/* ------------------------------------------------------------------ */
/* Grid of one molecule from its bounds (AlphaComplexParallel.cu:1120-1124,
 * evaluated in FP32 exactly as there), with the stencil half-widths that
 * make the enumeration and the emptiness boxes complete. */
struct GridPlan {
  GridP g;
  int64_t G = 0;
};
static GridPlan plan_grid(const Bounds& B, float alpha, int policy) {
  GridPlan P;
  GridP& g = P.g;
  std::memset(&g, 0, sizeof(g));
  volatile float m2 = B.max_raw_r * B.max_raw_r;
  volatile float m2a = m2 + alpha;
  volatile float pmr = sqrtf(m2a);
  float step = 2 * pmr;
  if (!(step > 0) || !std::isfinite(step))
    fail(PAC_ERR_INVALID, "grid step 2*sqrt(maxR^2+alpha) is not positive");
  volatile float qx = (B.maxx - B.minx) / step, qy = (B.maxy - B.miny) / step,
                 qz = (B.maxz - B.minz) / step;
  int64_t ex = (int64_t)(int)qx + 1, ey = (int64_t)(int)qy + 1,
          ez = (int64_t)(int)qz + 1;
  P.G = ex * ey * ez;
  if (P.G >= ((int64_t)1 << 31)) fail(PAC_ERR_RANGE, "grid has too many cells");
  g.minx = B.minx;
  g.miny = B.miny;
  g.minz = B.minz;
  g.step = step;
  g.ex = (int)ex;
  g.ey = (int)ey;
  g.ez = (int)ez;
  double maxw = (double)B.max_r * (double)B.max_r;
  double reach2 = maxw + (double)alpha;
  double reach = reach2 > 0 ? std::sqrt(reach2) * (1 + 1e-12) + 1e-12 : 0.0;
  double maxext = std::max({(double)B.maxx - B.minx, (double)B.maxy - B.miny,
                            (double)B.maxz - B.minz});
  double margin = (maxext + (double)step) * 3.0 * 0x1p-24 + 1e-9;
  if (policy == PAC_CELL_REFERENCE_RAW) {
    g.k_enum = 1;
    g.k_ac2 = 1;
    reach = (double)step * (1 - 1e-6);
    margin = 0;
  } else {
    g.k_enum = std::max(1, (int)std::ceil((2 * reach + 2 * margin) / step));
    g.k_ac2 = std::max(1, (int)std::ceil((reach + margin) / step));
  }
  g.reach = reach;
  g.margin = margin;
  return P;
}

/* One pipeline pass over n balls.  nmol == 1: one molecule; otherwise the
 * balls are nmol molecules, moff[0..nmol] (host, relative to raw), each with
 * its own grid, processed together by every kernel launch. */
static void run_pipeline(Engine* E, const float4* raw, int64_t n, float solvent,
                         float alpha, const pac_options& opt, pac_result* R,
                         hipStream_t s, Timer& T, KProf& KP,
                         const int64_t* moff = nullptr, int64_t nmol = 1) {
  Pool* P = &E->pool;
  PipeLaunch L{s, kBlock};
  pac_stats& st = R->st;
  st.n = n;
  if (n == 0) return;
  if (n > (int64_t)INT32_MAX) fail(PAC_ERR_OVERFLOW, "too many balls");
  const bool batch = nmol > 1;

  /* ---- bounds, grids (AlphaComplexParallel.cu:1089-1124) ---- */
  std::vector<Bounds> hb(std::max<int64_t>(nmol, 1));
  DBuf<uint32_t> dmoff, dmol;
  DBuf<GridP> dgrids;
  if (!batch) {
    DBuf<Bounds> bpart(P, 1024 * 2 + 4), bout(P, 1);
    launch_bounds(raw, n, solvent, bpart.get(), bout.get(), L);
    KCHK();
    hb[0] = readback(E, s, bout.get());
  } else {
    std::vector<uint32_t> mo(nmol + 1);
    for (int64_t m = 0; m <= nmol; ++m) mo[m] = (uint32_t)moff[m];
    dmoff = DBuf<uint32_t>(P, nmol + 1);
    HIPCHK(hipMemcpyAsync(dmoff.get(), mo.data(), (nmol + 1) * sizeof(uint32_t),
                          hipMemcpyHostToDevice, s));
    DBuf<Bounds> bout(P, nmol);
    launch_bounds_mol(raw, dmoff.get(), nmol, solvent, bout.get(), L);
    KCHK();
    HIPCHK(hipMemcpyAsync(hb.data(), bout.get(), nmol * sizeof(Bounds),
                          hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  std::vector<GridP> grids(hb.size());
  int64_t G = 0;
  int shift = 0;
  double mag = 1.0;
  bool no_edges = true;
  st.k_enum = st.k_ac2 = 0;
  for (size_t m = 0; m < hb.size(); ++m) {
    const Bounds& B = hb[m];
    if (B.nonfinite) fail(PAC_ERR_INVALID, "non-finite coordinate or radius");
    bool empty = batch && moff[m + 1] == moff[m];
    GridPlan gp;
    if (empty) { /* one empty cell, never visited */
      std::memset(&gp.g, 0, sizeof(gp.g));
      gp.g.step = 1;
      gp.g.ex = gp.g.ey = gp.g.ez = gp.g.k_enum = gp.g.k_ac2 = 1;
      gp.G = 1;
    } else {
      gp = plan_grid(B, alpha, opt.cell_policy);
      shift = std::max(shift, B.shift);
      mag = std::max({mag, (double)B.max_abs, (double)B.max_r});
      if ((double)B.max_r * (double)B.max_r + (double)alpha >= 0) no_edges = false;
    }
    gp.g.cbase = (uint32_t)G;
    G += gp.G;
    if (G >= ((int64_t)1 << 31)) fail(PAC_ERR_RANGE, "grids have too many cells");
    grids[m] = gp.g;
    st.k_enum = std::max(st.k_enum, gp.g.k_enum);
    st.k_ac2 = std::max(st.k_ac2, gp.g.k_ac2);
  }
  st.grid_cells = (int32_t)G;

  /* exact embedding scale (one for all molecules: any common scale at
   * which every input is an integer gives the same exact signs) */
  int S = std::max(shift, (neg_lsb_exp_host(alpha) + 1) / 2);
  mag = std::max(mag, std::sqrt(std::fabs((double)alpha)));
  double bits = std::log2(mag) + S + 2;
  if (8 * bits + 16 > 32.0 * WL - 2)
    fail(PAC_ERR_RANGE, "input exponent range exceeds the 512-bit exact stage");
  st.scale_bits = S;
  Stage SG;
  std::memset(&SG, 0, sizeof(SG));
  SG.n = n;
  SG.g = grids[0];
  if (batch) {
    dgrids = DBuf<GridP>(P, nmol);
    HIPCHK(hipMemcpyAsync(dgrids.get(), grids.data(), nmol * sizeof(GridP),
                          hipMemcpyHostToDevice, s));
    dmol = DBuf<uint32_t>(P, n);
    launch_mol_ids(dmoff.get(), nmol, n, dmol.get(), L);
    KCHK();
    SG.grids = dgrids.get();
    SG.mol = dmol.get();
  }
  SG.alpha = alpha;
  SG.ex.S = S;
  SG.ex.alpha = wi_from_scaled_host((double)alpha, 2 * S);
  SG.staged = opt.precision == PAC_PREC_STAGED;
  DBuf<unsigned long long> ctr(P, CTR_COUNT * CTR_SLOTS);
  hipLaunchKernelGGL(k_zero_u64, dim3((CTR_COUNT * CTR_SLOTS + 255) / 256),
                     dim3(256), 0, s, ctr.get(), CTR_COUNT * CTR_SLOTS);
  SG.ctr = ctr.get();
  DBuf<unsigned> qn(P, 4); /* [0] hot queue, [1] escalation queue */

  /* ---- binning: key, stable radix sort, gather, cell starts ---- */
  DBuf<uint32_t> key(P, n), idx(P, n), key2(P, n), idx2(P, n);
  launch_cell_key(raw, n, SG.g, SG.grids, SG.mol, key.get(), idx.get(), L);
  KCHK();
  int kbits = 0;
  while (kbits < 32 && ((int64_t)1 << kbits) < G) ++kbits;
  size_t rtb = radix_tmp_bytes(n);
  DBuf<uint8_t> rtmp(P, rtb);
  bool alt = false;
  KP.begin(KP_SORT);
  radix_sort_pairs(key.get(), idx.get(), key2.get(), idx2.get(), n, kbits,
                   rtmp.get(), rtb, s, &alt);
  KP.end();
  KCHK();
  const uint32_t* skey = alt ? key2.get() : key.get();
  const uint32_t* sidx = alt ? idx2.get() : idx.get();
  DBuf<float4> ball(P, n);
  DBuf<uint32_t> perm(P, n), cellk(P, n), cstart(P, G + 1);
  launch_gather(raw, sidx, skey, n, solvent, ball.get(), perm.get(),
                cellk.get(), L);
  launch_cell_starts(cellk.get(), n, G, cstart.get(), L);
  KCHK();
  rtmp.release();
  key.release();
  key2.release();
  idx.release();
  idx2.release();
  SG.ball = ball.get();
  SG.cellk = cellk.get();
  SG.cstart = cstart.get();
  T.mark(2);

  auto scan_tmp = [&](int64_t m) { return DBuf<uint8_t>(P, scan_tmp_bytes(m) + 64); };
  const bool staged = SG.staged != 0;

  /* precision ladder: hot kernel (q1) -> [staged: escalation kernel FP64 ->
   * double-double (q1 -> q2)] -> exact wide-integer kernel.  Queue
   * overflows rerun the producing kernel with a larger queue; only the
   * emptiness kernels can overflow, and they are idempotent. */
  auto ladder = [&](uint32_t cap1, int kp_hot, int kp_exact, auto&& hot,
                    auto&& esc, auto&& exact) {
    for (;;) {
      DBuf<QEntry> q1(P, cap1);
      zero_u32(qn.get(), 2, s);
      if (kp_hot >= 0) KP.begin(kp_hot);
      hot(Queue{q1.get(), qn.get(), cap1});
      if (kp_hot >= 0) KP.end();
      KCHK();
      uint32_t n1 = readback(E, s, qn.get());
      if (n1 > cap1) {
        cap1 = (uint32_t)std::min<int64_t>((int64_t)n1 + n1 / 4 + 1024,
                                           (int64_t)UINT32_MAX - 1);
        continue;
      }
      const QEntry* qx = q1.get();
      uint32_t nx = n1;
      DBuf<QEntry> q2;
      if (staged && n1) {
        uint32_t cap2 = (uint32_t)std::min<int64_t>(4 * (int64_t)n1 + 4096,
                                                    (int64_t)UINT32_MAX - 1);
        for (;;) {
          q2 = DBuf<QEntry>(P, cap2);
          zero_u32(qn.get() + 1, 1, s);
          esc(q1.get(), n1, Queue{q2.get(), qn.get() + 1, cap2});
          KCHK();
          uint32_t n2 = readback(E, s, qn.get() + 1);
          if (n2 <= cap2) {
            nx = n2;
            break;
          }
          cap2 = (uint32_t)std::min<int64_t>((int64_t)n2 + n2 / 4 + 1024,
                                             (int64_t)UINT32_MAX - 1);
        }
        qx = q2.get();
      }
      if (kp_exact >= 0) KP.begin(kp_exact);
      exact(qx, nx);
      if (kp_exact >= 0) KP.end();
      KCHK();
      return;
    }
  };

  /* alpha filter of one dimension over its enumerated candidates: the
   * ladder settles every flag, a flat scan of the flags gives both the
   * stable compaction positions and the survivors' CSR row starts. */
  int64_t host_evals[6] = {0, 0, 0, 0, 0, 0}; /* hot-stage evaluations booked on the host */
  auto filter_stage = [&](int kind, CandCtx C, int64_t nc, const uint32_t* cval,
                          const uint32_t* cseg, uint8_t* flag, int kp,
                          DBuf<uint64_t>& pos) -> int64_t {
    if (kind == 1) host_evals[0] += nc; /* tris / tets: slots != candidates */
    ladder(
        (uint32_t)std::max<int64_t>(nc, 1), kp, -1,
        [&](Queue q) { launch_filter(SG, kind, C, nc, cval, cseg, flag, q, L); },
        [&](const QEntry* q, uint32_t nq, Queue q2) {
          launch_filter_esc(SG, kind, C, q, nq, cval, flag, q2, L);
        },
        [&](const QEntry* q, uint32_t nq) {
          launch_filter_exact(SG, kind, C, q, nq, cval, flag, L);
        });
    pos = DBuf<uint64_t>(P, nc + 1);
    {
      auto tmp = scan_tmp(nc);
      scan_u8flag_to_u64(flag, pos.get(), nc, tmp.get(), 0, s);
    }
    KCHK();
    return (int64_t)readback(E, s, pos.get() + nc);
  };
  auto count_scan = [&](const uint32_t* cnt, int64_t nseg, uint64_t* off,
                        const char* what) -> int64_t {
    {
      auto tmp = scan_tmp(nseg);
      scan_u32_to_u64(cnt, off, nseg, tmp.get(), 0, s);
    }
    KCHK();
    int64_t tot = (int64_t)readback(E, s, off + nseg);
    if (tot >= ((int64_t)1 << 32))
      fail(PAC_ERR_OVERFLOW, std::string(what) + " candidates exceed 2^32");
    return tot;
  };

  /* ---- edges (AlphaComplexParallel.cu:1200-1263) ---- */
  int64_t Ecand = 0, Ealpha = 0;
  DBuf<uint64_t> erow(P, n + 1);
  DBuf<uint32_t> ecol, eleft;
  CandCtx C0{nullptr, nullptr, nullptr, nullptr};
  if (!no_edges) {
    DBuf<uint64_t> off(P, n + 1);
    {
      DBuf<uint32_t> cnt(P, n);
      launch_edge_count(SG, cnt.get(), L);
      Ecand = count_scan(cnt.get(), n, off.get(), "edge");
    }
    DBuf<uint32_t> cval(P, Ecand), cseg(P, Ecand);
    DBuf<uint8_t> flag(P, Ecand);
    launch_edge_enum(SG, off.get(), cval.get(), cseg.get(), L);
    DBuf<uint64_t> pos;
    Ealpha = filter_stage(1, C0, Ecand, cval.get(), cseg.get(), flag.get(),
                          KP_EDGE_FILL, pos);
    launch_seg_starts(n, off.get(), pos.get(), erow.get(), L);
    ecol = DBuf<uint32_t>(P, Ealpha);
    eleft = DBuf<uint32_t>(P, Ealpha);
    launch_compact_flat(1, Ecand, C0, cval.get(), cseg.get(), flag.get(),
                        pos.get(), ecol.get(), eleft.get(), nullptr, nullptr,
                        nullptr, nullptr, nullptr, L);
    KCHK();
  } else {
    hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)((n + 1 + 255) / 256)),
                       dim3(256), 0, s, (unsigned long long*)erow.get(),
                       (int)(n + 1));
    ecol = DBuf<uint32_t>(P, 1);
    eleft = DBuf<uint32_t>(P, 1);
  }
  st.edge_candidates = Ecand;
  st.edge_alpha = Ealpha;
  T.mark(3);

  /* ---- triangles (AlphaComplexParallel.cu:1279-1371) ---- */
  int64_t Tcand = 0, Tslots = 0, Talpha = 0;
  DBuf<uint64_t> trow(P, Ealpha + 1);
  DBuf<uint32_t> tk, tedge, teik, tejk;
  CandCtx C1{eleft.get(), ecol.get(), nullptr, nullptr};
  {
    DBuf<uint64_t> off(P, Ealpha + 1);
    {
      /* exact counts: a triangle's third vertex must be in both rows, and
       * the merge that counts it is cheap next to filtering empty slots */
      DBuf<uint32_t> cnt(P, Ealpha);
      launch_tri_count(Ealpha, erow.get(), ecol.get(), eleft.get(), cnt.get(), L);
      Tslots = count_scan(cnt.get(), Ealpha, off.get(), "triangle");
    }
    DBuf<uint32_t> cval(P, Tslots), cseg(P, Tslots), cea(P, Tslots),
        ceb(P, Tslots);
    DBuf<uint8_t> flag(P, Tslots);
    launch_tri_enum(Ealpha, erow.get(), ecol.get(), eleft.get(), off.get(),
                    cval.get(), cseg.get(), cea.get(), ceb.get(), ctr.get(), L);
    DBuf<uint64_t> pos;
    Talpha = filter_stage(2, C1, Tslots, cval.get(), cseg.get(), flag.get(),
                          KP_TRI_FILL, pos);
    launch_seg_starts(Ealpha, off.get(), pos.get(), trow.get(), L);
    tk = DBuf<uint32_t>(P, Talpha);
    tedge = DBuf<uint32_t>(P, Talpha);
    teik = DBuf<uint32_t>(P, Talpha);
    tejk = DBuf<uint32_t>(P, Talpha);
    launch_compact_flat(2, Tslots, C1, cval.get(), cseg.get(), flag.get(),
                        pos.get(), tk.get(), tedge.get(), nullptr, cea.get(),
                        ceb.get(), teik.get(), tejk.get(), L);
    KCHK();
  }
  st.tri_alpha = Talpha;
  T.mark(4);

  /* ---- tets (AlphaComplexParallel.cu:1381-1442) ---- */
  int64_t Kcand = 0, Kslots = 0, Kalpha = 0;
  DBuf<uint4> tets;
  CandCtx C2{eleft.get(), ecol.get(), tk.get(), tedge.get()};
  {
    DBuf<uint64_t> off(P, Talpha + 1);
    {
      DBuf<uint32_t> cap(P, Talpha);
      launch_tail_caps(Talpha, trow.get(), tedge.get(), cap.get(), L);
      Kslots = count_scan(cap.get(), Talpha, off.get(), "tet");
    }
    DBuf<uint32_t> cval(P, Kslots), cseg(P, Kslots);
    DBuf<uint8_t> flag(P, Kslots);
    launch_tet_enum(Talpha, trow.get(), tk.get(), tedge.get(), teik.get(),
                    tejk.get(), off.get(), cval.get(), cseg.get(), ctr.get(), L);
    DBuf<uint64_t> pos;
    Kalpha = filter_stage(3, C2, Kslots, cval.get(), cseg.get(), flag.get(),
                          KP_TET_FILL, pos);
    tets = DBuf<uint4>(P, Kalpha);
    launch_compact_flat(3, Kslots, C2, cval.get(), cseg.get(), flag.get(),
                        pos.get(), nullptr, nullptr, tets.get(), nullptr,
                        nullptr, nullptr, nullptr, L);
    KCHK();
  }
  (void)Tcand;
  (void)Kcand;
  st.tet_alpha = Kalpha;
  T.mark(5);

  auto ac2_cap = [](int64_t items) {
    return (uint32_t)std::min<int64_t>(std::max<int64_t>(1 << 20, items / 8),
                                       (int64_t)UINT32_MAX - 1);
  };

  /* ---- tet emptiness ---- */
  int budget = opt.ac2_budget > 0 ? opt.ac2_budget : (opt.ac2_budget < 0 ? 0 : 8);
  DBuf<uint32_t> rest(P, std::max({Kalpha, Talpha, Ealpha, (int64_t)1}));
  DBuf<uint8_t> kflag(P, Kalpha);
  if (opt.skip_tet_ac2) {
    fill_u8(kflag.get(), Kalpha, 1, s);
  } else if (Kalpha) {
    ladder(
        ac2_cap(Kalpha), KP_TET_AC2, KP_TET_AC2_EXACT,
        [&](Queue q) {
          zero_u32(qn.get() + 2, 1, s);
          launch_tet_ac2_hot(SG, Kalpha, tets.get(), kflag.get(), q, rest.get(),
                             qn.get() + 2, budget, L);
        },
        [&](const QEntry* q, uint32_t nq, Queue q2) {
          launch_tet_ac2_esc(SG, q, nq, tets.get(), kflag.get(), q2, L);
        },
        [&](const QEntry* q, uint32_t nq) {
          launch_tet_ac2_exact(SG, q, nq, tets.get(), perm.get(), kflag.get(), L);
        });
  }
  int64_t Kf = 0;
  DBuf<uint4> ftets;
  {
    DBuf<uint64_t> pos(P, Kalpha + 1);
    auto tmp = scan_tmp(Kalpha);
    scan_u8flag_to_u64(kflag.get(), pos.get(), Kalpha, tmp.get(), 0, s);
    KCHK();
    Kf = (int64_t)readback(E, s, pos.get() + Kalpha);
    ftets = DBuf<uint4>(P, Kf);
    launch_compact_tets(Kalpha, tets.get(), kflag.get(), pos.get(), ftets.get(), L);
    KCHK();
  }
  tets.release();
  T.mark(6);

  /* ---- triangle closure + dangling emptiness ---- */
  DBuf<uint8_t> tri_in(P, Talpha), tflag(P, Talpha);
  fill_u8(tri_in.get(), Talpha, 0, s);
  launch_mark_faces(Kf, ftets.get(), erow.get(), ecol.get(), trow.get(),
                    tk.get(), tri_in.get(), L);
  if (Talpha)
    ladder(
        ac2_cap(Talpha), KP_TRI_AC2, -1,
        [&](Queue q) {
          zero_u32(qn.get() + 2, 1, s);
          launch_low_ac2_hot(SG, 0, Talpha, ecol.get(), eleft.get(), tk.get(),
                             tedge.get(), tri_in.get(), tflag.get(), q,
                             rest.get(), qn.get() + 2, budget, L);
        },
        [&](const QEntry* q, uint32_t nq, Queue q2) {
          launch_low_ac2_esc(SG, 0, q, nq, ecol.get(), eleft.get(), tk.get(),
                             tedge.get(), tflag.get(), q2, L);
        },
        [&](const QEntry* q, uint32_t nq) {
          launch_low_ac2_exact(SG, 0, q, nq, ecol.get(), eleft.get(), tk.get(),
                               tedge.get(), tflag.get(), L);
        });
  int64_t Tf = 0;
  DBuf<uint32_t> ftri;
  {
    DBuf<uint64_t> pos(P, Talpha + 1);
    auto tmp = scan_tmp(Talpha);
    scan_u8flag_to_u64(tflag.get(), pos.get(), Talpha, tmp.get(), 0, s);
    KCHK();
    Tf = (int64_t)readback(E, s, pos.get() + Talpha);
    ftri = DBuf<uint32_t>(P, Tf);
    launch_compact_ids(Talpha, tflag.get(), pos.get(), ftri.get(), L);
    KCHK();
  }
  T.mark(7);

  /* ---- edge closure + dangling emptiness ---- */
  DBuf<uint8_t> edge_in(P, Ealpha), eflag(P, Ealpha);
  fill_u8(edge_in.get(), Ealpha, 0, s);
  launch_mark_edges(Tf, ftri.get(), tedge.get(), teik.get(), tejk.get(),
                    edge_in.get(), L);
  if (Ealpha)
    ladder(
        ac2_cap(Ealpha), KP_EDGE_AC2, -1,
        [&](Queue q) {
          zero_u32(qn.get() + 2, 1, s);
          launch_low_ac2_hot(SG, 1, Ealpha, ecol.get(), eleft.get(), nullptr,
                             nullptr, edge_in.get(), eflag.get(), q,
                             rest.get(), qn.get() + 2, budget, L);
        },
        [&](const QEntry* q, uint32_t nq, Queue q2) {
          launch_low_ac2_esc(SG, 1, q, nq, ecol.get(), eleft.get(), nullptr,
                             nullptr, eflag.get(), q2, L);
        },
        [&](const QEntry* q, uint32_t nq) {
          launch_low_ac2_exact(SG, 1, q, nq, ecol.get(), eleft.get(), nullptr,
                               nullptr, eflag.get(), L);
        });
  int64_t Ef = 0;
  DBuf<uint64_t> pos_e(P, Ealpha + 1);
  {
    auto tmp = scan_tmp(Ealpha);
    scan_u8flag_to_u64(eflag.get(), pos_e.get(), Ealpha, tmp.get(), 0, s);
    KCHK();
    Ef = (int64_t)readback(E, s, pos_e.get() + Ealpha);
  }

  /* ---- emit final lists (original ids, reference order) ---- */
  R->lists[0] = DBuf<int32_t>(P, 2 * Ef);
  R->lists[1] = DBuf<int32_t>(P, 3 * Tf);
  R->lists[2] = DBuf<int32_t>(P, 4 * Kf);
  launch_emit_edges(Ealpha, ecol.get(), eleft.get(), eflag.get(), pos_e.get(),
                    perm.get(), R->lists[0].get(), L);
  launch_emit_tris(Tf, ftri.get(), ecol.get(), eleft.get(), tk.get(),
                   tedge.get(), perm.get(), R->lists[1].get(), L);
  launch_emit_tets(Kf, ftets.get(), perm.get(), R->lists[2].get(), L);
  KCHK();
  R->counts[1] = Ef;
  R->counts[2] = Tf;
  R->counts[3] = Kf;
  st.edges_final = Ef;
  st.tris_final = Tf;
  st.tets_final = Kf;
  T.mark(8);

  unsigned long long hs[CTR_COUNT * CTR_SLOTS], h[CTR_COUNT];
  HIPCHK(hipMemcpyAsync(hs, ctr.get(), sizeof(hs), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (int c = 0; c < CTR_COUNT; ++c) {
    h[c] = 0;
    for (int k = 0; k < CTR_SLOTS; ++k) h[c] += hs[c * CTR_SLOTS + k];
  }
  for (int f = 0; f < 6; ++f)
    for (int q = 0; q < 4; ++q)
      st.stage_evals[f][q] = (int64_t)h[CTR_STAGE + 4 * f + q];
  /* hot-stage evaluations: one per candidate for the filters, one per
   * (simplex, ball) pair tested for the emptiness scans */
  const int hot = SG.staged ? 0 : 1;
  st.tri_candidates = (int64_t)h[CTR_CAND + 1];
  st.tet_candidates = (int64_t)h[CTR_CAND + 2];
  host_evals[1] = st.tri_candidates;
  host_evals[2] = st.tet_candidates;
  for (int f = 0; f < 3; ++f) st.stage_evals[f][hot] += host_evals[f];
  for (int q = 0; q < 3; ++q) st.stage_evals[3 + q][hot] += (int64_t)h[CTR_PAIRS + q];
  st.sos_calls = (int64_t)h[CTR_SOS];
  for (int q = 0; q < 3; ++q) st.ac2_pairs[q] = (int64_t)h[CTR_PAIRS + q];
}

/* copy raw input with solvent-added radii, original order (kept for rho /
 * adjoint queries) */
__global__ void k_balls_orig(const float4* raw, int64_t n, float solvent,
                             double4* out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 a = raw[i];
  out[i] = make_double4(a.x, a.y, a.z, (double)__fadd_rn(a.w, solvent));
}

/* canonical order: vertices of each row by original id, rows lexicographic */
static void canonicalize(Engine* E, pac_result* R, int64_t n, hipStream_t s) {
  int idbits = 1;
  while (idbits < 32 && ((int64_t)1 << idbits) < n) ++idbits;
  for (int d = 1; d <= 3; ++d) {
    int64_t m = R->counts[d];
    if (m <= 0) continue;
    DBuf<uint32_t> a(&E->pool, m), b(&E->pool, m), c(&E->pool, m),
        e(&E->pool, m);
    DBuf<int32_t> copy(&E->pool, m * (d + 1));
    size_t rtb = radix_tmp_bytes(m);
    DBuf<uint8_t> rt(&E->pool, rtb);
    canonical_rows(d + 1, m, R->lists[d - 1].get(), idbits, a.get(), b.get(),
                   c.get(), e.get(), rt.get(), rtb, copy.get(), s);
    KCHK();
    HIPCHK(hipStreamSynchronize(s));
  }
}

static void fill_phase_ms(Timer& T, pac_stats& st) {
  for (int q = 0; q < 8; ++q) st.phase_ms[q] = T.ms(q, q + 1);
}

static void compute_f32(const float* xyzr, int64_t n, float solvent,
                        float alpha, const pac_options* o, pac_result** out) {
  if (!out) fail(PAC_ERR_INVALID, "out is NULL");
  if (n < 0 || (n > 0 && !xyzr)) fail(PAC_ERR_INVALID, "bad input array");
  if (!std::isfinite(solvent) || !std::isfinite(alpha))
    fail(PAC_ERR_INVALID, "non-finite solvent or alpha");
  pac_options opt;
  pac_options_init(&opt);
  if (o) opt = *o;
  Engine* E = engine_for(opt.device);
  std::lock_guard<std::mutex> lk(E->run_mu);
  HIPCHK(hipSetDevice(E->device));
  hipStream_t s = opt.stream ? (hipStream_t)opt.stream : E->stream;
  std::unique_ptr<pac_result> R(new pac_result());
  R->eng = E;
  R->alpha = alpha;
  std::memset(&R->st, 0, sizeof(R->st));
  R->counts[0] = n;
  Timer T(s);
  KProf KP;
  KP.on = opt.profile != 0;
  KP.s = s;
  T.mark(0);
  auto t_host0 = std::chrono::steady_clock::now();
  DBuf<float4> raw_own;
  const float4* raw = (const float4*)xyzr;
  if (n > 0 && !opt.device_inputs) {
    raw_own = DBuf<float4>(&E->pool, n);
    HIPCHK(hipMemcpyAsync(raw_own.get(), xyzr, (size_t)n * 16,
                          hipMemcpyHostToDevice, s));
    raw = raw_own.get();
  }
  T.mark(1);
  if (n > 0) {
    R->balls = DBuf<double4>(&E->pool, n);
    hipLaunchKernelGGL(k_balls_orig, dim3((unsigned)((n + 255) / 256)),
                       dim3(256), 0, s, raw, n, solvent, R->balls.get());
    KCHK();
    run_pipeline(E, raw, n, solvent, alpha, opt, R.get(), s, T, KP);
    if (opt.order == PAC_ORDER_CANONICAL) canonicalize(E, R.get(), n, s);
  } else {
    for (int i = 2; i <= 8; ++i) T.mark(i);
  }
  HIPCHK(hipStreamSynchronize(s));
  auto t_host1 = std::chrono::steady_clock::now();
  pac_stats& st = R->st;
  fill_phase_ms(T, st);
  st.total_ms =
      std::chrono::duration<double, std::milli>(t_host1 - t_host0).count();
  KP.collect(st.kernel_ms);
  *out = R.release();
}

/* ------------------------------------------------------------------ */
/* adjoint of the alpha values                                         */
/* ------------------------------------------------------------------ */
static bool is_device_ptr(const void* p) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

/* dL/d(x,y,z) and dL/dr of every ball (input order) from dL/drho of the
 * output triangles and tets; either gradient input may be NULL (zero).
 * Pointers may be host or device memory. */
static void alpha_adjoint(const pac_result* r, const double* gtri,
                          const double* gtet, double* dxyz, double* dr) {
  if (!r || !dxyz || !dr) fail(PAC_ERR_INVALID, "NULL argument");
  Engine* E = r->eng;
  std::lock_guard<std::mutex> lk(E->run_mu);
  HIPCHK(hipSetDevice(E->device));
  hipStream_t s = E->stream;
  Pool* P = &E->pool;
  const int64_t n = r->counts[0], T = r->counts[2], K = r->counts[3];
  const bool dev_out = is_device_ptr(dxyz);
  DBuf<double> gx(P, 3 * std::max<int64_t>(n, 1)), gr(P, std::max<int64_t>(n, 1));
  if (n > 0) {
    const int64_t F = 3 * T + 4 * K;
    DBuf<double> g(P, T + K + 1);
    const double* src[2] = {gtri, gtet};
    const int64_t cnt[2] = {T, K}, at[2] = {0, T};
    for (int q = 0; q < 2; ++q) {
      if (!cnt[q]) continue;
      if (src[q])
        HIPCHK(hipMemcpyAsync(g.get() + at[q], src[q], cnt[q] * sizeof(double),
                              is_device_ptr(src[q]) ? hipMemcpyDeviceToDevice
                                                    : hipMemcpyHostToDevice, s));
      else
        hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)((cnt[q] + 255) / 256)),
                           dim3(256), 0, s,
                           (unsigned long long*)(g.get() + at[q]), (int)cnt[q]);
    }
    DBuf<uint32_t> key(P, F + 1), val(P, F + 1), key2(P, F + 1), val2(P, F + 1);
    DBuf<double4> contrib(P, F + 1);
    launch_adj_contrib(2, T, r->lists[1].get(), r->balls.get(), g.get(), 0,
                       key.get(), val.get(), contrib.get(), s);
    launch_adj_contrib(3, K, r->lists[2].get(), r->balls.get(), g.get() + T,
                       3 * T, key.get(), val.get(), contrib.get(), s);
    KCHK();
    int idbits = 1;
    while (idbits < 32 && ((int64_t)1 << idbits) < n) ++idbits;
    size_t rtb = radix_tmp_bytes(std::max<int64_t>(F, 1));
    DBuf<uint8_t> rt(P, rtb);
    bool alt = false;
    if (F > 0)
      radix_sort_pairs(key.get(), val.get(), key2.get(), val2.get(), F, idbits,
                       rt.get(), rtb, s, &alt);
    const uint32_t* sk = alt ? key2.get() : key.get();
    const uint32_t* sv = alt ? val2.get() : val.get();
    DBuf<uint32_t> vstart(P, n + 1);
    launch_cell_starts(sk, F, n, vstart.get(), PipeLaunch{s, kBlock});
    launch_adj_gather(n, vstart.get(), sv, contrib.get(), gx.get(), gr.get(), s);
    KCHK();
  }
  HIPCHK(hipMemcpyAsync(dxyz, gx.get(), 3 * n * sizeof(double),
                        dev_out ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(dr, gr.get(), n * sizeof(double),
                        is_device_ptr(dr) ? hipMemcpyDeviceToDevice
                                          : hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
}

/* ------------------------------------------------------------------ */
/* batched ensembles                                                   */
/* ------------------------------------------------------------------ */
struct pac_batch_result {
  Engine* eng = nullptr;
  int64_t nmol = 0;
  std::vector<int64_t> off;                        /* ball offsets, nmol+1 */
  std::vector<std::unique_ptr<pac_result>> groups; /* one per pipeline pass */
  std::vector<int64_t> group_first;                /* first molecule of each group */
  std::vector<int32_t> mol_group;
  std::vector<int64_t> counts;                     /* nmol x 4 */
  std::vector<uint64_t> row0[3];                   /* first row of each molecule */
  pac_stats st;
};

static void add_stats(pac_stats& a, const pac_stats& b) {
  a.n += b.n;
  a.edge_candidates += b.edge_candidates;
  a.edge_alpha += b.edge_alpha;
  a.edges_final += b.edges_final;
  a.tri_candidates += b.tri_candidates;
  a.tri_alpha += b.tri_alpha;
  a.tris_final += b.tris_final;
  a.tet_candidates += b.tet_candidates;
  a.tet_alpha += b.tet_alpha;
  a.tets_final += b.tets_final;
  for (int f = 0; f < 6; ++f)
    for (int q = 0; q < 4; ++q) a.stage_evals[f][q] += b.stage_evals[f][q];
  a.sos_calls += b.sos_calls;
  for (int q = 0; q < 3; ++q) a.ac2_pairs[q] += b.ac2_pairs[q];
  a.k_enum = std::max(a.k_enum, b.k_enum);
  a.k_ac2 = std::max(a.k_ac2, b.k_ac2);
  a.scale_bits = std::max(a.scale_bits, b.scale_bits);
  a.grid_cells = (int32_t)std::min<int64_t>((int64_t)a.grid_cells + b.grid_cells,
                                            INT32_MAX);
  for (int q = 0; q < 8; ++q) {
    a.phase_ms[q] += b.phase_ms[q];
    a.kernel_ms[q] += b.kernel_ms[q];
  }
}

/* Molecules are packed, in order, into passes of at most `cap` balls (a
 * molecule larger than cap gets a pass of its own); each pass runs the
 * whole pipeline once over its concatenated balls, every molecule with its
 * own grid. */
static void compute_batch(const float* xyzr, const int64_t* offsets,
                          int64_t nmol, float solvent, float alpha,
                          const pac_options* o, pac_batch_result** out) {
  if (!out) fail(PAC_ERR_INVALID, "out is NULL");
  if (nmol < 0 || (nmol > 0 && !offsets)) fail(PAC_ERR_INVALID, "bad offsets");
  if (!std::isfinite(solvent) || !std::isfinite(alpha))
    fail(PAC_ERR_INVALID, "non-finite solvent or alpha");
  if (nmol > 0 && offsets[0] != 0) fail(PAC_ERR_INVALID, "offsets[0] must be 0");
  for (int64_t m = 0; m < nmol; ++m)
    if (offsets[m + 1] < offsets[m])
      fail(PAC_ERR_INVALID, "offsets must be non-decreasing");
  const int64_t ntot = nmol > 0 ? offsets[nmol] : 0;
  if (ntot > 0 && !xyzr) fail(PAC_ERR_INVALID, "bad input array");
  pac_options opt;
  pac_options_init(&opt);
  if (o) opt = *o;
  const int64_t cap = opt.batch_balls > 0 ? opt.batch_balls : (int64_t)4 << 20;
  Engine* E = engine_for(opt.device);
  std::lock_guard<std::mutex> lk(E->run_mu);
  HIPCHK(hipSetDevice(E->device));
  hipStream_t s = opt.stream ? (hipStream_t)opt.stream : E->stream;
  std::unique_ptr<pac_batch_result> B(new pac_batch_result());
  B->eng = E;
  B->nmol = nmol;
  B->off.assign(offsets, offsets + nmol + 1);
  B->counts.assign(4 * nmol, 0);
  B->mol_group.assign(nmol, 0);
  for (auto& r : B->row0) r.assign(nmol, 0);
  std::memset(&B->st, 0, sizeof(B->st));
  auto t0 = std::chrono::steady_clock::now();
  DBuf<float4> raw_own;
  const float4* raw = (const float4*)xyzr;
  if (ntot > 0 && !opt.device_inputs) {
    raw_own = DBuf<float4>(&E->pool, ntot);
    HIPCHK(hipMemcpyAsync(raw_own.get(), xyzr, (size_t)ntot * 16,
                          hipMemcpyHostToDevice, s));
    raw = raw_own.get();
  }
  for (int64_t m0 = 0; m0 < nmol;) {
    int64_t m1 = m0 + 1;
    while (m1 < nmol && offsets[m1 + 1] - offsets[m0] <= cap) ++m1;
    const int64_t gm = m1 - m0, b0 = offsets[m0], gn = offsets[m1] - b0;
    std::vector<int64_t> rel(gm + 1);
    for (int64_t q = 0; q <= gm; ++q) rel[q] = offsets[m0 + q] - b0;
    std::unique_ptr<pac_result> R(new pac_result());
    R->eng = E;
    R->alpha = alpha;
    std::memset(&R->st, 0, sizeof(R->st));
    R->counts[0] = gn;
    Timer T(s);
    KProf KP;
    KP.on = opt.profile != 0;
    KP.s = s;
    T.mark(0);
    T.mark(1);
    if (gn > 0) {
      run_pipeline(E, raw + b0, gn, solvent, alpha, opt, R.get(), s, T, KP,
                   rel.data(), gm);
      if (opt.order == PAC_ORDER_CANONICAL) canonicalize(E, R.get(), gn, s);
    } else {
      for (int i = 2; i <= 8; ++i) T.mark(i);
    }
    /* per-molecule row ranges of the group's lists */
    if (gm == 1) {
      for (int d = 1; d <= 3; ++d) B->row0[d - 1][m0] = 0;
      for (int d = 1; d <= 3; ++d) B->counts[4 * m0 + d] = R->counts[d];
    } else {
      std::vector<uint32_t> mo(gm + 1);
      for (int64_t q = 0; q <= gm; ++q) mo[q] = (uint32_t)rel[q];
      DBuf<uint32_t> dmo(&E->pool, gm + 1);
      DBuf<uint64_t> rows(&E->pool, 3 * (gm + 1));
      HIPCHK(hipMemcpyAsync(dmo.get(), mo.data(), mo.size() * sizeof(uint32_t),
                            hipMemcpyHostToDevice, s));
      for (int d = 1; d <= 3; ++d)
        launch_mol_rows(R->lists[d - 1].get(), R->counts[d], d + 1, dmo.get(),
                        gm, rows.get() + (d - 1) * (gm + 1), s);
      KCHK();
      std::vector<uint64_t> hr(3 * (gm + 1));
      HIPCHK(hipMemcpyAsync(hr.data(), rows.get(), hr.size() * sizeof(uint64_t),
                            hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      for (int d = 1; d <= 3; ++d)
        for (int64_t q = 0; q < gm; ++q) {
          const uint64_t* h = hr.data() + (d - 1) * (gm + 1);
          B->row0[d - 1][m0 + q] = h[q];
          B->counts[4 * (m0 + q) + d] = (int64_t)(h[q + 1] - h[q]);
        }
    }
    for (int64_t q = 0; q < gm; ++q) {
      B->counts[4 * (m0 + q)] = rel[q + 1] - rel[q];
      B->mol_group[m0 + q] = (int32_t)B->groups.size();
    }
    HIPCHK(hipStreamSynchronize(s));
    fill_phase_ms(T, R->st);
    KP.collect(R->st.kernel_ms);
    add_stats(B->st, R->st);
    B->group_first.push_back(m0);
    B->groups.push_back(std::move(R));
    m0 = m1;
  }
  HIPCHK(hipStreamSynchronize(s));
  B->st.total_ms = std::chrono::duration<double, std::milli>(
                       std::chrono::steady_clock::now() - t0)
                       .count();
  *out = B.release();
}

/* ------------------------------------------------------------------ */
/* C ABI                                                               */
/* ------------------------------------------------------------------ */
extern "C" {

void pac_options_init(pac_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->cell_policy = PAC_CELL_EXACT;
  o->precision = PAC_PREC_STAGED;
  o->order = PAC_ORDER_REFERENCE;
}

int pac_compute(const float* xyzr, int64_t n, float solvent, float alpha,
                const pac_options* opt, pac_result** out) {
  return guarded([&] { compute_f32(xyzr, n, solvent, alpha, opt, out); });
}

int pac_compute_f64(const double*, int64_t, double, double, const pac_options*,
                    pac_result**) {
  g_err = "pac_compute_f64: not built in this version";
  return PAC_ERR_INVALID;
}

int pac_result_counts(const pac_result* r, int64_t counts[4]) {
  if (!r || !counts) return PAC_ERR_INVALID;
  for (int i = 0; i < 4; ++i) counts[i] = r->counts[i];
  return PAC_OK;
}

static int copy_list(const pac_result* r, int dim, int32_t* dst) {
  if (!r || !dst || dim < 1 || dim > 3) return PAC_ERR_INVALID;
  return guarded([&] {
    size_t bytes = (size_t)r->counts[dim] * (dim + 1) * sizeof(int32_t);
    if (!bytes) return;
    HIPCHK(hipSetDevice(r->eng->device));
    HIPCHK(hipMemcpy(dst, r->lists[dim - 1].get(), bytes,
                     hipMemcpyDeviceToHost));
  });
}
int pac_result_copy_edges(const pac_result* r, int32_t* dst) {
  return copy_list(r, 1, dst);
}
int pac_result_copy_triangles(const pac_result* r, int32_t* dst) {
  return copy_list(r, 2, dst);
}
int pac_result_copy_tets(const pac_result* r, int32_t* dst) {
  return copy_list(r, 3, dst);
}
const int32_t* pac_result_device_list(const pac_result* r, int dim) {
  if (!r || dim < 1 || dim > 3) return nullptr;
  return r->lists[dim - 1].get();
}
int pac_result_copy_alpha(const pac_result* r, int dim, double* rho) {
  if (!r || !rho || dim < 1 || dim > 3) return PAC_ERR_INVALID;
  return guarded([&] {
    int64_t m = r->counts[dim];
    if (m <= 0) return;
    Engine* E = r->eng;
    std::lock_guard<std::mutex> lk(E->run_mu);
    HIPCHK(hipSetDevice(E->device));
    DBuf<double> d(&E->pool, m);
    launch_rho(dim, m, r->lists[dim - 1].get(), r->balls.get(), d.get(),
               E->stream);
    KCHK();
    HIPCHK(hipMemcpyAsync(rho, d.get(), (size_t)m * sizeof(double),
                          hipMemcpyDeviceToHost, E->stream));
    HIPCHK(hipStreamSynchronize(E->stream));
  });
}
int pac_result_stats(const pac_result* r, pac_stats* st) {
  if (!r || !st) return PAC_ERR_INVALID;
  *st = r->st;
  return PAC_OK;
}
void pac_result_free(pac_result* r) { delete r; }

int pac_compute_batch(const float* xyzr, const int64_t* offsets, int64_t n_mol,
                      float solvent, float alpha, const pac_options* opt,
                      pac_batch_result** out) {
  return guarded([&] {
    compute_batch(xyzr, offsets, n_mol, solvent, alpha, opt, out);
  });
}
int pac_batch_counts(const pac_batch_result* r, int64_t* out) {
  if (!r || !out) return PAC_ERR_INVALID;
  std::memcpy(out, r->counts.data(), r->counts.size() * sizeof(int64_t));
  return PAC_OK;
}
int pac_batch_copy(const pac_batch_result* r, int64_t m, int dim, int32_t* dst) {
  if (!r || !dst || dim < 1 || dim > 3 || m < 0 || m >= r->nmol)
    return PAC_ERR_INVALID;
  return guarded([&] {
    int64_t cnt = r->counts[4 * m + dim];
    if (cnt <= 0) return;
    const pac_result* g = r->groups[r->mol_group[m]].get();
    size_t k = (size_t)(dim + 1);
    HIPCHK(hipSetDevice(r->eng->device));
    HIPCHK(hipMemcpy(dst, g->lists[dim - 1].get() + k * r->row0[dim - 1][m],
                     (size_t)cnt * k * sizeof(int32_t), hipMemcpyDeviceToHost));
    const int32_t base = (int32_t)(r->off[m] - r->off[r->group_first[r->mol_group[m]]]);
    for (size_t q = 0; q < (size_t)cnt * k; ++q) dst[q] -= base;
  });
}
int pac_batch_stats(const pac_batch_result* r, pac_stats* st) {
  if (!r || !st) return PAC_ERR_INVALID;
  *st = r->st;
  return PAC_OK;
}
void pac_batch_free(pac_batch_result* r) { delete r; }
int pac_alpha_adjoint(const pac_result* r, const double* dl_drho_tri,
                      const double* dl_drho_tet, double* dl_dxyz,
                      double* dl_dr) {
  return guarded([&] { alpha_adjoint(r, dl_drho_tri, dl_drho_tet, dl_dxyz, dl_dr); });
}

const char* pac_last_error(void) { return g_err.c_str(); }
const char* pac_version(void) { return "parallelac-mi355x 0.1"; }

/* test hook: the SoS term table compiled into the device code */
int pac_internal_sos_table(int32_t* out3, int cap) {
  return sos_table_host(out3, cap);
}

}  // extern "C"
