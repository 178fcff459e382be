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
#include <type_traits>
#include <vector>

#include "pac_host.h"

using namespace pac;

/* ------------------------------------------------------------------ */
/* errors                                                              */
/* ------------------------------------------------------------------ */
static thread_local std::string g_err;
void pac_set_error(const char* m) { g_err = m ? m : ""; }


/* ------------------------------------------------------------------ */
/* engines                                                             */
/* ------------------------------------------------------------------ */
static std::mutex g_eng_mu;
static std::map<int, Engine*> g_engines;
namespace pac {
Engine* engine_for(int dev) {
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
void fill_u8(uint8_t* p, int64_t n, uint8_t v, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_u8, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     s, p, n, v);
}
void zero_u32(unsigned* p, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_u32, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
}
void zero_u64(unsigned long long* p, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)((n + 255) / 256)), dim3(256),
                     0, s, p, (int)n);
}
}  // namespace pac


/* copy raw input with solvent-added radii, original order (kept for rho /
 * adjoint queries) */
__device__ __forceinline__ double add_r_orig(float a, float s) {
  return (double)__fadd_rn(a, s);
}
__device__ __forceinline__ double add_r_orig(double a, double s) { return a + s; }
template <class B, class R>
__global__ void k_balls_orig(const B* raw, int64_t n, R solvent, double4* out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  B a = raw[i];
  out[i] = make_double4(a.x, a.y, a.z, add_r_orig(a.w, solvent));
}

static void run_variant(Engine* E, const float4* raw, int64_t n, float solvent,
                        float alpha, const pac_options& opt, pac_result* R,
                        hipStream_t s, Timer& T, KProf& KP) {
  vf32::run_pipeline(E, raw, n, solvent, alpha, opt, R, s, T, KP);
}
static void run_variant(Engine* E, const double4* raw, int64_t n,
                        double solvent, double alpha, const pac_options& opt,
                        pac_result* R, hipStream_t s, Timer& T, KProf& KP) {
  vf64::run_pipeline(E, raw, n, solvent, alpha, opt, R, s, T, KP);
}
static void run_chunked_variant(Engine* E, const float4* raw, int64_t n,
                                float solvent, float alpha,
                                const pac_options& opt, pac_result* R,
                                hipStream_t s, Timer& T, KProf& KP) {
  vf32::run_chunked(E, raw, n, solvent, alpha, opt, R, s, T, KP);
}
static void run_chunked_variant(Engine* E, const double4* raw, int64_t n,
                                double solvent, double alpha,
                                const pac_options& opt, pac_result* R,
                                hipStream_t s, Timer& T, KProf& KP) {
  vf64::run_chunked(E, raw, n, solvent, alpha, opt, R, s, T, KP);
}

/* canonical order: vertices of each row by original id, rows lexicographic */
static void canonicalize(Engine* E, pac_result* R, int64_t n, hipStream_t s) {
  int idbits = 1;
  while (idbits < 32 && ((int64_t)1 << idbits) < n) ++idbits;
  for (int d = 1; d <= 3; ++d) {
    int64_t m = R->counts[d];
    if (m <= 0) continue;
    Pool* P = pool_of(E);
    DBuf<uint32_t> a(P, m), b(P, m), c(P, m), e(P, m);
    DBuf<int32_t> copy(P, m * (d + 1));
    size_t rtb = radix_tmp_bytes(m);
    DBuf<uint8_t> rt(P, rtb);
    canonical_rows(d + 1, m, R->lists[d - 1].get(), idbits, a.get(), b.get(),
                   c.get(), e.get(), rt.get(), rtb, copy.get(), s);
    KCHK();
    HIPCHK(hipStreamSynchronize(s));
  }
}

static void fill_phase_ms(Timer& T, pac_stats& st) {
  for (int q = 0; q < 8; ++q) st.phase_ms[q] = T.ms(q, q + 1);
}

/* one molecule; F is float (FP32 balls, the reference's input) or double
 * (FP64 balls) */
template <class F>
static void compute_one(const F* xyzr, int64_t n, F solvent, F alpha,
                        const pac_options* o, pac_result** out) {
  using B = typename std::conditional<sizeof(F) == 4, float4, double4>::type;
  if (!out) fail(PAC_ERR_INVALID, "out is NULL");
  if (n < 0 || (n > 0 && !xyzr)) fail(PAC_ERR_INVALID, "bad input array");
  if (!std::isfinite(solvent) || !std::isfinite(alpha))
    fail(PAC_ERR_INVALID, "non-finite solvent or alpha");
  pac_options opt;
  pac_options_init(&opt);
  if (o) opt = *o;
  if (opt.chunk_balls < 0) fail(PAC_ERR_INVALID, "chunk_balls < 0");
  if (opt.filtration && opt.order != PAC_ORDER_REFERENCE)
    fail(PAC_ERR_INVALID, "filtration keeps the reference row order");
  if (opt.filtration && opt.skip_tet_ac2)
    fail(PAC_ERR_INVALID, "filtration needs the tet emptiness test");
  Engine* E = engine_for(opt.device);
  std::lock_guard<std::mutex> lk(E->run_mu);
  HIPCHK(hipSetDevice(E->device));
  hipStream_t s = opt.stream ? (hipStream_t)opt.stream : E->stream;
  std::unique_ptr<pac_result> R(new pac_result());
  R->eng = E;
  R->alpha = alpha;
  R->filt = opt.filtration != 0;
  std::memset(&R->st, 0, sizeof(R->st));
  R->counts[0] = n;
  Timer T(s);
  KProf KP;
  KP.on = opt.profile != 0;
  KP.s = s;
  T.mark(0);
  auto t_host0 = std::chrono::steady_clock::now();
  DBuf<B> raw_own;
  const B* raw = (const B*)xyzr;
  if (n > 0 && !opt.device_inputs) {
    raw_own = DBuf<B>(&E->pool, n);
    HIPCHK(hipMemcpyAsync(raw_own.get(), xyzr, (size_t)n * sizeof(B),
                          hipMemcpyHostToDevice, s));
    raw = raw_own.get();
  }
  T.mark(1);
  if (n > 0) {
    R->balls = DBuf<double4>(&E->pool, n);
    hipLaunchKernelGGL((k_balls_orig<B, F>), dim3((unsigned)((n + 255) / 256)),
                       dim3(256), 0, s, raw, n, solvent, R->balls.get());
    KCHK();
    if (opt.chunk_balls > 0)
      run_chunked_variant(E, raw, n, solvent, alpha, opt, R.get(), s, T, KP);
    else
      run_variant(E, raw, n, solvent, alpha, opt, R.get(), s, T, KP);
    if (opt.order == PAC_ORDER_CANONICAL) canonicalize(E, R.get(), n, s);
  } else {
    for (int i = 2; i <= 8; ++i) T.mark(i);
  }
  HIPCHK(hipStreamSynchronize(s));
  auto t_host1 = std::chrono::steady_clock::now();
  pac_stats& st = R->st;
  fill_phase_ms(T, st);
  for (int q = 0; q < 8; ++q) st.phase_ms[q] += R->phase_add[q];
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
        zero_u64((unsigned long long*)(g.get() + at[q]), cnt[q], s);
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
    vf32::launch_cell_starts(sk, F, n, vstart.get(), PipeLaunch{s, kBlock});
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
void pac::add_stats(pac_stats& a, const pac_stats& b) {
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
void pac::compute_batch(const float* xyzr, const int64_t* offsets,
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
  if (opt.chunk_balls != 0)
    fail(PAC_ERR_INVALID, "chunk_balls applies to single inputs, not batches");
  if (opt.filtration != 0)
    fail(PAC_ERR_INVALID, "filtration applies to single inputs, not batches");
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
      vf32::run_pipeline(E, raw + b0, gn, solvent, alpha, opt, R.get(), s, T, KP,
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
/* autotuner                                                           */
/* ------------------------------------------------------------------ */
/* Times each candidate (phase-one pair budget x second-phase grid) on the
 * sample, best of two runs after a warm-up run, wall clock around the
 * whole call (what a caller sees), and keeps the fastest. */
static void autotune(const float* xyzr, int64_t n, const int64_t* offsets,
                     int64_t nmol, float solvent, float alpha, pac_options* o,
                     double* best_ms) {
  if (!o) fail(PAC_ERR_INVALID, "opt is NULL");
  pac_options base = *o;
  int64_t ns = n, ms = nmol;
  if (offsets) {
    if (nmol <= 0 || offsets[0] != 0) fail(PAC_ERR_INVALID, "bad offsets");
    ms = 1;
    while (ms < nmol && offsets[ms + 1] <= ((int64_t)1 << 20)) ++ms;
    ns = offsets[ms];
  }
  if (ns <= 0) {
    if (best_ms) *best_ms = 0;
    return;
  }
  Engine* E = engine_for(base.device);
  DBuf<float4> dev;
  const float* src = xyzr;
  if (!base.device_inputs) {
    {
      std::lock_guard<std::mutex> lk(E->run_mu);
      HIPCHK(hipSetDevice(E->device));
      dev = DBuf<float4>(&E->pool, ns);
      HIPCHK(hipMemcpy(dev.get(), xyzr, (size_t)ns * 16, hipMemcpyHostToDevice));
    }
    src = (const float*)dev.get();
    base.device_inputs = 1;
  }
  /* device time of one run with settings t: best of two after a warm-up */
  auto time_one = [&](const pac_options& t) {
    double tmin = 1e300;
    for (int rep = 0; rep < 3; ++rep) {
      auto t0 = std::chrono::steady_clock::now();
      if (offsets) {
        pac_batch_result* r = nullptr;
        compute_batch(src, offsets, ms, solvent, alpha, &t, &r);
        delete r;
      } else {
        pac_result* r = nullptr;
        compute_one<float>(src, ns, solvent, alpha, &t, &r);
        delete r;
      }
      double dt = std::chrono::duration<double, std::milli>(
                      std::chrono::steady_clock::now() - t0).count();
      if (rep > 0) tmin = std::min(tmin, dt);
    }
    return tmin;
  };
  /* the scan grid and the quick rejection pass together (at the default
   * budget and grid), then the phase-one budget x second-phase grid over
   * them */
  static const int scans[][2] = {{2, 0}, {4, 0}, {2, 1}, {4, 1}, {-1, 0}};
  static const int budgets[] = {-1, 2, 4, 8, 16};
  static const int grids[] = {1024, 2048, 4096};
  double best = 1e300;
  pac_options pick = base;
  pick.profile = 0;
  if (base.cell_policy == PAC_CELL_EXACT)
    for (const auto& sc : scans) {
      pac_options t = pick;
      t.ac2_fine = sc[0];
      t.ac2_quick = sc[1];
      const double tm = time_one(t);
      if (tm < best) {
        best = tm;
        pick.ac2_fine = sc[0];
        pick.ac2_quick = sc[1];
      }
    }
  const pac_options fixed = pick;
  best = 1e300;
  for (int b : budgets)
    for (int g : grids) {
      if (b < 0 && g != grids[0]) continue; /* no second phase: no grid */
      pac_options t = fixed;
      t.ac2_budget = b;
      t.ac2_grid = g;
      const double tm = time_one(t);
      if (tm < best) {
        best = tm;
        pick = t;
      }
    }
  /* block shape and queue granularity of the persistent second phase
   * (with a phase-one budget: no second phase otherwise) */
  if (pick.ac2_budget >= 0) {
    static const int blocks[] = {64, 128, 256};
    static const int chunks[] = {64, 256, 1024};
    const pac_options tuned = pick;
    for (int bl : blocks)
      for (int ch : chunks) {
        pac_options t = tuned;
        t.block = bl;
        t.ac2_chunk = ch;
        const double tm = time_one(t);
        if (tm < best) {
          best = tm;
          pick = t;
        }
      }
  }
  o->ac2_budget = pick.ac2_budget;
  o->ac2_grid = pick.ac2_grid;
  o->ac2_fine = pick.ac2_fine;
  o->ac2_quick = pick.ac2_quick;
  o->block = pick.block;
  o->ac2_chunk = pick.ac2_chunk;
  if (best_ms) *best_ms = best;
}

/* ------------------------------------------------------------------ */
/* allocation-free entry point                                         */
/* ------------------------------------------------------------------ */
/* Device scratch of pac_compute_noalloc: every intermediate buffer of its
 * runs is carved from this array (Pool arena mode).  Each GPU the library
 * runs on holds its own copy until the process ends.  In host code the
 * name is a host-side stand-in, not a device address: host code uses only
 * Engine::scratch, the address hipGetSymbolAddress returned with that
 * device current. */
static __device__ __align__(256) unsigned char scratch[PAC_NOALLOC_SCRATCH_BYTES];

/* the device's scratch address, looked up once per device (engine lock
 * held, device current) */
static void* scratch_of(Engine* E) {
  if (!E->scratch) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) fail(PAC_ERR_DEVICE, "scratch symbol has no device address");
    E->scratch = p;
  }
  return E->scratch;
}

/* pac_compute over the device scratch into the caller's lists: no device
 * or pinned allocation, no hipEvent.  The engine lock serializes calls in
 * this process; the stream is drained before returning, also on errors,
 * so no kernel still uses the scratch when the next call reuses it. */
static void compute_noalloc(const float* xyzr, int64_t n, float solvent,
                            float alpha, const pac_options* o, int32_t* lists[3],
                            const int64_t caps[3], int64_t* counts) {
  if (!counts) fail(PAC_ERR_INVALID, "counts is NULL");
  if (n < 0 || (n > 0 && !xyzr)) fail(PAC_ERR_INVALID, "bad input array");
  if (n > PAC_NOALLOC_MAX_BALLS)
    fail(PAC_ERR_INVALID, "n = " + std::to_string(n) + " exceeds the "
         "allocation-free entry point's maximum of " +
         std::to_string(PAC_NOALLOC_MAX_BALLS) + " balls (use pac_compute)");
  for (int d = 0; d < 3; ++d)
    if (caps[d] < 0 || (caps[d] > 0 && !lists[d]))
      fail(PAC_ERR_INVALID, "bad output list or capacity");
  if (!std::isfinite(solvent) || !std::isfinite(alpha))
    fail(PAC_ERR_INVALID, "non-finite solvent or alpha");
  pac_options opt;
  pac_options_init(&opt);
  if (o) opt = *o;
  if (opt.chunk_balls != 0 || opt.filtration != 0 || opt.profile != 0)
    fail(PAC_ERR_INVALID, "chunk_balls, filtration and profile are not "
                          "available without allocation");
  Engine* E = engine_for(opt.device);
  std::lock_guard<std::mutex> lk(E->run_mu);
  HIPCHK(hipSetDevice(E->device));
  void* base = scratch_of(E);
  hipStream_t s = opt.stream ? (hipStream_t)opt.stream : E->stream;
  E->arena.set_arena(base, PAC_NOALLOC_SCRATCH_BYTES);
  E->run_pool = &E->arena;
  struct Restore {
    Engine* E;
    hipStream_t s;
    ~Restore() {
      (void)hipStreamSynchronize(s); /* no kernel may still use the scratch */
      E->run_pool = nullptr;
    }
  } restore{E, s};
  counts[0] = n;
  counts[1] = counts[2] = counts[3] = 0;
  if (n == 0) return;
  pac_result R;
  R.eng = E;
  R.alpha = alpha;
  std::memset(&R.st, 0, sizeof(R.st));
  R.counts[0] = n;
  Timer T(s, false);
  KProf KP;
  DBuf<float4> raw_own;
  const float4* raw = (const float4*)xyzr;
  if (!opt.device_inputs) {
    raw_own = DBuf<float4>(&E->arena, n);
    HIPCHK(hipMemcpyAsync(raw_own.get(), xyzr, (size_t)n * sizeof(float4),
                          hipMemcpyHostToDevice, s));
    raw = raw_own.get();
  }
  run_variant(E, raw, n, solvent, alpha, opt, &R, s, T, KP);
  if (opt.order == PAC_ORDER_CANONICAL) canonicalize(E, &R, n, s);
  bool over = false;
  for (int d = 1; d <= 3; ++d) {
    counts[d] = R.counts[d];
    if (R.counts[d] > caps[d - 1]) over = true;
  }
  if (over)
    fail(PAC_ERR_OVERFLOW, "an output list exceeds its capacity (counts[] "
                           "holds the sizes needed)");
  for (int d = 1; d <= 3; ++d)
    if (R.counts[d] > 0)
      HIPCHK(hipMemcpyAsync(lists[d - 1], R.lists[d - 1].get(),
                            (size_t)R.counts[d] * (d + 1) * sizeof(int32_t),
                            hipMemcpyDefault, s));
  HIPCHK(hipStreamSynchronize(s));
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
  return guarded([&] { compute_one<float>(xyzr, n, solvent, alpha, opt, out); });
}

int pac_compute_f64(const double* xyzr, int64_t n, double solvent,
                    double alpha, const pac_options* opt, pac_result** out) {
  return guarded([&] { compute_one<double>(xyzr, n, solvent, alpha, opt, out); });
}

int pac_compute_noalloc(const float* xyzr, int64_t n, float solvent,
                        float alpha, const pac_options* opt, int32_t* edges,
                        int64_t cap_edges, int32_t* triangles,
                        int64_t cap_triangles, int32_t* tets, int64_t cap_tets,
                        int64_t counts[4]) {
  int32_t* lists[3] = {edges, triangles, tets};
  const int64_t caps[3] = {cap_edges, cap_triangles, cap_tets};
  return guarded([&] {
    compute_noalloc(xyzr, n, solvent, alpha, opt, lists, caps, counts);
  });
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
    /* host or device destination (unified addressing picks the copy) */
    HIPCHK(hipMemcpy(dst, r->lists[dim - 1].get(), bytes, hipMemcpyDefault));
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
    if (is_device_ptr(rho)) { /* device destination: written in place */
      launch_rho(dim, m, r->lists[dim - 1].get(), r->balls.get(), rho,
                 E->stream);
      KCHK();
    } else {
      DBuf<double> d(&E->pool, m);
      launch_rho(dim, m, r->lists[dim - 1].get(), r->balls.get(), d.get(),
                 E->stream);
      KCHK();
      HIPCHK(hipMemcpyAsync(rho, d.get(), (size_t)m * sizeof(double),
                            hipMemcpyDeviceToHost, E->stream));
    }
    HIPCHK(hipStreamSynchronize(E->stream));
  });
}
/* filtration keys of a result, made once (engine lock held) */
static const uint64_t* filtration_keys(const pac_result* r) {
  if (!r->filt) fail(PAC_ERR_INVALID, "result computed without opt.filtration");
  Engine* E = r->eng;
  if (!r->bkey_done) {
    const int64_t m = r->counts[0] + r->counts[1] + r->counts[2] + r->counts[3];
    r->bkey = DBuf<uint64_t>(&E->pool, std::max<int64_t>(m, 1));
    DBuf<unsigned> miss(&E->pool, 1);
    zero_u32(miss.get(), 1, E->stream);
    const int32_t* lists[3] = {r->lists[0].get(), r->lists[1].get(),
                               r->lists[2].get()};
    const uint8_t* gab[2] = {r->gab[0].get(), r->gab[1].get()};
    birth_keys(lists, r->counts, r->balls.get(), gab, r->rank.get(),
               r->bkey.get(), miss.get(), E->stream);
    KCHK();
    unsigned h = 0;
    HIPCHK(hipMemcpyAsync(&h, miss.get(), sizeof(h), hipMemcpyDeviceToHost,
                          E->stream));
    HIPCHK(hipStreamSynchronize(E->stream));
    if (h) fail(PAC_ERR_DEVICE, "internal: a face missing from the lists");
    r->bkey_done = true;
  }
  return r->bkey.get();
}
int pac_result_copy_birth(const pac_result* r, int dim, double* value) {
  if (!r || !value || dim < 0 || dim > 3) return PAC_ERR_INVALID;
  return guarded([&] {
    Engine* E = r->eng;
    std::lock_guard<std::mutex> lk(E->run_mu);
    HIPCHK(hipSetDevice(E->device));
    const uint64_t* key = filtration_keys(r);
    const int64_t m = r->counts[dim];
    if (m <= 0) return;
    int64_t off = 0;
    for (int d = 0; d < dim; ++d) off += r->counts[d];
    DBuf<double> out(&E->pool, m);
    DBuf<unsigned> miss(&E->pool, 1);
    zero_u32(miss.get(), 1, E->stream);
    birth_values(m, key + off, out.get(), miss.get(), E->stream);
    KCHK();
    unsigned h = 0;
    HIPCHK(hipMemcpyAsync(value, out.get(), (size_t)m * sizeof(double),
                          hipMemcpyDeviceToHost, E->stream));
    HIPCHK(hipMemcpyAsync(&h, miss.get(), sizeof(h), hipMemcpyDeviceToHost,
                          E->stream));
    HIPCHK(hipStreamSynchronize(E->stream));
    if (h) fail(PAC_ERR_DEVICE, "internal: a simplex without a filtration value");
  });
}
int pac_result_copy_filtration(const pac_result* r, int32_t* dim,
                               int64_t* index, double* value) {
  if (!r || !dim || !index || !value) return PAC_ERR_INVALID;
  return guarded([&] {
    Engine* E = r->eng;
    std::lock_guard<std::mutex> lk(E->run_mu);
    HIPCHK(hipSetDevice(E->device));
    const uint64_t* key = filtration_keys(r);
    const int64_t m = r->counts[0] + r->counts[1] + r->counts[2] + r->counts[3];
    if (m <= 0) return;
    if (m >= ((int64_t)1 << 32)) fail(PAC_ERR_OVERFLOW, "filtration too large");
    DBuf<uint32_t> a(&E->pool, m), b(&E->pool, m), c(&E->pool, m),
        e(&E->pool, m);
    const size_t rtb = radix_tmp_bytes(m);
    DBuf<uint8_t> rt(&E->pool, rtb);
    DBuf<int32_t> dd(&E->pool, m);
    DBuf<int64_t> di(&E->pool, m);
    DBuf<double> dv(&E->pool, m);
    filtration_order(m, key, r->counts, dd.get(), di.get(), dv.get(), a.get(),
                     b.get(), c.get(), e.get(), rt.get(), rtb, E->stream);
    KCHK();
    HIPCHK(hipMemcpyAsync(dim, dd.get(), (size_t)m * sizeof(int32_t),
                          hipMemcpyDeviceToHost, E->stream));
    HIPCHK(hipMemcpyAsync(index, di.get(), (size_t)m * sizeof(int64_t),
                          hipMemcpyDeviceToHost, E->stream));
    HIPCHK(hipMemcpyAsync(value, dv.get(), (size_t)m * sizeof(double),
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
int pac_autotune(const float* xyzr, int64_t n, const int64_t* offsets,
                 int64_t n_mol, float solvent, float alpha, pac_options* opt,
                 double* best_ms) {
  return guarded([&] {
    autotune(xyzr, n, offsets, n_mol, solvent, alpha, opt, best_ms);
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

int pac_memory_stats(int device, int64_t* in_use, int64_t* peak,
                     int64_t* held, int reset_peak) {
  Engine* E = nullptr;
  {
    std::lock_guard<std::mutex> g(g_eng_mu);
    auto it = g_engines.find(device);
    if (it != g_engines.end()) E = it->second;
  }
  int64_t u = 0, p = 0, h = 0;
  if (E) {
    std::lock_guard<std::mutex> g(E->pool.m);
    u = E->pool.in_use;
    p = E->pool.peak;
    h = E->pool.held;
    if (reset_peak) E->pool.peak = E->pool.in_use;
  }
  if (in_use) *in_use = u;
  if (peak) *peak = p;
  if (held) *held = h;
  return PAC_OK;
}
int pac_memory_trim(int device) {
  Engine* E = nullptr;
  {
    std::lock_guard<std::mutex> g(g_eng_mu);
    auto it = g_engines.find(device);
    if (it != g_engines.end()) E = it->second;
  }
  if (!E) return PAC_OK;
  return guarded([&] {
    std::lock_guard<std::mutex> lk(E->run_mu);
    HIPCHK(hipSetDevice(E->device));
    HIPCHK(hipStreamSynchronize(E->stream));
    E->pool.trim();
  });
}
const char* pac_last_error(void) { return g_err.c_str(); }
const char* pac_version(void) { return "parallelac-mi355x 0.1"; }

/* test hook: the SoS term table compiled into the device code */
int pac_internal_sos_table(int32_t* out3, int cap) {
  return vf32::sos_table_host(out3, cap);
}

}  // extern "C"
