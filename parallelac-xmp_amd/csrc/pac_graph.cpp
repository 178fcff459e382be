/*
 * pac_graph.cpp -- a batch recorded once and replayed as one HIP graph.
 *
 * The eager pipeline reads counts back between its stages to size the
 * next stage's buffers.  A recorded batch instead fixes every size up
 * front: an eager run of the batch (compute_batch) measures each pass's
 * counts and largest queue fills, each count gets a capacity of that size
 * plus headroom, and run_pipeline records the pass against those
 * capacities (PassRec): item counts stay on the device (Cnt), queue views
 * are built by a kernel, each precision ladder runs once, and overflows or
 * an input that no longer fits the plan set status bits instead of being
 * retried.  All passes go into one stream capture, so a run of the whole
 * batch is one hipGraphLaunch.  Passes run one after another and share one
 * set of scratch buffers (the graph's own pool, whose blocks the capture
 * hands out in stream order); each pass keeps its own output lists, counts
 * and per-molecule row starts.
 *
 * No counterpart in the reference, which runs one molecule per process
 * with a host round trip after every phase (AlphaComplexParallel.cu:
 * 1126-1806).
 */
#include <hip/hip_runtime.h>

#include <chrono>
#include <memory>
#include <string>
#include <vector>

#include "pac_host.h"

using namespace pac;

namespace {

struct GPass {
  int64_t m0 = 0, gm = 0, b0 = 0, gn = 0;
  std::vector<int64_t> rel;  /* molecule offsets relative to b0, gm+1 */
  int64_t probe[RC_COUNT];   /* counts of the eager run */
  int64_t cap[RC_COUNT];
  vf32::PassPlan plan;
  DBuf<Bounds> bref, bnow;
  DBuf<uint64_t> dcnt;       /* RC_COUNT device counts */
  DBuf<uint32_t> dmo;        /* rel as uint32, for the row starts */
  DBuf<uint64_t> rows;       /* 3 x (gm+1) first row of each molecule */
  std::unique_ptr<pac_result> R; /* output lists (capacities) */
  pac_stats st;              /* per-pass plan statistics */
};

}  // namespace

struct pac_graph {
  /* the pool first: destroyed last, after every buffer went back to it */
  Pool pool;
  Engine* E = nullptr;
  hipStream_t s = nullptr;       /* capture stream */
  hipStream_t last = nullptr;    /* stream of the last launch */
  pac_options opt;
  float solvent = 0, alpha = 0;
  int64_t nmol = 0, ntot = 0;
  std::vector<int64_t> off;
  DBuf<float4> in;               /* owned input copy (host inputs) */
  const float4* raw = nullptr;   /* what the graph reads */
  std::vector<std::unique_ptr<GPass>> passes;
  DBuf<unsigned> status;         /* [0] REC_* bits, [1..2] largest fills */
  DBuf<unsigned long long> ctr;
  DBuf<unsigned> qn;
  DBuf<QView> views;
  DBuf<QEntry> q1, q2;
  uint32_t c1 = 0, c2 = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool launched = false;
  int64_t nodes = 0;

  ~pac_graph() {
    if (E) (void)hipSetDevice(E->device);
    (void)hipDeviceSynchronize();
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    passes.clear();
    in.release();
    status.release();
    ctr.release();
    qn.release();
    views.release();
    q1.release();
    q2.release();
    pool.trim();
    if (s) (void)hipStreamDestroy(s);
  }
};

namespace {

int64_t with_headroom(int64_t v, double h) {
  return v + (int64_t)std::ceil((double)v * h) + 256;
}

/* Enqueue every pass on G->s in recorded form (also what the capture
 * records).  Each pass's previous outputs go back to the graph's pool
 * first, so a second recording takes exactly the blocks the first one
 * left there and allocates nothing. */
void record_all(pac_graph* G) {
  hipStream_t s = G->s;
  zero_u32(G->status.get(), 4, s);
  zero_u32(G->qn.get(), QN_WORDS, s); /* then kept clear by the views */
  zero_u64(G->ctr.get(), CTR_COUNT * CTR_SLOTS, s);
  Timer T(s);
  T.on = false;
  KProf KP;
  for (auto& pp : G->passes) {
    GPass& p = *pp;
    p.R.reset();
    if (p.gn == 0) continue;
    p.R.reset(new pac_result());
    p.R->eng = G->E;
    p.R->alpha = G->alpha;
    std::memset(&p.R->st, 0, sizeof(p.R->st));
    PassRec rec;
    rec.pool = &G->pool;
    rec.plan = &p.plan;
    rec.bref = p.bref.get();
    rec.bnow = p.bnow.get();
    for (int q = 0; q < RC_COUNT; ++q) rec.cap[q] = p.cap[q];
    rec.dcnt = p.dcnt.get();
    rec.status = G->status.get();
    rec.fill = G->status.get() + 1;
    rec.ctr = G->ctr.get();
    rec.qn = G->qn.get();
    rec.views = G->views.get();
    rec.q1 = G->q1.get();
    rec.q2 = G->q2.get();
    rec.c1 = G->c1;
    rec.c2 = G->c2;
    vf32::run_pipeline(G->E, G->raw + p.b0, p.gn, G->solvent, G->alpha,
                       G->opt, p.R.get(), s, T, KP, p.rel.data(), p.gm, &rec);
    static const int fin[3] = {RC_EF, RC_TF, RC_KF};
    for (int d = 1; d <= 3; ++d)
      launch_mol_rows(p.R->lists[d - 1].get(),
                      Cnt(p.cap[fin[d - 1]], p.dcnt.get() + fin[d - 1]), d + 1,
                      p.dmo.get(), p.gm, p.rows.get() + (d - 1) * (p.gm + 1), s);
    KCHK();
  }
}

void check_status(pac_graph* G, const char* when) {
  unsigned h[4];
  HIPCHK(hipMemcpy(h, G->status.get(), sizeof(h), hipMemcpyDeviceToHost));
  if (!h[0]) return;
  std::string w = std::string(when) + ": ";
  if (h[0] & (1u << REC_STALE))
    fail(PAC_ERR_STALE, w + "the input no longer fits the recorded grids "
                            "(a molecule's bounds changed); record again");
  if (h[0] & (1u << REC_Q1))
    fail(PAC_ERR_OVERFLOW, w + "hot queue overflow (" + std::to_string(h[1]) +
                               " entries in one sub-queue of " +
                               std::to_string(G->c1) + "); record again");
  if (h[0] & (1u << REC_Q2))
    fail(PAC_ERR_OVERFLOW, w + "escalation queue overflow (" +
                               std::to_string(h[2]) + " of " +
                               std::to_string(G->c2) + "); record again");
  static const char* names[RC_COUNT] = {
      "edge candidates", "alpha edges", "triangle slots", "alpha triangles",
      "tet slots", "alpha tets", "final tets", "final triangles", "final edges"};
  for (int q = 0; q < RC_COUNT; ++q)
    if (h[0] & (1u << (REC_CAP + q)))
      fail(PAC_ERR_OVERFLOW, w + std::string(names[q]) +
                                 " exceed the recorded capacity; record again");
  fail(PAC_ERR_OVERFLOW, w + "recorded capacity exceeded");
}

void graph_create(const float* xyzr, const int64_t* offsets, int64_t nmol,
                  float solvent, float alpha, const pac_options* o,
                  double headroom, pac_graph** out) {
  if (!out) fail(PAC_ERR_INVALID, "out is NULL");
  *out = nullptr;
  pac_options opt;
  pac_options_init(&opt);
  if (o) opt = *o;
  if (opt.order != PAC_ORDER_REFERENCE)
    fail(PAC_ERR_INVALID, "a recorded batch keeps the reference row order");
  if (!(headroom >= 0)) headroom = 0.125;
  if (headroom > 64) fail(PAC_ERR_INVALID, "headroom above 64");
  /* A recorded batch keeps every pass's outputs and shares one set of
   * scratch buffers sized for its largest pass: passes of 2 Mi balls
   * (unless opt.batch_balls says otherwise) hold that scratch to half of
   * the eager default's at the cost of more graph nodes. */
  if (opt.batch_balls <= 0) opt.batch_balls = 1 << 21;
  /* the eager run: validates the arguments, fixes the passes, measures */
  pac_options popt = opt;
  popt.profile = 0;
  pac_batch_result* probe_raw = nullptr;
  compute_batch(xyzr, offsets, nmol, solvent, alpha, &popt, &probe_raw);
  std::unique_ptr<pac_batch_result> probe(probe_raw);

  std::unique_ptr<pac_graph> G(new pac_graph());
  Engine* E = engine_for(opt.device);
  std::lock_guard<std::mutex> lk(E->run_mu);
  HIPCHK(hipSetDevice(E->device));
  G->E = E;
  G->opt = opt;
  G->opt.profile = 0;
  G->solvent = solvent;
  G->alpha = alpha;
  G->nmol = nmol;
  G->off.assign(offsets, offsets + nmol + 1);
  G->ntot = nmol > 0 ? offsets[nmol] : 0;
  HIPCHK(hipStreamCreateWithFlags(&G->s, hipStreamNonBlocking));
  HIPCHK(hipEventCreate(&G->ev0));
  HIPCHK(hipEventCreate(&G->ev1));
  hipStream_t s = G->s;
  if (opt.device_inputs || G->ntot == 0) {
    G->raw = (const float4*)xyzr;
  } else {
    G->in = DBuf<float4>(&G->pool, G->ntot);
    HIPCHK(hipMemcpyAsync(G->in.get(), xyzr, (size_t)G->ntot * 16,
                          hipMemcpyHostToDevice, s));
    G->raw = G->in.get();
  }

  /* passes exactly as the eager run made them */
  int64_t fill1 = 0, fill2 = 0;
  const size_t ng = probe->groups.size();
  for (size_t g = 0; g < ng; ++g) {
    std::unique_ptr<GPass> P(new GPass());
    GPass& p = *P;
    p.m0 = probe->group_first[g];
    const int64_t m1 = g + 1 < ng ? probe->group_first[g + 1] : nmol;
    p.gm = m1 - p.m0;
    p.b0 = offsets[p.m0];
    p.gn = offsets[m1] - p.b0;
    p.rel.resize(p.gm + 1);
    for (int64_t q = 0; q <= p.gm; ++q) p.rel[q] = offsets[p.m0 + q] - p.b0;
    const pac_result& pr = *probe->groups[g];
    for (int q = 0; q < RC_COUNT; ++q) {
      p.probe[q] = pr.sizes[q];
      p.cap[q] = with_headroom(pr.sizes[q], headroom);
      if (p.cap[q] >= ((int64_t)1 << 32) - 1)
        p.cap[q] = ((int64_t)1 << 32) - 2; /* survivors index with uint32 */
    }
    fill1 = std::max(fill1, pr.sizes[9]);
    fill2 = std::max(fill2, pr.sizes[10]);
    p.st = pr.st;
    if (p.gn > 0) {
      vf32::plan_pass(E, G->raw + p.b0, p.gn, solvent, alpha, G->opt, s,
                      p.rel.data(), p.gm, p.plan);
      const size_t nb = p.plan.bounds.size();
      p.bref = DBuf<Bounds>(&G->pool, nb);
      p.bnow = DBuf<Bounds>(&G->pool, nb);
      HIPCHK(hipMemcpyAsync(p.bref.get(), p.plan.bounds.data(),
                            nb * sizeof(Bounds), hipMemcpyHostToDevice, s));
      p.dcnt = DBuf<uint64_t>(&G->pool, RC_COUNT);
      std::vector<uint32_t> mo(p.gm + 1);
      for (int64_t q = 0; q <= p.gm; ++q) mo[q] = (uint32_t)p.rel[q];
      p.dmo = DBuf<uint32_t>(&G->pool, p.gm + 1);
      HIPCHK(hipMemcpyAsync(p.dmo.get(), mo.data(), mo.size() * sizeof(uint32_t),
                            hipMemcpyHostToDevice, s));
      p.rows = DBuf<uint64_t>(&G->pool, 3 * (p.gm + 1));
      HIPCHK(hipStreamSynchronize(s));
    }
    G->passes.push_back(std::move(P));
  }
  /* queues: the largest sub-queue fill seen, plus headroom; an
   * escalation kernel pushes one entry per consumed entry except for
   * full-scan entries, so the escalation queue is at least the hot one */
  const int64_t qmax = (int64_t)(UINT32_MAX / QSLOTS) - 1;
  G->c1 = (uint32_t)std::min(with_headroom(fill1, headroom) + 256, qmax);
  G->c2 = (uint32_t)std::min(
      std::max(with_headroom(fill2, headroom) + 256, (int64_t)G->c1 + 512), qmax);
  G->status = DBuf<unsigned>(&G->pool, 4);
  G->ctr = DBuf<unsigned long long>(&G->pool, CTR_COUNT * CTR_SLOTS);
  G->qn = DBuf<unsigned>(&G->pool, QN_WORDS);
  G->views = DBuf<QView>(&G->pool, 2);
  G->q1 = DBuf<QEntry>(&G->pool, (size_t)G->c1 * QSLOTS);
  G->q2 = DBuf<QEntry>(&G->pool, (size_t)G->c2 * QSLOTS);

  /* one uncaptured run of the recording: it must reproduce the eager
   * counts exactly */
  record_all(G.get());
  HIPCHK(hipStreamSynchronize(s));
  check_status(G.get(), "recorded run");
  for (auto& pp : G->passes) {
    GPass& p = *pp;
    if (p.gn == 0) continue;
    uint64_t h[RC_COUNT];
    HIPCHK(hipMemcpy(h, p.dcnt.get(), sizeof(h), hipMemcpyDeviceToHost));
    for (int q = 0; q < RC_COUNT; ++q)
      if ((int64_t)h[q] != p.probe[q])
        fail(PAC_ERR_DEVICE, "recorded run count " + std::to_string(q) + " = " +
                                 std::to_string(h[q]) + ", eager run " +
                                 std::to_string(p.probe[q]));
  }

  /* the capture: the same enqueue sequence, which now finds every block
   * it needs in the graph's pool */
  HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
  try {
    record_all(G.get());
  } catch (...) {
    hipGraph_t broken = nullptr;
    (void)hipStreamEndCapture(s, &broken);
    if (broken) (void)hipGraphDestroy(broken);
    throw;
  }
  HIPCHK(hipStreamEndCapture(s, &G->graph));
  HIPCHK(hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0));
  size_t nn = 0;
  HIPCHK(hipGraphGetNodes(G->graph, nullptr, &nn));
  G->nodes = (int64_t)nn;
  *out = G.release();
}

void graph_set_input(pac_graph* G, const float* xyzr, void* stream) {
  if (!G || (!xyzr && G->ntot > 0)) fail(PAC_ERR_INVALID, "bad argument");
  if (G->opt.device_inputs)
    fail(PAC_ERR_INVALID, "the graph reads the caller's device array: write "
                          "that array instead");
  if (G->ntot == 0) return;
  HIPCHK(hipSetDevice(G->E->device));
  hipStream_t st = stream ? (hipStream_t)stream : G->s;
  hipPointerAttribute_t a;
  bool dev = hipPointerGetAttributes(&a, xyzr) == hipSuccess &&
             a.type == hipMemoryTypeDevice;
  (void)hipGetLastError();
  HIPCHK(hipMemcpyAsync(G->in.get(), xyzr, (size_t)G->ntot * 16,
                        dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                        st));
  if (!dev) HIPCHK(hipStreamSynchronize(st)); /* host memory may be pageable */
}

void graph_launch(pac_graph* G, void* stream) {
  if (!G) fail(PAC_ERR_INVALID, "graph is NULL");
  HIPCHK(hipSetDevice(G->E->device));
  hipStream_t st = stream ? (hipStream_t)stream : G->s;
  HIPCHK(hipEventRecord(G->ev0, st));
  HIPCHK(hipGraphLaunch(G->exec, st));
  HIPCHK(hipEventRecord(G->ev1, st));
  G->last = st;
  G->launched = true;
}

void graph_result(pac_graph* G, pac_batch_result** out) {
  if (!G || !out) fail(PAC_ERR_INVALID, "bad argument");
  if (!G->launched) fail(PAC_ERR_INVALID, "the graph has not been launched");
  Engine* E = G->E;
  HIPCHK(hipSetDevice(E->device));
  HIPCHK(hipStreamSynchronize(G->last));
  check_status(G, "graph launch");
  std::unique_ptr<pac_batch_result> B(new pac_batch_result());
  B->eng = E;
  B->nmol = G->nmol;
  B->off = G->off;
  B->counts.assign(4 * G->nmol, 0);
  B->mol_group.assign(G->nmol, 0);
  for (auto& r : B->row0) r.assign(G->nmol, 0);
  std::memset(&B->st, 0, sizeof(B->st));
  int64_t ecand = 0;
  for (auto& pp : G->passes) {
    GPass& p = *pp;
    std::unique_ptr<pac_result> R(new pac_result());
    R->eng = E;
    R->alpha = G->alpha;
    std::memset(&R->st, 0, sizeof(R->st));
    R->counts[0] = p.gn;
    std::vector<uint64_t> hr(3 * (p.gm + 1), 0);
    if (p.gn > 0) {
      uint64_t h[RC_COUNT];
      HIPCHK(hipMemcpy(h, p.dcnt.get(), sizeof(h), hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(hr.data(), p.rows.get(), hr.size() * sizeof(uint64_t),
                       hipMemcpyDeviceToHost));
      static const int fin[3] = {RC_EF, RC_TF, RC_KF};
      for (int d = 1; d <= 3; ++d) {
        const int64_t m = (int64_t)h[fin[d - 1]];
        R->counts[d] = m;
        R->lists[d - 1] = DBuf<int32_t>(&E->pool, (size_t)m * (d + 1));
        if (m)
          HIPCHK(hipMemcpyAsync(R->lists[d - 1].get(), p.R->lists[d - 1].get(),
                                (size_t)m * (d + 1) * sizeof(int32_t),
                                hipMemcpyDeviceToDevice, G->s));
      }
      for (int q = 0; q < RC_COUNT; ++q) R->sizes[q] = (int64_t)h[q];
      pac_stats& st = R->st;
      st.n = p.gn;
      st.edge_candidates = (int64_t)h[RC_ECAND];
      st.edge_alpha = (int64_t)h[RC_EALPHA];
      st.tri_alpha = (int64_t)h[RC_TALPHA];
      st.tet_alpha = (int64_t)h[RC_KALPHA];
      st.edges_final = (int64_t)h[RC_EF];
      st.tris_final = (int64_t)h[RC_TF];
      st.tets_final = (int64_t)h[RC_KF];
      st.k_enum = p.st.k_enum;
      st.k_ac2 = p.st.k_ac2;
      st.scale_bits = p.st.scale_bits;
      st.grid_cells = p.st.grid_cells;
      ecand += st.edge_candidates;
    }
    for (int d = 1; d <= 3; ++d)
      for (int64_t q = 0; q < p.gm; ++q) {
        const uint64_t* h = hr.data() + (d - 1) * (p.gm + 1);
        B->row0[d - 1][p.m0 + q] = h[q];
        B->counts[4 * (p.m0 + q) + d] = (int64_t)(h[q + 1] - h[q]);
      }
    for (int64_t q = 0; q < p.gm; ++q) {
      B->counts[4 * (p.m0 + q)] = p.rel[q + 1] - p.rel[q];
      B->mol_group[p.m0 + q] = (int32_t)B->groups.size();
    }
    add_stats(B->st, R->st);
    B->group_first.push_back(p.m0);
    B->groups.push_back(std::move(R));
  }
  std::vector<unsigned long long> hs(CTR_COUNT * CTR_SLOTS);
  HIPCHK(hipMemcpyAsync(hs.data(), G->ctr.get(), hs.size() * sizeof(hs[0]),
                        hipMemcpyDeviceToHost, G->s));
  HIPCHK(hipStreamSynchronize(G->s));
  stats_from_counters(hs.data(), G->opt.precision == PAC_PREC_STAGED, ecand,
                      B->st);
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, G->ev0, G->ev1));
  B->st.total_ms = ms;
  *out = B.release();
}

}  // namespace

extern "C" {

int pac_graph_create(const float* xyzr, const int64_t* offsets, int64_t n_mol,
                     float solvent, float alpha, const pac_options* opt,
                     double headroom, pac_graph** out) {
  return guarded([&] {
    graph_create(xyzr, offsets, n_mol, solvent, alpha, opt, headroom, out);
  });
}
int pac_graph_set_input(pac_graph* g, const float* xyzr, void* stream) {
  return guarded([&] { graph_set_input(g, xyzr, stream); });
}
int pac_graph_launch(pac_graph* g, void* stream) {
  return guarded([&] { graph_launch(g, stream); });
}
int pac_graph_result(pac_graph* g, pac_batch_result** out) {
  return guarded([&] { graph_result(g, out); });
}
int pac_graph_info(const pac_graph* g, int64_t* nodes, int64_t* bytes) {
  if (!g) return PAC_ERR_INVALID;
  if (nodes) *nodes = g->nodes;
  if (bytes) {
    int64_t b = 0;
    for (auto& kv : const_cast<pac_graph*>(g)->pool.cls_) b += (int64_t)kv.second;
    *bytes = b;
  }
  return PAC_OK;
}
void pac_graph_free(pac_graph* g) { delete g; }

}  // extern "C"
