/*
 * pac_pipeline.cpp -- the pipeline driver: one pass of binning,
 * enumeration + alpha filters, emptiness tests, closure and emission on one
 * stream (replaces the phase sequence of the reference's main(),
 * AlphaComplexParallel.cu:1126-1806).  Compiled once per ball type
 * (pac_variant.h): pac::vf32::run_pipeline, pac::vf64::run_pipeline.
 */
#include "pac_host.h"

namespace pac {
namespace PAC_V {

/* v * 2^shift as a WL-limb two's-complement integer (v*2^shift integral) */
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
static GridPlan plan_grid(const Bounds& B, Real alpha, int policy) {
  GridPlan P;
  GridP& g = P.g;
  std::memset(&g, 0, sizeof(g));
  /* FP32 throughout, as the reference (FP64 balls: on their FP32
   * roundings, the values their cell keys are computed from) */
  const float mr = (float)B.max_raw_r, fal = (float)alpha;
  const float fminx = (float)B.minx, fminy = (float)B.miny, fminz = (float)B.minz;
  volatile float m2 = mr * mr;
  volatile float m2a = m2 + fal;
  volatile float pmr = sqrtf(m2a);
  float step = 2 * pmr;
  if (!(step > 0) || !std::isfinite(step))
    fail(PAC_ERR_INVALID, "grid step 2*sqrt(maxR^2+alpha) is not positive");
  volatile float qx = ((float)B.maxx - fminx) / step,
                 qy = ((float)B.maxy - fminy) / step,
                 qz = ((float)B.maxz - fminz) / step;
  int64_t ex = (int64_t)(int)qx + 1, ey = (int64_t)(int)qy + 1,
          ez = (int64_t)(int)qz + 1;
  P.G = ex * ey * ez;
  if (P.G >= ((int64_t)1 << 31)) fail(PAC_ERR_RANGE, "grid has too many cells");
  g.minx = fminx;
  g.miny = fminy;
  g.minz = fminz;
  g.step = step;
  g.ex = (int)ex;
  g.ey = (int)ey;
  g.ez = (int)ez;
  double maxw = (double)B.max_r * (double)B.max_r;
  double reach2 = maxw + (double)alpha;
  double reach = reach2 > 0 ? std::sqrt(reach2) * (1 + 1e-12) + 1e-12 : 0.0;
  double maxext = std::max({(double)B.maxx - B.minx, (double)B.maxy - B.miny,
                            (double)B.maxz - B.minz});
  /* key rounding: FP32 subtract and divide (and for FP64 balls the
   * rounding of the coordinate itself), in length units */
  double margin = (maxext + (double)step + (kBallF64 ? B.max_abs : 0.0)) *
                      3.0 * 0x1p-24 + 1e-9;
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
  g.inv_step = 1.0 / (double)step;
  g.inv_yz = g.inv_step;
  g.xsplit = 1.0f;
  g.maxw = policy == PAC_CELL_REFERENCE_RAW ? -1.0 : maxw;
  return P;
}

/* Grids of one pass (bounds reduced on the device, read back once), the
 * exact scale and the device-side molecule tables of a batch pass. */
void plan_pass(Engine* E, const Ball* raw, int64_t n, Real solvent,
               Real alpha, const pac_options& opt, hipStream_t s,
               const int64_t* moff, int64_t nmol, PassPlan& pl,
               const Bounds* gridb) {
  Pool* P = pool_of(E);
  PipeLaunch L{s, kBlock};
  const bool batch = nmol > 1;
  std::vector<Bounds> hb(std::max<int64_t>(nmol, 1));
  if (!batch) {
    DBuf<Bounds> bpart(P, 1024 * 2 + 4), bout(P, 1);
    launch_bounds(raw, n, solvent, bpart.get(), bout.get(), L);
    KCHK();
    hb[0] = readback(E, s, bout.get());
  } else {
    std::vector<uint32_t> mo(nmol + 1);
    for (int64_t m = 0; m <= nmol; ++m) mo[m] = (uint32_t)moff[m];
    pl.dmoff = DBuf<uint32_t>(P, nmol + 1);
    HIPCHK(hipMemcpyAsync(pl.dmoff.get(), mo.data(), (nmol + 1) * sizeof(uint32_t),
                          hipMemcpyHostToDevice, s));
    DBuf<Bounds> bout(P, nmol);
    launch_bounds_mol(raw, pl.dmoff.get(), nmol, solvent, bout.get(), L);
    KCHK();
    HIPCHK(hipMemcpyAsync(hb.data(), bout.get(), nmol * sizeof(Bounds),
                          hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  if (gridb && !batch) {
    /* a sub-input of the chunked driver: plan on the grid of the whole
     * input, so that cell keys, stencils and the row order are those of a
     * single pass (the exact scale stays the sub-input's own) */
    Bounds& B = hb[0];
    B.minx = gridb->minx;
    B.miny = gridb->miny;
    B.minz = gridb->minz;
    B.maxx = gridb->maxx;
    B.maxy = gridb->maxy;
    B.maxz = gridb->maxz;
    B.max_raw_r = gridb->max_raw_r;
    B.max_r = gridb->max_r;
    B.max_abs = gridb->max_abs;
  }
  pl.bounds = hb;
  pl.grids.assign(hb.size(), GridP());
  pl.G = 0;
  int shift = 0;
  double mag = 1.0;
  pl.no_edges = true;
  pl.k_enum = pl.k_ac2 = 0;
  for (size_t m = 0; m < hb.size(); ++m) {
    const Bounds& B = hb[m];
    if (B.nonfinite) fail(PAC_ERR_INVALID, "non-finite coordinate or radius");
    bool empty = batch && moff[m + 1] == moff[m];
    GridPlan gp;
    if (empty) { /* one empty cell, never visited */
      std::memset(&gp.g, 0, sizeof(gp.g));
      gp.g.step = 1;
      gp.g.inv_step = 1;
      gp.g.inv_yz = 1;
      gp.g.xsplit = 1.0f;
      gp.g.maxw = -1;
      gp.g.ex = gp.g.ey = gp.g.ez = gp.g.k_enum = gp.g.k_ac2 = 1;
      gp.G = 1;
    } else {
      gp = plan_grid(B, alpha, opt.cell_policy);
      shift = std::max(shift, B.shift);
      mag = std::max({mag, (double)B.max_abs, (double)B.max_r});
      if ((double)B.max_r * (double)B.max_r + (double)alpha >= 0) pl.no_edges = false;
    }
    gp.g.cbase = (uint32_t)pl.G;
    pl.G += gp.G;
    if (pl.G >= ((int64_t)1 << 31)) fail(PAC_ERR_RANGE, "grids have too many cells");
    pl.grids[m] = gp.g;
    pl.k_enum = std::max(pl.k_enum, gp.g.k_enum);
    pl.k_ac2 = std::max(pl.k_ac2, gp.g.k_ac2);
  }
  /* Fine scan grids (exact policy): each reference cell split into kx
   * cells along x (same origin, rows and y, z cells), so that each scan
   * row is cut to the sphere's x-extent finely at no cost in rows. */
  pl.kf = 0;
  pl.Gf = 0;
  pl.sgrids.clear();
  const int kx = opt.ac2_fine > 0 ? std::min(opt.ac2_fine, kMaxFine) : 2;
  if (opt.cell_policy == PAC_CELL_EXACT && opt.ac2_fine >= 0 && kx >= 2 &&
      pl.G * kx < ((int64_t)1 << 31)) {
    pl.kf = kx;
    pl.Gf = pl.G * kx;
    pl.sgrids = pl.grids;
    for (GridP& f : pl.sgrids) {
      f.inv_yz = 1.0 / (double)f.step;
      f.inv_step = kx / (double)f.step;
      f.xsplit = (float)kx;
      f.ex *= kx;
      f.cbase *= (uint32_t)kx;
    }
  }
  /* exact embedding scale (one for all molecules: any common scale at
   * which every input is an integer gives the same exact signs) */
  pl.S = std::max(shift, (neg_lsb_exp_host(alpha) + 1) / 2);
  mag = std::max(mag, std::sqrt(std::fabs((double)alpha)));
  double bits = std::log2(mag) + pl.S + 2;
  if (8 * bits + 16 > 32.0 * WL - 2)
    fail(PAC_ERR_RANGE, "input exponent range exceeds the exact stage");
  pl.alpha_wi = wi_from_scaled_host((double)alpha, 2 * pl.S);
  if (batch) {
    pl.dgrids = DBuf<GridP>(P, nmol);
    HIPCHK(hipMemcpyAsync(pl.dgrids.get(), pl.grids.data(), nmol * sizeof(GridP),
                          hipMemcpyHostToDevice, s));
    pl.dmol = DBuf<uint32_t>(P, n);
    launch_mol_ids(pl.dmoff.get(), nmol, n, pl.dmol.get(), L);
    KCHK();
    if (pl.kf) {
      pl.dsgrids = DBuf<GridP>(P, nmol);
      HIPCHK(hipMemcpyAsync(pl.dsgrids.get(), pl.sgrids.data(),
                            nmol * sizeof(GridP), hipMemcpyHostToDevice, s));
    }
  }
}

/* One pipeline pass over n balls.  nmol == 1: one molecule; otherwise the
 * balls are nmol molecules, moff[0..nmol] (host, relative to raw), each with
 * its own grid, processed together by every kernel launch.
 *
 * rec == nullptr: eager; counts are read back as they are produced and
 * every buffer has its exact size.  rec set: the pass is recorded with the
 * capacities rec->cap for capture into a graph; nothing is read back, the
 * counts stay at rec->dcnt, the plan is rec->plan (checked on the device
 * against the current input) and each precision ladder runs once over
 * queues of fixed size, overflows setting status bits. */
void run_pipeline(Engine* E, const Ball* raw, int64_t n, Real solvent,
                  Real alpha, const pac_options& opt, pac_result* R,
                  hipStream_t s, Timer& T, KProf& KP, const int64_t* moff,
                  int64_t nmol, const PassRec* rec, const SubInput* sub) {
  Pool* P = rec ? rec->pool : pool_of(E);
  PipeLaunch L{s, kBlock};
  pac_stats& st = R->st;
  st.n = n;
  if (n == 0) return;
  if (n > (int64_t)INT32_MAX) fail(PAC_ERR_OVERFLOW, "too many balls");

  /* ---- bounds, grids (AlphaComplexParallel.cu:1089-1124) ---- */
  PassPlan own;
  if (!rec)
    plan_pass(E, raw, n, solvent, alpha, opt, s, moff, nmol, own,
              sub ? &sub->grid : nullptr);
  const PassPlan& pl = rec ? *(const PassPlan*)rec->plan : own;
  /* counts left on the device by the scans (RC_*) */
  DBuf<uint64_t> dcnt_own;
  uint64_t* dcnt = rec ? rec->dcnt : nullptr;
  if (!rec) {
    dcnt_own = DBuf<uint64_t>(P, RC_COUNT);
    dcnt = dcnt_own.get();
  }
  if (rec) {
    zero_u64((unsigned long long*)dcnt, RC_COUNT, s);
    /* does the current input still fit the recorded plan? */
    if (nmol > 1) {
      launch_bounds_mol(raw, pl.dmoff.get(), nmol, solvent, rec->bnow, L);
    } else {
      DBuf<Bounds> bpart(P, 1024 * 2 + 4);
      launch_bounds(raw, n, solvent, bpart.get(), rec->bnow, L);
    }
    launch_plan_check(rec->bnow, rec->bref, (int64_t)pl.bounds.size(), pl.S,
                      rec->status, 1u << REC_STALE, s);
    KCHK();
  }
  const int64_t G = pl.G;
  const int S = pl.S;
  const bool no_edges = pl.no_edges;
  st.k_enum = pl.k_enum;
  st.k_ac2 = pl.k_ac2;
  st.grid_cells = (int32_t)G;
  st.scale_bits = S;
  Stage SG;
  std::memset(&SG, 0, sizeof(SG));
  SG.n = n;
  SG.g = pl.grids[0];
  if (nmol > 1) {
    SG.grids = pl.dgrids.get();
    SG.mol = pl.dmol.get();
  }
  SG.alpha = alpha;
  SG.ex.S = S;
  SG.ex.alpha = pl.alpha_wi;
  /* FP64 balls are not exact in FP32: their ladder starts at FP64 */
  SG.staged = opt.precision == PAC_PREC_STAGED && !kBallF64;
  SG.bgrid = opt.ac2_grid > 0 ? std::min(opt.ac2_grid, 1 << 16) : 2048;
  /* persistent second emptiness phase: block shape and hand-out size */
  SG.bblock = opt.block == 64 || opt.block == 128 ? opt.block : kBlock;
  SG.bchunk = opt.ac2_chunk > 0
                  ? (unsigned)std::min((opt.ac2_chunk + 63) / 64 * 64, 1 << 16)
                  : 64u;
  if (sub) {
    SG.chunk = 1;
    SG.own_lo = sub->own_lo;
    SG.own_hi = sub->own_hi;
  }
  DBuf<unsigned long long> ctr_own;
  if (rec) {
    SG.ctr = rec->ctr;
  } else {
    ctr_own = DBuf<unsigned long long>(P, CTR_COUNT * CTR_SLOTS);
    zero_u64(ctr_own.get(), CTR_COUNT * CTR_SLOTS, s);
    SG.ctr = ctr_own.get();
  }
  unsigned long long* const ctr = SG.ctr;
  /* [0, QSLOTS) hot queue, [QSLOTS, 2 QSLOTS) escalation queue, then the
   * LSLOTS counters of the list of simplices left for the second emptiness
   * phase and those of the list of tets a quick rejection pass left */
  DBuf<unsigned> qn_own;
  if (!rec) {
    qn_own = DBuf<unsigned>(P, QN_WORDS);
    zero_u32(qn_own.get(), QN_WORDS, s); /* then kept clear by the views */
  }
  unsigned* const qn = rec ? rec->qn : qn_own.get();
  unsigned* const nrest = qn + 2 * QSLOTS;
  unsigned* const nnext = nrest + LSLOTS;
  /* eager ladders: the device view of q1 and the state read back once */
  DBuf<QView> ev_view(P, 1);
  DBuf<unsigned> lad(P, kLadWords);

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
  DBuf<Ball> ball(P, n);
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

  /* ---- fine scan grid: reference cells split along x ---- */
  DBuf<Ball> fball;
  DBuf<uint32_t> fid, fstart;
  if (pl.kf) {
    const GridP* dsg = nmol > 1 ? pl.dsgrids.get() : nullptr;
    fball = DBuf<Ball>(P, n);
    fid = DBuf<uint32_t>(P, n);
    fstart = DBuf<uint32_t>(P, pl.Gf + 1);
    launch_fine_split(SG.ball, SG.cstart, G, pl.kf, pl.sgrids[0], dsg, SG.mol,
                      fball.get(), fid.get(), fstart.get(), L);
    KCHK();
    SG.fine = 1;
    SG.quick = opt.ac2_quick > 0;
    SG.sg = pl.sgrids[0];
    SG.sgrids = dsg;
    SG.fball = fball.get();
    SG.fid = fid.get();
    SG.fstart = fstart.get();
  }
  T.mark(2);

  auto scan_tmp = [&](int64_t m) { return DBuf<uint8_t>(P, scan_tmp_bytes(m) + 64); };
  const bool staged = SG.staged != 0;

  /* the count a scan left at dcnt[q]: read back, or (recorded) kept on the
   * device under its capacity, an excess setting status bit q */
  auto settle = [&](int q) -> Cnt {
    if (!rec) return Cnt((int64_t)readback(E, s, dcnt + q));
    launch_cap_check(dcnt + q, (uint64_t)rec->cap[q], rec->status,
                     1u << (REC_CAP + q), s);
    return Cnt(rec->cap[q], dcnt + q);
  };

  /* Precision ladder: hot kernel (q1) -> [staged: escalation kernel FP64 ->
   * double-double (q1 -> q2)] -> exact wide-integer kernel, then `tail`,
   * the scan that leaves the ladder's count at dcnt[q].
   * Eager: the escalation reads q1 through a view built on the device, and
   * the tail runs before the exact stage on the chance that the last queue
   * (q2; q1 without escalation) is empty, which it almost always is; one
   * read-back then brings that queue's counts, the count and the overflow
   * status.  Only when the last queue holds entries are they read into a
   * host view (for the exact stage, sorted when long), and the tail run
   * again after it.  An overflow reruns the ladder with larger queues (the
   * stages are idempotent).
   * Recorded: views built on the device, queues of fixed size, overflows
   * reported through the status word. */
  const int64_t qmax = (int64_t)(UINT32_MAX / QSLOTS) - 1;
  /* sub-queue capacity for up to `total` entries spread over the slots;
   * a one-item-per-thread producer of n items never exceeds n/QSLOTS + 512
   * per slot (its blocks of 256 are dealt round-robin) */
  auto slot_cap = [&](int64_t total) {
    return (uint32_t)std::min<int64_t>(total / QSLOTS + 512, qmax);
  };
  auto grow = [&](uint32_t need) {
    return (uint32_t)std::min<int64_t>((int64_t)need + need / 4 + 256, qmax);
  };
  auto ladder = [&](int64_t cap1_total, int kp_hot, int kp_exact, auto&& hot,
                    auto&& esc, auto&& exact, auto&& tail, int q) -> Cnt {
    if (rec) { /* once, over the recorded queues; views built on the device */
      hot(Queue{rec->q1, qn, rec->c1});
      launch_make_view(rec->q1, rec->c1, qn, rec->views, rec->status,
                       1u << REC_Q1, rec->fill, nrest, false, s);
      QRef r{QView{}, rec->views, QSLOTS * rec->c1};
      if (staged) {
        esc(r, Queue{rec->q2, qn + QSLOTS, rec->c2});
        launch_make_view(rec->q2, rec->c2, qn + QSLOTS, rec->views + 1,
                         rec->status, 1u << REC_Q2, rec->fill + 1, nullptr,
                         false, s);
        r = QRef{QView{}, rec->views + 1, QSLOTS * rec->c2};
      }
      exact(r);
      KCHK();
      tail();
      KCHK();
      return settle(q);
    }
    uint32_t c1 = slot_cap(cap1_total);
    /* q2 as large as q1: an escalated pair queues at most itself; a
     * whole-simplex entry queues its undecided pairs, and an overflow
     * reruns the ladder */
    uint32_t c2 = staged ? slot_cap((int64_t)c1 * QSLOTS + 4096) : 0;
    DBuf<QEntry> q1, q2;
    for (;;) {
      q1 = DBuf<QEntry>(P, (size_t)c1 * QSLOTS);
      zero_u32(nrest, LSLOTS, s);
      if (kp_hot >= 0) KP.begin(kp_hot);
      hot(Queue{q1.get(), qn, c1});
      if (kp_hot >= 0) KP.end();
      KCHK();
      const QEntry* last = q1.get();
      uint32_t clast = c1;
      unsigned* cnt_last = qn;
      if (staged) {
        q2 = DBuf<QEntry>(P, (size_t)c2 * QSLOTS);
        launch_make_view(q1.get(), c1, qn, ev_view.get(), lad.get() + kLadStatus,
                         1u, lad.get() + kLadFill, nullptr, true, s);
        esc(QRef{QView{}, ev_view.get(), QSLOTS * c1},
            Queue{q2.get(), qn + QSLOTS, c2});
        KCHK();
        last = q2.get();
        clast = c2;
        cnt_last = qn + QSLOTS;
      }
      tail();
      launch_ladder_state(cnt_last, dcnt + q, lad.get(), s);
      KCHK();
      unsigned h[kLadWords];
      HIPCHK(hipMemcpyAsync(E->pinned, lad.get(), sizeof(h),
                            hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      std::memcpy(h, E->pinned, sizeof(h));
      QView vx;
      vx.q = last;
      vx.cap = clast;
      uint64_t acc = 0;
      unsigned mx = 0;
      for (int k = 0; k < QSLOTS; ++k) {
        vx.off[k] = (uint32_t)acc;
        acc += std::min(h[kLadCnt + k], clast);
        mx = std::max(mx, h[kLadCnt + k]);
      }
      vx.off[QSLOTS] = (uint32_t)acc;
      const unsigned fill1 = staged ? h[kLadFill] : mx;
      R->sizes[9] = std::max<int64_t>(R->sizes[9], fill1);
      if (staged) R->sizes[10] = std::max<int64_t>(R->sizes[10], mx);
      if (fill1 > c1) { /* q1 overflowed: everything again, larger */
        c1 = grow(fill1);
        if (staged) c2 = std::max(c2, slot_cap((int64_t)c1 * QSLOTS + 4096));
        continue;
      }
      if (mx > clast) { /* q2 overflowed */
        c2 = grow(mx);
        continue;
      }
      if (!vx.size()) {
        uint64_t cnt;
        std::memcpy(&cnt, h + kLadCount, sizeof(cnt));
        return Cnt((int64_t)cnt);
      }
      if (kp_exact >= 0) KP.begin(kp_exact);
      exact(QRef{vx, nullptr, 0});
      if (kp_exact >= 0) KP.end();
      KCHK();
      tail();
      KCHK();
      return settle(q);
    }
  };

  /* hot-stage evaluations booked on the host: one per edge candidate
   * (triangle and tet candidates are counted by their kernels); the
   * filters' first stages run fused with the enumerations, each a ladder
   * with the escalation and exact kernels of its kind */
  int64_t host_evals[6] = {0, 0, 0, 0, 0, 0};
  auto count_scan = [&](const uint32_t* cnt, Cnt nseg, uint64_t* off, int q,
                        const char* what) -> Cnt {
    {
      auto tmp = scan_tmp(nseg.cap);
      scan_u32_to_u64(cnt, off, nseg, tmp.get(), 0, s, dcnt + q);
    }
    KCHK();
    Cnt tot = settle(q);
    if (tot.cap >= ((int64_t)1 << 32))
      fail(PAC_ERR_OVERFLOW, std::string(what) + " candidates exceed 2^32");
    return tot;
  };
  /* Every buffer below is sized by a count's cap: the exact count (eager)
   * or its recorded capacity.  Enumerations stop at their arrays' ends,
   * compactions at their outputs' ends, and CSR row starts are clamped to
   * the row arrays, so a recorded pass that overflows stays in bounds. */

  /* ---- edges (AlphaComplexParallel.cu:1200-1263) ---- */
  Cnt Ecand(0), Ealpha(0);
  DBuf<uint64_t> erow(P, n + 1);
  DBuf<uint32_t> ecol, eleft;
  CandCtx C0{nullptr, nullptr, nullptr, nullptr, nullptr};
  if (!no_edges) {
    DBuf<uint64_t> off(P, n + 1);
    {
      DBuf<uint32_t> cnt(P, n);
      launch_edge_count(SG, cnt.get(), L);
      Ecand = count_scan(cnt.get(), n, off.get(), RC_ECAND, "edge");
    }
    DBuf<uint32_t> cval(P, Ecand.cap), cseg(P, Ecand.cap);
    DBuf<uint8_t> flag(P, Ecand.cap);
    /* enumeration and first filter stage in one kernel (each ball's kept
     * and undecided candidates compacted from its first slot, their count
     * in bcnt); the ladder settles the undecided ones; the kept counts'
     * scan is the alpha-edge CSR (erow) and the kept edges are copied out
     * in order */
    DBuf<uint32_t> bcnt(P, n), keep(P, n);
    auto tmp = scan_tmp(n);
    host_evals[0] += Ecand.cap;
    Ealpha = ladder(
        std::min<int64_t>(Ecand.cap, Ecand.cap / 32 + 65536), KP_EDGE_FILL, -1,
        [&](Queue qq) {
          launch_edge_enum_filter(SG, off.get(), cval.get(), cseg.get(),
                                  flag.get(), bcnt.get(), qq, L,
                                  (uint64_t)Ecand.cap);
        },
        [&](QRef qr, Queue q2) {
          launch_filter_esc(SG, 1, C0, qr, cval.get(), flag.get(), q2, L);
        },
        [&](QRef qr) {
          launch_filter_exact(SG, 1, C0, qr, cval.get(), flag.get(), L);
        },
        [&] {
          launch_seg_keep(n, off.get(), bcnt.get(), flag.get(), keep.get(), L);
          scan_u32_to_u64(keep.get(), erow.get(), Cnt(n), tmp.get(), 0, s,
                          dcnt + RC_EALPHA);
        },
        RC_EALPHA);
    ecol = DBuf<uint32_t>(P, Ealpha.cap);
    eleft = DBuf<uint32_t>(P, Ealpha.cap);
    launch_edge_copy(n, off.get(), bcnt.get(), cval.get(), flag.get(),
                     erow.get(), ecol.get(), eleft.get(), L,
                     (uint64_t)Ealpha.cap);
    KCHK();
  } else {
    zero_u64((unsigned long long*)erow.get(), n + 1, s);
    ecol = DBuf<uint32_t>(P, 1);
    eleft = DBuf<uint32_t>(P, 1);
  }
  T.mark(3);

  /* ---- triangles (AlphaComplexParallel.cu:1279-1371) ---- */
  Cnt Tslots(0), Talpha(0);
  DBuf<uint64_t> trow(P, Ealpha.cap + 1);
  /* the tet slots of each alpha edge's triangles (their row tails), for
   * the tets' slot offsets in closed form; chunk sub-inputs give unowned
   * triangles no slots and keep per-triangle offsets */
  const bool tail_slots = !SG.chunk;
  DBuf<uint64_t> tpair(P, tail_slots ? std::max<int64_t>(Ealpha.cap, 1) : 1);
  DBuf<uint32_t> tk, tedge, teik, tejk;
  CandCtx C1{eleft.get(), ecol.get(), nullptr, nullptr, nullptr};
  {
    DBuf<uint64_t> off(P, Ealpha.cap + 1);
    {
      /* exact counts: a triangle's third vertex must be in both rows, and
       * the merge that counts it is cheap next to filtering empty slots */
      DBuf<uint32_t> cnt(P, Ealpha.cap);
      launch_tri_count(Ealpha, erow.get(), ecol.get(), eleft.get(), cnt.get(), L);
      Tslots = count_scan(cnt.get(), Ealpha, off.get(), RC_TSLOTS, "triangle");
    }
    DBuf<uint32_t> cval(P, Tslots.cap), cseg(P, Tslots.cap),
        cea(P, Tslots.cap), ceb(P, Tslots.cap);
    DBuf<uint8_t> flag(P, Tslots.cap);
    /* enumeration and first filter stage in one kernel, each block's kept
     * and undecided candidates compacted in place (their count in bcnt);
     * the ladder settles the undecided ones; then the kept triangles of
     * each block are counted, scanned and copied out in order with the
     * row starts (as for tets) */
    const int64_t nblk = (Ealpha.cap + kBlock - 1) / kBlock;
    DBuf<uint32_t> bcnt(P, nblk + 1), bkeep(P, nblk + 1);
    DBuf<uint64_t> bout(P, nblk + 1);
    auto tmp = scan_tmp(nblk);
    /* a rerun (queue overflow, eager only) recounts the candidates from
     * zero */
    int runs = 0;
    Talpha = ladder(
        std::min<int64_t>(Tslots.cap, Tslots.cap / 32 + 65536), KP_TRI_FILL, -1,
        [&](Queue qq) {
          if (runs++ && !rec)
            zero_u64(ctr + (CTR_CAND + 1) * CTR_SLOTS, CTR_SLOTS, s);
          launch_tri_enum_filter(SG, Ealpha, erow.get(), ecol.get(), eleft.get(),
                                 off.get(), cval.get(), cseg.get(), cea.get(),
                                 ceb.get(), flag.get(), bcnt.get(), bkeep.get(),
                                 qq, L, (uint64_t)Tslots.cap);
        },
        [&](QRef qr, Queue q2) {
          launch_filter_esc(SG, 2, C1, qr, cval.get(), flag.get(), q2, L);
        },
        [&](QRef qr) {
          launch_filter_exact(SG, 2, C1, qr, cval.get(), flag.get(), L);
        },
        [&] {
          launch_tet_block_keep(Ealpha, off.get(), kBlock, bcnt.get(), flag.get(),
                                bkeep.get(), L);
          scan_u32_to_u64(bkeep.get(), bout.get(), Cnt(nblk), tmp.get(), 0, s,
                          dcnt + RC_TALPHA);
        },
        RC_TALPHA);
    tk = DBuf<uint32_t>(P, Talpha.cap);
    tedge = DBuf<uint32_t>(P, Talpha.cap);
    teik = DBuf<uint32_t>(P, Talpha.cap);
    tejk = DBuf<uint32_t>(P, Talpha.cap);
    launch_tri_block_copy(Ealpha, off.get(), bcnt.get(), cval.get(), cseg.get(),
                          cea.get(), ceb.get(), flag.get(), bout.get(), tk.get(),
                          tedge.get(), teik.get(), tejk.get(), trow.get(),
                          tail_slots ? tpair.get() : nullptr, L,
                          (uint64_t)Talpha.cap);
    KCHK();
  }
  T.mark(4);

  /* ---- tets (AlphaComplexParallel.cu:1381-1442) ---- */
  Cnt Kslots(0), Kalpha(0);
  DBuf<uint4> tets;
  CandCtx C2{eleft.get(), ecol.get(), tk.get(), tedge.get(), nullptr};
  {
    /* slot offsets: per alpha edge in closed form (roff), or per alpha
     * triangle from its cap (off, chunk sub-inputs) */
    DBuf<uint64_t> off, roff;
    if (tail_slots) {
      roff = DBuf<uint64_t>(P, Ealpha.cap + 1);
      {
        auto tmp = scan_tmp(Ealpha.cap);
        scan_u64_to_u64(tpair.get(), roff.get(), Ealpha, tmp.get(), 0, s,
                        dcnt + RC_KSLOTS);
      }
      KCHK();
      Kslots = settle(RC_KSLOTS);
      if (Kslots.cap >= ((int64_t)1 << 32))
        fail(PAC_ERR_OVERFLOW, "tet candidates exceed 2^32");
    } else {
      off = DBuf<uint64_t>(P, Talpha.cap + 1);
      DBuf<uint32_t> cap(P, Talpha.cap);
      launch_tet_caps(SG, Talpha, trow.get(), tedge.get(), eleft.get(),
                      ecol.get(), tk.get(), cap.get(), L);
      Kslots = count_scan(cap.get(), Talpha, off.get(), RC_KSLOTS, "tet");
    }
    tpair.release();
    DBuf<uint4> ctet(P, Kslots.cap);
    DBuf<uint8_t> flag(P, Kslots.cap);
    C2.ctet = ctet.get();
    /* enumeration and first filter stage in one kernel, each block's kept
     * and undecided tets compacted in place (its count in bcnt); the
     * ladder settles the undecided ones; then the kept tets of each block
     * are counted, scanned and copied out in order */
    const int64_t nblk = (Talpha.cap + kBlock - 1) / kBlock;
    DBuf<uint32_t> bcnt(P, nblk + 1), bkeep(P, nblk + 1);
    DBuf<uint64_t> bout(P, nblk + 1), bbase(P, nblk + 1);
    auto tmp = scan_tmp(nblk);
    /* a rerun (queue overflow, eager only) recounts the candidates from
     * zero */
    int runs = 0;
    Kalpha = ladder(
        std::min<int64_t>(Kslots.cap, Kslots.cap / 32 + 65536), KP_TET_FILL, -1,
        [&](Queue qq) {
          if (runs++ && !rec)
            zero_u64(ctr + (CTR_CAND + 2) * CTR_SLOTS, CTR_SLOTS, s);
          launch_tet_enum_filter(SG, Talpha, trow.get(), tk.get(), tedge.get(),
                                 teik.get(), tejk.get(), eleft.get(), ecol.get(),
                                 tail_slots ? nullptr : off.get(),
                                 tail_slots ? roff.get() : nullptr, bbase.get(),
                                 ctet.get(), flag.get(), bcnt.get(), bkeep.get(),
                                 qq, L, (uint64_t)Kslots.cap);
        },
        [&](QRef qr, Queue q2) {
          launch_filter_esc(SG, 3, C2, qr, nullptr, flag.get(), q2, L);
        },
        [&](QRef qr) { launch_filter_exact(SG, 3, C2, qr, nullptr, flag.get(), L); },
        [&] {
          launch_tet_block_keep(Talpha, bbase.get(), 1, bcnt.get(), flag.get(),
                                bkeep.get(), L);
          scan_u32_to_u64(bkeep.get(), bout.get(), Cnt(nblk), tmp.get(), 0, s,
                          dcnt + RC_KALPHA);
        },
        RC_KALPHA);
    tets = DBuf<uint4>(P, Kalpha.cap);
    launch_tet_block_copy(Talpha, bbase.get(), 1, bcnt.get(), ctet.get(),
                          flag.get(), bout.get(), tets.get(), L,
                          (uint64_t)Kalpha.cap);
    KCHK();
  }
  T.mark(5);

  auto ac2_cap = [](int64_t items) {
    return (uint32_t)std::min<int64_t>(std::max<int64_t>(1 << 20, items / 8),
                                       (int64_t)UINT32_MAX - 1);
  };

  /* A long host view of exact pairs (degenerate inputs) as (item, aux)
   * arrays sorted by item, so that an item's pairs share its exact
   * per-simplex terms; false (nothing done) below 65536 entries or for a
   * device-built view.  items < nitems. */
  auto sorted_view = [&](const QRef& q, int64_t nitems, DBuf<uint32_t>& k1,
                         DBuf<uint32_t>& v1, DBuf<uint32_t>& k2,
                         DBuf<uint32_t>& v2, bool& alt) -> uint32_t {
    const uint32_t ne = q.dv ? 0u : q.v.size();
    if (ne < (1u << 16)) return 0;
    k1 = DBuf<uint32_t>(P, ne);
    v1 = DBuf<uint32_t>(P, ne);
    k2 = DBuf<uint32_t>(P, ne);
    v2 = DBuf<uint32_t>(P, ne);
    launch_view_split(q.v, k1.get(), v1.get(), L);
    int bits = 1;
    while (bits < 32 && ((int64_t)1 << bits) < nitems) ++bits;
    const size_t rtb = radix_tmp_bytes(ne);
    DBuf<uint8_t> rt(P, rtb);
    alt = false;
    radix_sort_pairs(k1.get(), v1.get(), k2.get(), v2.get(), ne, bits, rt.get(),
                     rtb, s, &alt);
    return ne;
  };

  /* ---- tet emptiness ---- */
  /* phase-one budget: 0 = the default (4), < 0 = no second phase */
  int budget = opt.ac2_budget > 0 ? opt.ac2_budget : (opt.ac2_budget < 0 ? 0 : 4);
  /* the left-over and quick-pass lists: LSLOTS sub-lists sized for the
   * largest first-phase grid */
  SG.lcap = list_slot_cap(
      (std::max({Kalpha.cap, Talpha.cap, Ealpha.cap, (int64_t)1}) + kBlock - 1) /
      kBlock);
  DBuf<uint32_t> rest(P, (size_t)LSLOTS * SG.lcap);
  DBuf<uint32_t> next(P, pl.kf || sub ? (size_t)LSLOTS * SG.lcap : 1);
  DBuf<uint8_t> kflag(P, Kalpha.cap);
  Cnt Kf(0);
  DBuf<uint4> ftets;
  {
    auto tmp = scan_tmp(Kalpha.cap);
    const uint64_t* koffs = nullptr; /* tile offsets of kflag, in tmp */
    auto tail = [&] {
      koffs = flag_tile_offsets(kflag.get(), Kalpha, tmp.get(), s, dcnt + RC_KF);
    };
    if (opt.skip_tet_ac2 || !Kalpha.cap) {
      fill_u8(kflag.get(), Kalpha.cap, 1, s);
      tail();
      KCHK();
      Kf = settle(RC_KF);
    } else {
      Kf = ladder(
          ac2_cap(Kalpha.cap), KP_TET_AC2, KP_TET_AC2_EXACT,
          [&](Queue q) {
            launch_tet_ac2_hot(SG, Kalpha, tets.get(), kflag.get(), q,
                               rest.get(), nrest, next.get(), nnext,
                               budget, L);
          },
          [&](QRef q, Queue q2) {
            launch_tet_ac2_esc(SG, q, tets.get(), kflag.get(), q2, L);
          },
          [&](QRef q) {
            /* many exact pairs (degenerate inputs): sorted by tet first, so
             * that the pairs of a tet share its exact D and N */
            DBuf<uint32_t> k1, v1, k2, v2;
            bool alt = false;
            const uint32_t ne = sorted_view(q, Kalpha.cap, k1, v1, k2, v2, alt);
            if (!ne) {
              launch_tet_ac2_exact(SG, q, tets.get(), perm.get(), kflag.get(), L);
              return;
            }
            launch_tet_ac2_exact_sorted(SG, alt ? k2.get() : k1.get(),
                                        alt ? v2.get() : v1.get(), ne,
                                        tets.get(), perm.get(), kflag.get(), L);
          },
          tail, RC_KF);
    }
    ftets = DBuf<uint4>(P, Kf.cap);
    launch_compact_tets(Kalpha, tets.get(), kflag.get(), koffs, ftets.get(), L,
                        (uint64_t)Kf.cap);
    KCHK();
  }
  tets.release();
  T.mark(6);

  /* ---- triangle closure + dangling emptiness ---- */
  /* filtration: every final triangle and edge gets its emptiness test (an
   * empty smallest orthosphere: the simplex enters at its own rho), and
   * the closure marks join the final flags only after the test */
  const bool filt = opt.filtration != 0 && !rec;
  DBuf<uint8_t> tri_in(P, Talpha.cap), tflag(P, Talpha.cap);
  fill_u8(tri_in.get(), Talpha.cap, 0, s);
  launch_mark_faces(Kf, ftets.get(), erow.get(), ecol.get(), trow.get(),
                    tk.get(), tri_in.get(), L);
  DBuf<uint8_t> tri_skip_own, tfin;
  uint8_t* tri_skip = tri_in.get(); /* passes without a test */
  if (filt) {
    tri_skip_own = DBuf<uint8_t>(P, Talpha.cap);
    fill_u8(tri_skip_own.get(), Talpha.cap, 0, s);
    tri_skip = tri_skip_own.get();
    tfin = DBuf<uint8_t>(P, Talpha.cap);
  }
  launch_own_mark(SG, 0, Talpha, ecol.get(), eleft.get(), tk.get(), tedge.get(),
                  tri_skip, L);
  const uint8_t* tfinal = filt ? tfin.get() : tflag.get();
  Cnt Tf(0);
  DBuf<uint32_t> ftri;
  {
    auto tmp = scan_tmp(Talpha.cap);
    const uint64_t* toffs = nullptr; /* tile offsets of tfinal, in tmp */
    auto tail = [&] {
      if (filt) launch_or_flags(Talpha, tflag.get(), tri_in.get(), tfin.get(), L);
      toffs = flag_tile_offsets(tfinal, Talpha, tmp.get(), s, dcnt + RC_TF);
    };
    if (!Talpha.cap) {
      tail();
      KCHK();
      Tf = settle(RC_TF);
    } else {
      Tf = ladder(
          ac2_cap(Talpha.cap), KP_TRI_AC2, -1,
          [&](Queue q) {
            launch_low_ac2_hot(SG, 0, Talpha, ecol.get(), eleft.get(), tk.get(),
                               tedge.get(), tri_skip, tflag.get(), q,
                               rest.get(), nrest, budget, L);
          },
          [&](QRef q, Queue q2) {
            launch_low_ac2_esc(SG, 0, q, ecol.get(), eleft.get(), tk.get(),
                               tedge.get(), tflag.get(), q2, L);
          },
          [&](QRef q) {
            DBuf<uint32_t> k1, v1, k2, v2;
            bool alt = false;
            const uint32_t ne = sorted_view(q, Talpha.cap, k1, v1, k2, v2, alt);
            if (!ne)
              launch_low_ac2_exact(SG, 0, q, ecol.get(), eleft.get(), tk.get(),
                                   tedge.get(), tflag.get(), L);
            else
              launch_low_ac2_exact_sorted(
                  SG, 0, alt ? k2.get() : k1.get(), alt ? v2.get() : v1.get(),
                  ne, ecol.get(), eleft.get(), tk.get(), tedge.get(),
                  tflag.get(), L);
          },
          tail, RC_TF);
    }
    ftri = DBuf<uint32_t>(P, Tf.cap);
    launch_compact_ids(Talpha, tfinal, toffs, ftri.get(), L, (uint64_t)Tf.cap);
    if (filt) {
      R->gab[1] = DBuf<uint8_t>(P, Tf.cap);
      launch_compact_u8(Talpha, tfinal, toffs, tflag.get(), R->gab[1].get(), L,
                        (uint64_t)Tf.cap);
    }
    KCHK();
  }
  T.mark(7);

  /* ---- edge closure + dangling emptiness ---- */
  DBuf<uint8_t> edge_in(P, Ealpha.cap), eflag(P, Ealpha.cap);
  fill_u8(edge_in.get(), Ealpha.cap, 0, s);
  launch_mark_edges(Tf, ftri.get(), tedge.get(), teik.get(), tejk.get(),
                    edge_in.get(), L);
  DBuf<uint8_t> edge_skip_own, efin;
  uint8_t* edge_skip = edge_in.get();
  if (filt) {
    edge_skip_own = DBuf<uint8_t>(P, Ealpha.cap);
    fill_u8(edge_skip_own.get(), Ealpha.cap, 0, s);
    edge_skip = edge_skip_own.get();
    efin = DBuf<uint8_t>(P, Ealpha.cap);
  }
  launch_own_mark(SG, 1, Ealpha, ecol.get(), eleft.get(), nullptr, nullptr,
                  edge_skip, L);
  const uint8_t* efinal = filt ? efin.get() : eflag.get();
  Cnt Ef(0);
  auto etmp = scan_tmp(Ealpha.cap);
  const uint64_t* eoffs = nullptr; /* tile offsets of efinal, in etmp */
  {
    auto tail = [&] {
      if (filt) launch_or_flags(Ealpha, eflag.get(), edge_in.get(), efin.get(), L);
      eoffs = flag_tile_offsets(efinal, Ealpha, etmp.get(), s, dcnt + RC_EF);
    };
    if (!Ealpha.cap) {
      tail();
      KCHK();
      Ef = settle(RC_EF);
    } else {
      Ef = ladder(
          ac2_cap(Ealpha.cap), KP_EDGE_AC2, -1,
          [&](Queue q) {
            launch_low_ac2_hot(SG, 1, Ealpha, ecol.get(), eleft.get(), nullptr,
                               nullptr, edge_skip, eflag.get(), q,
                               rest.get(), nrest, budget, L);
          },
          [&](QRef q, Queue q2) {
            launch_low_ac2_esc(SG, 1, q, ecol.get(), eleft.get(), nullptr,
                               nullptr, eflag.get(), q2, L);
          },
          [&](QRef q) {
            DBuf<uint32_t> k1, v1, k2, v2;
            bool alt = false;
            const uint32_t ne = sorted_view(q, Ealpha.cap, k1, v1, k2, v2, alt);
            if (!ne)
              launch_low_ac2_exact(SG, 1, q, ecol.get(), eleft.get(), nullptr,
                                   nullptr, eflag.get(), L);
            else
              launch_low_ac2_exact_sorted(
                  SG, 1, alt ? k2.get() : k1.get(), alt ? v2.get() : v1.get(),
                  ne, ecol.get(), eleft.get(), nullptr, nullptr, eflag.get(),
                  L);
          },
          tail, RC_EF);
    }
  }

  /* ---- emit final lists (original ids, reference order) ---- */
  R->lists[0] = DBuf<int32_t>(P, 2 * Ef.cap);
  R->lists[1] = DBuf<int32_t>(P, 3 * Tf.cap);
  R->lists[2] = DBuf<int32_t>(P, 4 * Kf.cap);
  launch_emit_edges(Ealpha, ecol.get(), eleft.get(), efinal, eoffs, perm.get(),
                    R->lists[0].get(), L, (uint64_t)Ef.cap);
  if (filt) {
    R->gab[0] = DBuf<uint8_t>(P, Ef.cap);
    launch_compact_u8(Ealpha, efinal, eoffs, eflag.get(), R->gab[0].get(), L,
                      (uint64_t)Ef.cap);
    R->rank = DBuf<uint32_t>(P, n);
    launch_rank_of(perm.get(), n, R->rank.get(), s);
  }
  launch_emit_tris(Tf, ftri.get(), ecol.get(), eleft.get(), tk.get(),
                   tedge.get(), perm.get(), R->lists[1].get(), L);
  launch_emit_tets(Kf, ftets.get(), perm.get(), R->lists[2].get(), L);
  KCHK();
  if (rec) return; /* counts and statistics are read after the graph runs */
  R->counts[1] = Ef.cap;
  R->counts[2] = Tf.cap;
  R->counts[3] = Kf.cap;
  {
    const Cnt* sz[9] = {&Ecand, &Ealpha, &Tslots, &Talpha, &Kslots,
                        &Kalpha, &Kf, &Tf, &Ef};
    for (int q = 0; q < 9; ++q) R->sizes[q] = sz[q]->cap;
    R->sizes[11] = n;
  }
  st.edge_candidates = Ecand.cap;
  st.edge_alpha = Ealpha.cap;
  st.tri_alpha = Talpha.cap;
  st.tet_alpha = Kalpha.cap;
  st.edges_final = Ef.cap;
  st.tris_final = Tf.cap;
  st.tets_final = Kf.cap;
  T.mark(8);

  unsigned long long hs[CTR_COUNT * CTR_SLOTS];
  HIPCHK(hipMemcpyAsync(hs, ctr, sizeof(hs), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  stats_from_counters(hs, SG.staged != 0, host_evals[0], st);
}

/* The LD driver (AlphaComplexLargeData.cu:1227-1992).  The cell-sorted
 * balls are walked in chunks of C = 512 * 2^k >= opt.chunk_balls sorted
 * ids (:1228-1233); a chunk owns the simplices whose lowest sorted vertex
 * lies in it (:1232, :775-778).
 *
 * Here each chunk is one ordinary pipeline run over a sub-input: the balls
 * of the chunk's z-layers of reference cells plus H layers on either side,
 * in input order, planned on the grid of the whole input.  Its cell-sorted
 * order is then the whole input's restricted, so the rows it owns are one
 * contiguous run of each output list, already in reference order; they are
 * mapped back to input ids and concatenated chunk after chunk.  Every
 * decision about a simplex with lowest vertex v -- its alpha filter, its
 * emptiness test, and the cofaces that close it -- involves only balls
 * within D = 2 sqrt(maxR^2 + alpha) of v (alpha edges are at most D long;
 * orthocentres lie within D/2 of every vertex; a ball of power <= rho lies
 * within D/2 of the centre), and with the reference's own 27-cell stencils
 * only balls within D/2 plus three cells; H covers both with a margin for
 * the FP32 cell key, so the lists equal those of one pass.  (LD keeps every
 * ball resident and lets out-of-chunk vertices through unchecked,
 * :210-216, :383-399; this driver instead bounds each run's memory by its
 * sub-input.)  A sub-input skips the emptiness tests no owned simplex
 * depends on (Stage::chunk); candidate and escalation counts include the
 * halo's repeated enumeration. */
void run_chunked(Engine* E, const Ball* raw, int64_t n, Real solvent,
                 Real alpha, const pac_options& opt, pac_result* R,
                 hipStream_t s, Timer& T, KProf& KP) {
  int64_t C = 512;
  while (C < (int64_t)opt.chunk_balls) C *= 2;
  if (n <= C) {
    run_pipeline(E, raw, n, solvent, alpha, opt, R, s, T, KP);
    return;
  }
  if (n > (int64_t)INT32_MAX) fail(PAC_ERR_OVERFLOW, "too many balls");
  Pool* P = pool_of(E);
  PipeLaunch L{s, kBlock};

  /* the whole input's grid and cell-sorted order */
  PassPlan gp;
  plan_pass(E, raw, n, solvent, alpha, opt, s, nullptr, 1, gp);
  const GridP g = gp.grids[0];
  const Bounds gb = gp.bounds[0];
  DBuf<uint32_t> rank(P, n);
  std::vector<uint32_t> hkey(n);
  {
    DBuf<uint32_t> key(P, n), idx(P, n), key2(P, n), idx2(P, n);
    launch_cell_key(raw, n, g, nullptr, nullptr, key.get(), idx.get(), L);
    KCHK();
    int kbits = 0;
    while (kbits < 32 && ((int64_t)1 << kbits) < gp.G) ++kbits;
    size_t rtb = radix_tmp_bytes(n);
    DBuf<uint8_t> rtmp(P, rtb);
    bool alt = false;
    radix_sort_pairs(key.get(), idx.get(), key2.get(), idx2.get(), n, kbits,
                     rtmp.get(), rtb, s, &alt);
    KCHK();
    launch_rank_of(alt ? idx2.get() : idx.get(), n, rank.get(), s);
    KCHK();
    HIPCHK(hipMemcpyAsync(hkey.data(), alt ? key2.get() : key.get(),
                          (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                          s));
    HIPCHK(hipStreamSynchronize(s));
  }
  for (int q = 2; q <= 8; ++q) T.mark(q);

  const uint64_t exy = (uint64_t)g.ex * (uint64_t)g.ey;
  const double reach2 = (double)gb.max_r * (double)gb.max_r + (double)alpha;
  const double D = reach2 > 0 ? 2 * std::sqrt(reach2) : 0.0;
  const int64_t H =
      opt.cell_policy == PAC_CELL_REFERENCE_RAW
          ? (int64_t)std::ceil(0.5 * D / (double)g.step) + 4
          : (int64_t)std::ceil(D / (double)g.step) + 2;
  auto first_at = [&](uint64_t k) -> int64_t { /* first sorted id, key >= k */
    return (int64_t)(std::lower_bound(hkey.begin(), hkey.end(), k,
                                      [](uint32_t a, uint64_t b) {
                                        return (uint64_t)a < b;
                                      }) -
                     hkey.begin());
  };

  pac_stats acc;
  std::memset(&acc, 0, sizeof(acc));
  std::vector<DBuf<int32_t>> piece[3];
  std::vector<DBuf<uint8_t>> gpiece[2]; /* filtration flags of the rows */
  std::vector<int64_t> prows[3];
  DBuf<uint8_t> flag(P, n);
  DBuf<uint32_t> pos(P, n + 1);
  DBuf<uint8_t> stmp(P, scan_tmp_bytes(n) + 64);
  DBuf<uint64_t> rb(P, 6);
  uint64_t hb6[6];
  for (int64_t lo = 0; lo < n; lo += C) {
    const int64_t hi = std::min(n, lo + C);
    const int64_t czl = (int64_t)(hkey[lo] / exy), czh = (int64_t)(hkey[hi - 1] / exy);
    const int64_t za = std::max<int64_t>(0, czl - H), zb = czh + H;
    const int64_t slo = first_at((uint64_t)za * exy);
    const int64_t shi = first_at((uint64_t)(zb + 1) * exy);
    const int64_t m = shi - slo;
    if (slo > lo || shi < hi || m <= 0) fail(PAC_ERR_DEVICE, "internal: chunk halo");
    Timer Tc(s);
    Tc.mark(0);
    launch_rank_flag(rank.get(), n, (uint32_t)slo, (uint32_t)shi, flag.get(), s);
    scan_u8flag_to_u32(flag.get(), pos.get(), Cnt(n), stmp.get(), 0, s);
    DBuf<Ball> sraw(P, m);
    DBuf<uint32_t> sid(P, m);
    launch_select_balls(raw, flag.get(), pos.get(), n, m, sraw.get(), sid.get(), s);
    KCHK();
    Tc.mark(1);
    pac_result Rs;
    Rs.eng = E;
    Rs.alpha = (double)alpha;
    std::memset(&Rs.st, 0, sizeof(Rs.st));
    SubInput sub;
    sub.grid = gb;
    sub.own_lo = (uint32_t)(lo - slo);
    sub.own_hi = (uint32_t)(hi - slo);
    run_pipeline(E, sraw.get(), m, solvent, alpha, opt, &Rs, s, Tc, KP,
                 nullptr, 1, nullptr, &sub);
    for (int d = 1; d <= 3; ++d)
      launch_row_bounds(Rs.lists[d - 1].get(), Rs.counts[d], d + 1, sid.get(),
                        rank.get(), (uint32_t)lo, (uint32_t)hi,
                        rb.get() + 2 * (d - 1), s);
    KCHK();
    HIPCHK(hipMemcpyAsync(hb6, rb.get(), sizeof(hb6), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int d = 1; d <= 3; ++d) {
      const int64_t r0 = (int64_t)hb6[2 * (d - 1)], r1 = (int64_t)hb6[2 * d - 1];
      if (r0 > r1 || r1 > Rs.counts[d]) fail(PAC_ERR_DEVICE, "internal: chunk rows");
      if (r1 == r0) continue;
      DBuf<int32_t> out(P, (r1 - r0) * (d + 1));
      launch_map_ids(Rs.lists[d - 1].get() + r0 * (d + 1), (r1 - r0) * (d + 1),
                     sid.get(), out.get(), s);
      piece[d - 1].push_back(std::move(out));
      prows[d - 1].push_back(r1 - r0);
      if (opt.filtration && d <= 2) {
        DBuf<uint8_t> g(P, r1 - r0);
        HIPCHK(hipMemcpyAsync(g.get(), Rs.gab[d - 1].get() + r0, (size_t)(r1 - r0),
                              hipMemcpyDeviceToDevice, s));
        gpiece[d - 1].push_back(std::move(g));
      }
    }
    KCHK();
    Tc.mark(9);
    HIPCHK(hipStreamSynchronize(s));
    for (int q = 2; q < 8; ++q) Rs.st.phase_ms[q] = Tc.ms(q, q + 1);
    Rs.st.phase_ms[1] = Tc.ms(0, 2);  /* selection + binning */
    Rs.st.phase_ms[0] = Tc.ms(8, 9);  /* owned rows out */
    add_stats(acc, Rs.st);
  }

  /* concatenate the chunks' rows */
  for (int d = 1; d <= 3; ++d) {
    int64_t rows = 0;
    for (int64_t r : prows[d - 1]) rows += r;
    R->lists[d - 1] = DBuf<int32_t>(P, rows * (d + 1));
    int64_t at = 0;
    for (size_t k = 0; k < piece[d - 1].size(); ++k) {
      const int64_t len = prows[d - 1][k] * (d + 1);
      HIPCHK(hipMemcpyAsync(R->lists[d - 1].get() + at, piece[d - 1][k].get(),
                            (size_t)len * sizeof(int32_t),
                            hipMemcpyDeviceToDevice, s));
      at += len;
    }
    R->counts[d] = rows;
    if (opt.filtration && d <= 2) {
      R->gab[d - 1] = DBuf<uint8_t>(P, rows);
      int64_t g = 0;
      for (size_t k = 0; k < gpiece[d - 1].size(); ++k) {
        const int64_t len = prows[d - 1][k];
        HIPCHK(hipMemcpyAsync(R->gab[d - 1].get() + g, gpiece[d - 1][k].get(),
                              (size_t)len, hipMemcpyDeviceToDevice, s));
        g += len;
      }
    }
  }
  if (opt.filtration) R->rank = std::move(rank);
  HIPCHK(hipStreamSynchronize(s));
  for (int q = 0; q < 8; ++q) R->phase_add[q] = acc.phase_ms[q];
  acc.n = n;
  acc.edges_final = R->counts[1];
  acc.tris_final = R->counts[2];
  acc.tets_final = R->counts[3];
  acc.grid_cells = (int32_t)gp.G;
  R->st = acc;
}
}  // namespace PAC_V
}  // namespace pac
