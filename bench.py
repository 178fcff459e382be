#!/usr/bin/env python3
"""Benchmark: classified simplices/s on the 100,000-ball globule, staged
precision path vs plain FP64 filter path, bit-exact.

One step = one full pipeline run (binning -> enumeration + alpha filters ->
emptiness tests -> closure -> output lists on the device) over the
device-resident 100k-ball input.  Classified simplices = candidate edges +
candidate triangles + candidate tetrahedra (each decided by an alpha
filter; survivors additionally by an emptiness test).

Prints one JSON line.  `value` is the staged path's throughput; the plain
FP64 path runs on the same input in the same process and its throughput,
the ratio, the per-stage escalation counts and the bit-exactness checks
(staged == plain == CPU oracle) are reported beside it.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "parallelac-xmp_amd"))

METRIC = "classified simplices/s at 100k balls, staged vs plain FP64 filter, bit-exact"
# algorithmic work per unit (flops counted from the kernels' source) and
# MI355X peaks: bench/work_model.json
WORK_MODEL = json.load(open(os.path.join(ROOT, "bench", "work_model.json")))
PEAK_TFLOPS = {"f32": WORK_MODEL["peaks"]["valu_f32_TFLOPs"],
               "f64": WORK_MODEL["peaks"]["valu_f64_TFLOPs"]}
KERNEL_KEYS = ("edge_fill", "tri_fill", "tet_fill", "tet_ac2", "tri_ac2", "edge_ac2")


def kernel_flops(key, units, alpha_tets, fp64):
    """(FP32 flops, FP64 flops) of one launch of kernel `key` over `units`
    work units; the tet emptiness scan adds its per-alpha-tet setup.  On the
    plain FP64 path every FP32 flop runs in FP64."""
    f = WORK_MODEL["flops"][key]
    f32, f64 = f["f32"] * units, f["f64"] * units
    if key == "tet_ac2":
        su = WORK_MODEL["flops"]["tet_ac2_setup"]
        f32 += su["f32"] * alpha_tets
        f64 += su["f64"] * alpha_tets
    return (0.0, f32 + f64) if fp64 else (float(f32), float(f64))


def valu_roofline(f32, f64, ms):
    """achieved TFLOP/s and the peak of this FP32/FP64 mix: the time the
    mix takes at the published peaks, t* = f32/P32 + f64/P64, gives the
    mix's peak (f32 + f64)/t*; frac = t*/t."""
    total = f32 + f64
    t_star = f32 / (PEAK_TFLOPS["f32"] * 1e12) + f64 / (PEAK_TFLOPS["f64"] * 1e12)
    achieved = total / (ms * 1e-3) / 1e12 if ms > 0 else 0.0
    peak = total / t_star / 1e12 if t_star > 0 else PEAK_TFLOPS["f32"]
    return achieved, peak


def cpu_info(threads):
    """the host CPU the baseline ran on: model string, logical CPUs seen by
    the OS and by this process, threads used and why."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": affinity,
            "threads": threads,
            "threads_note": "OpenMP threads in one process = OMP_NUM_THREADS if set "
                            "(the host's CPU share per GPU), else every CPU this "
                            "process may run on"}


def cpu_threads(args):
    if args.cpu_threads:
        return args.cpu_threads
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=100000)
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU oracle leg (baseline + bit-exact check)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--budget", type=int, default=0,
                    help="emptiness phase-one pair budget (with --grid: skip "
                         "the autotuner, e.g. for profiler runs)")
    ap.add_argument("--grid", type=int, default=0,
                    help="second emptiness phase grid (blocks)")
    ap.add_argument("--fine", type=int, default=0,
                    help="emptiness scan grid x split (with --budget and --grid)")
    ap.add_argument("--quick", type=int, default=0,
                    help="quick rejection pass (with --budget and --grid)")
    ap.add_argument("--ac2-block", type=int, default=0,
                    help="threads per block of the persistent second "
                         "emptiness phase (with --budget and --grid)")
    ap.add_argument("--ac2-chunk", type=int, default=0,
                    help="queue granularity of the persistent second "
                         "emptiness phase (with --budget and --grid)")
    ap.add_argument("--graph-pass-balls", type=int, default=0,
                    help="batch workload: balls per pass of the recorded "
                         "graph (0 = the library's default, 2 Mi)")
    ap.add_argument("--chunk", type=int, default=100000,
                    help="stress workload: also run the chunked (LD) driver "
                         "with this chunk_balls (0 = skip)")
    ap.add_argument("--workload", choices=("globule", "batch", "stress", "copy"),
                    default="globule",
                    help="globule: the 100k headline (default); batch: the "
                         "4096-molecule ensemble; stress: 2M balls with a "
                         "degenerate lattice block; copy: a 1 GiB device "
                         "copy, the reference kernel for HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this engine runs on one GPU; multi-GPU runs are not "
              "part of this build (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    log("importing torch")
    import torch
    import pac

    if args.workload == "batch":
        return bench_batch(args, torch, pac)
    if args.workload == "stress":
        return bench_stress(args, torch, pac)
    if args.workload == "copy":
        return bench_copy(args, torch)

    n = args.n
    radius = (3.0 * n / (4.0 * np.pi * 0.05)) ** (1.0 / 3.0)
    if n == 100000:
        radius = 78.2
    balls = pac.synth_globule(n, radius, 1.3, seed=2, decimals=3)
    solvent, alpha = 1.4, 0.0
    dev = torch.from_numpy(balls).to("cuda:0")
    torch.cuda.synchronize()
    ptr = dev.data_ptr()

    # the library's autotuner picks the emptiness settings per path
    # (untimed; outputs do not depend on them)
    tuned = {}
    for prec in (pac.PREC_STAGED, pac.PREC_PLAIN_FP64):
        if args.budget and args.grid:
            tuned[prec] = {"ac2_budget": args.budget, "ac2_grid": args.grid,
                           "ac2_fine": args.fine, "ac2_quick": args.quick,
                           "block": args.ac2_block, "ac2_chunk": args.ac2_chunk}
        else:
            tuned[prec], _ = pac.autotune(ptr, solvent, alpha, device_inputs=True,
                                          n=n, precision=prec)
    log("settings: staged %s plain %s" % (tuned[pac.PREC_STAGED],
                                           tuned[pac.PREC_PLAIN_FP64]))

    def run(prec, profile=False):
        return pac.compute_device(ptr, n, solvent, alpha, precision=prec,
                                  profile=profile, **tuned[prec])

    log("warmup")
    for _ in range(args.warmup):
        run(pac.PREC_STAGED)
        run(pac.PREC_PLAIN_FP64)
    log("timing")

    def timed(prec):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = None
        for _ in range(args.steps):
            r = None
            r = run(prec)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        return r, (t1 - t0) / args.steps

    r_st, t_st = timed(pac.PREC_STAGED)
    r_pl, t_pl = timed(pac.PREC_PLAIN_FP64)
    s_st, s_pl = r_st.stats(), r_pl.stats()
    classified = int(s_st.edge_candidates + s_st.tri_candidates + s_st.tet_candidates)
    L_st, L_pl = r_st.lists(), r_pl.lists()
    same_paths = all(np.array_equal(L_st[d], L_pl[d]) for d in (1, 2, 3))

    # live per-kernel device times (hipEvents on the library's stream)
    p_st = run(pac.PREC_STAGED, profile=True).stats()
    p_pl = run(pac.PREC_PLAIN_FP64, profile=True).stats()
    kms = {pac.KERNELS[i]: p_st.kernel_ms[i] for i in range(8)}
    units = {"edge_fill": s_st.edge_candidates, "tri_fill": s_st.tri_candidates,
             "tet_fill": s_st.tet_candidates, "tet_ac2": s_st.ac2_pairs[0],
             "tri_ac2": s_st.ac2_pairs[1], "edge_ac2": s_st.ac2_pairs[2]}
    dom = max(KERNEL_KEYS, key=lambda k: kms[k])
    f32, f64 = kernel_flops(dom, units[dom], s_st.tet_alpha, fp64=False)
    achieved, peak = valu_roofline(f32, f64, kms[dom])
    # the same model for every filter kernel (vector FP32/FP64 rate), both paths
    kms_pl = {pac.KERNELS[i]: p_pl.kernel_ms[i] for i in range(8)}
    units_pl = {"edge_fill": s_pl.edge_candidates, "tri_fill": s_pl.tri_candidates,
                "tet_fill": s_pl.tet_candidates, "tet_ac2": s_pl.ac2_pairs[0],
                "tri_ac2": s_pl.ac2_pairs[1], "edge_ac2": s_pl.ac2_pairs[2]}
    rates, fracs = {}, {}
    for k in KERNEL_KEYS:
        for tag, km, un, fp64 in (("staged", kms, units, False),
                                  ("plain", kms_pl, units_pl, True)):
            a32, a64 = kernel_flops(k, un[k], (s_st if not fp64 else s_pl).tet_alpha, fp64)
            ach, pk = valu_roofline(a32, a64, km[k])
            rates.setdefault(k, {})[tag] = round(ach, 3) if km[k] > 0 else None
            fracs.setdefault(k, {})[tag] = round(ach / pk, 4) if km[k] > 0 else None
    traffic, traffic_src = None, None
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic_globule100k.json")
    if os.path.exists(tpath) and dom in ("tet_ac2",):
        tj = json.load(open(tpath))
        traffic = tj.get("bytes_per_run", {}).get(dom)
        traffic_src = "profiles/hbm_traffic_globule100k.json (rocprofv3 --pmc " \
                      "FETCH_SIZE / WRITE_SIZE, scaled by a 1 GiB copy)"
    roofline = {"bound": "valu", "achieved": achieved, "peak": peak,
                "unit": "TFLOP/s", "frac": achieved / peak if peak else None,
                "traffic": traffic, "traffic_source": traffic_src,
                "kernel": dom, "kernel_ms": kms[dom],
                "work_units": int(units[dom]),
                "flops_per_unit": WORK_MODEL["flops"][dom],
                "setup_units": int(s_st.tet_alpha) if dom == "tet_ac2" else 0,
                "flops_per_setup": WORK_MODEL["flops"]["tet_ac2_setup"]
                if dom == "tet_ac2" else None,
                "flops_f32": f32, "flops_f64": f64,
                "note": "vector ALU (no MFMA: nothing on this path is a dense "
                        "contraction); peak = the published MI355X vector peaks "
                        "for this kernel's FP32/FP64 mix (FP32 157.3, FP64 78.6 "
                        "TFLOP/s), t* = f32/157.3 + f64/78.6; without packed FP32 "
                        "(not used in this build) the attainable FP32 rate is 78.6; "
                        "work model: bench/work_model.json"}

    # the same run recorded as one graph (one molecule): device time of the
    # pipeline without its host read-backs; buffer capacities come from an
    # earlier run on this input, so this is reported beside `value`, not as it
    g = pac.Graph(ptr, np.array([0, n], dtype=np.int64), solvent, alpha,
                  device_inputs=True, precision=pac.PREC_STAGED,
                  **tuned[pac.PREC_STAGED])
    for _ in range(args.warmup):
        g.launch()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.launch()
    torch.cuda.synchronize()
    t_gr = (time.perf_counter() - t0) / args.steps
    gl = g.result().lists(0)
    graph_eq = all(np.array_equal(gl[d], L_st[d]) for d in (1, 2, 3))
    del g
    log("gpu legs done: staged %.3f ms plain %.3f ms graph %.3f ms"
        % (t_st * 1e3, t_pl * 1e3, t_gr * 1e3))

    # the same run with the filtration (entry values of every simplex and
    # the sorted filtration copied to the host)
    def filt_run():
        rf = pac.compute_device(ptr, n, solvent, alpha, precision=pac.PREC_STAGED,
                                filtration=True, **tuned[pac.PREC_STAGED])
        return rf, rf.filtration()
    filt_run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rf, (fd, fi, fv) = filt_run()
    torch.cuda.synchronize()
    t_f = time.perf_counter() - t0
    sf = rf.stats()
    filt = {"ms": t_f * 1e3, "entries": int(len(fd)),
            "lists_eq": all(np.array_equal(rf.lists()[d], L_st[d]) for d in (1, 2, 3)),
            "tri_ac2_pairs_vs_plain_run": sf.ac2_pairs[1] / max(1, s_st.ac2_pairs[1]),
            "edge_ac2_pairs_vs_plain_run": sf.ac2_pairs[2] / max(1, s_st.ac2_pairs[2])}
    del rf, fd, fi, fv
    cpu = None
    oracle_exact = None
    if not args.no_cpu:
        log("cpu oracle leg")
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        threads = cpu_threads(args)
        t0 = time.perf_counter()
        o, ost, _ = oracle.run(balls, solvent, alpha, nthreads=threads)
        t1 = time.perf_counter()
        ocl = ost["E_cand"] + ost["T_cand"] + ost["K_cand"]
        cpu = {"value": ocl / (t1 - t0), "unit": "classified simplices/s",
               "cores": threads, "kind": "port",
               "sample": "full 100k-ball workload, one run of the C++ oracle "
                         "(FP64 filter + exact integer fallback), %.2f s" % (t1 - t0),
               "host": cpu_info(threads)}
        oracle_exact = all(np.array_equal(L_st[d], o[d]) for d in (1, 2, 3))

    esc = {pac.FAMILIES[f]: {pac.STAGES[s]: int(s_st.stage_evals[f][s]) for s in range(4)}
           for f in range(6)}
    line = {
        "metric": METRIC,
        "value": classified / t_st,
        "unit": "classified simplices/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_st * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32->f64->dd->int512",
        "data": "synthetic",
        "config": {"workload": "globule-100k: %d balls, Poisson-disk min 1.3 A in a "
                               "sphere r=%.1f A, radii {1.70,1.55,1.52,1.80}, coords "
                               "rounded 0.001 A, FP32, solvent 1.4, alpha 0" % (n, radius),
                   "n_balls": n, "solvent": solvent, "alpha": alpha, "seed": 2,
                   "cell_policy": "exact", "order": "reference"},
        "plain_fp64": {"value": classified / t_pl, "ms_per_step": t_pl * 1e3,
                       "dtype": "f64->int512", "tuned": tuned[pac.PREC_PLAIN_FP64]},
        "tuned": tuned[pac.PREC_STAGED],
        "staged_over_plain": t_pl / t_st,
        "staged_one_graph_launch": {"ms_per_step": t_gr * 1e3, "eq_eager": graph_eq},
        "filtration": filt,
        "bit_exact": {"staged_eq_plain": same_paths, "staged_eq_oracle": oracle_exact},
        "counts": {"classified": classified,
                   "edge_cand": int(s_st.edge_candidates), "tri_cand": int(s_st.tri_candidates),
                   "tet_cand": int(s_st.tet_candidates), "edges": int(s_st.edges_final),
                   "triangles": int(s_st.tris_final), "tets": int(s_st.tets_final)},
        "escalations": esc, "sos_calls": int(s_st.sos_calls),
        "phase_ms": {pac.PHASE_LABELS[i]: round(s_st.phase_ms[i], 3) for i in range(8)},
        "kernel_ms_staged": {k: round(v, 3) for k, v in kms.items()},
        "kernel_ms_plain": {pac.KERNELS[i]: round(p_pl.kernel_ms[i], 3) for i in range(8)},
        "kernel_tflops": rates,
        "kernel_roofline_frac": fracs,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


def _classified(st):
    return int(st.edge_candidates + st.tri_candidates + st.tet_candidates)


def _esc(pac, st):
    return {pac.FAMILIES[f]: {pac.STAGES[s]: int(st.stage_evals[f][s]) for s in range(4)}
            for f in range(6)}


def bench_batch(args, torch, pac):
    """4096 molecules, sizes log-uniform in [200, 20000] (seed 3), each in
    its own sphere at 0.05 balls/A^3; one pac_compute_batch call per step
    over the device-resident concatenated array.  A sample of molecules is
    checked against single-molecule runs (and two against the oracle)."""
    n_mol = 4096
    xyzr, off = pac.synth_batch(n_mol, 200, 20000, seed=3)
    ntot = int(off[-1])
    dev = torch.from_numpy(xyzr).to("cuda:0")
    torch.cuda.synchronize()
    ptr = dev.data_ptr()

    tuned = {}
    for prec in (pac.PREC_STAGED, pac.PREC_PLAIN_FP64):
        tuned[prec], _ = pac.autotune(ptr, 1.4, 0.0, offsets=off,
                                      device_inputs=True, precision=prec)
    log("autotuned on the leading molecules: %s" % tuned)

    def run(prec):
        return pac.compute_batch(ptr, off, 1.4, 0.0, precision=prec,
                                 device_inputs=True, **tuned[prec])

    log("batch: %d molecules, %d balls; warmup" % (n_mol, ntot))
    for _ in range(args.warmup):
        run(pac.PREC_STAGED)
    res = {}
    for prec in (pac.PREC_STAGED, pac.PREC_PLAIN_FP64):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = None
        for _ in range(args.steps):
            r = None
            r = run(prec)
        torch.cuda.synchronize()
        res[prec] = (r, (time.perf_counter() - t0) / args.steps)
    r_st, t_st = res[pac.PREC_STAGED]
    r_pl, t_pl = res[pac.PREC_PLAIN_FP64]
    st = r_st.stats()
    rng = np.random.default_rng(3)
    sample = sorted(rng.choice(n_mol, 24, replace=False).tolist())
    same_single = True
    for m in sample:
        single = pac.compute(xyzr[off[m]:off[m + 1]], 1.4, 0.0).lists()
        got = r_st.lists(m)
        got_pl = r_pl.lists(m)
        same_single &= all(np.array_equal(got[d], single[d]) and
                           np.array_equal(got_pl[d], single[d]) for d in (1, 2, 3))
    oracle_ok = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        oracle_ok = True
        for m in sample[:2]:
            o, _, _ = oracle.run(xyzr[off[m]:off[m + 1]], 1.4, 0.0,
                                 nthreads=cpu_threads(args))
            got = r_st.lists(m)
            oracle_ok &= all(np.array_equal(got[d], o[d]) for d in (1, 2, 3))
    # the same batch recorded once as one HIP graph: each step is one
    # hipGraphLaunch of every pass (recording untimed)
    log("graph: recording the batch")
    t0 = time.perf_counter()
    g = pac.Graph(ptr, off, 1.4, 0.0, precision=pac.PREC_STAGED,
                  device_inputs=True, batch_balls=args.graph_pass_balls,
                  **tuned[pac.PREC_STAGED])
    t_rec = time.perf_counter() - t0
    nodes, gbytes = g.info()
    for _ in range(args.warmup):
        g.launch()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.launch()
    torch.cuda.synchronize()
    t_gr = (time.perf_counter() - t0) / args.steps
    rg = g.result()
    graph_eq = bool(np.array_equal(rg.counts(), r_st.counts()))
    for m in sample:
        got, want = rg.lists(m), r_st.lists(m)
        graph_eq &= all(np.array_equal(got[d], want[d]) for d in (1, 2, 3))
    log("graph: %.1f ms per launch (eager %.1f ms), %d nodes, %.1f GB, eq %s"
        % (t_gr * 1e3, t_st * 1e3, nodes, gbytes / 1e9, graph_eq))
    del g
    classified = _classified(st)
    line = {
        "metric": "classified simplices/s, 4096-molecule ensemble, bit-exact per molecule",
        "value": classified / t_st, "unit": "classified simplices/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t_st * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32->f64->dd->int512", "data": "synthetic",
        "config": {"workload": "ensemble-4096: sizes log-uniform [200, 20000] seed 3, "
                               "Poisson-disk min 1.3 A at 0.05 balls/A^3, FP32, "
                               "solvent 1.4, alpha 0", "molecules": n_mol, "balls": ntot},
        "molecules_per_s": n_mol / t_st, "balls_per_s": ntot / t_st,
        "tuned": tuned[pac.PREC_STAGED],
        "plain_fp64": {"value": classified / t_pl, "ms_per_step": t_pl * 1e3},
        "graph": {"value": classified / t_gr, "ms_per_step": t_gr * 1e3,
                  "launches_per_step": 1, "nodes": nodes,
                  "device_bytes": gbytes, "record_s": round(t_rec, 2),
                  "device_ms_last_launch": round(rg.stats().total_ms, 3),
                  "eq_eager": graph_eq},
        "check": {"sampled_molecules": len(sample), "eq_single_runs": bool(same_single),
                  "eq_oracle_2": oracle_ok},
        "counts": {"classified": classified, "edges": int(st.edges_final),
                   "triangles": int(st.tris_final), "tets": int(st.tets_final)},
        "escalations": _esc(pac, st),
        "phase_ms": {pac.PHASE_LABELS[i]: round(st.phase_ms[i], 3) for i in range(8)},
    }
    print(json.dumps(line))


def bench_copy(args, torch):
    """1 GiB device-to-device copy (torch): a kernel of known traffic
    (1 GiB read + 1 GiB written) against which the FETCH_SIZE / WRITE_SIZE
    counters of the pipeline's kernels are scaled."""
    n = 1 << 28
    x = torch.ones(n, dtype=torch.float32, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(args.warmup):
        y.copy_(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        y.copy_(x)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / args.steps
    print(json.dumps({"metric": "device copy bandwidth", "value": 2 * 4 * n / t / 1e9,
                      "unit": "GB/s", "bytes_read": 4 * n, "bytes_written": 4 * n,
                      "ms_per_step": t * 1e3, "steps": args.steps}))


def bench_stress(args, torch, pac):
    """2,000,000 balls in one pass: Poisson-disk in a sphere r = 212.2 A with
    a central block replaced by a 2.0 A cubic lattice of r = 1.70 (10% of
    the balls).  Reports the fallback counts; checked against the oracle."""
    n = 2000000
    balls = pac.synth_stress(n, 212.2, seed=5, lattice_frac=0.1, spacing=2.0,
                             lattice_r=1.70)
    dev = torch.from_numpy(balls).to("cuda:0")
    torch.cuda.synchronize()
    ptr = dev.data_ptr()
    # autotuned per path on a 200k-ball sample of the same construction
    # (10% lattice block); the full input is too large to sweep
    sample = pac.synth_stress(200000, None, seed=6, lattice_frac=0.1,
                              spacing=2.0, lattice_r=1.70)
    tuned = {}
    for prec in (pac.PREC_STAGED, pac.PREC_PLAIN_FP64):
        if args.budget and args.grid:
            tuned[prec] = {"ac2_budget": args.budget, "ac2_grid": args.grid,
                           "ac2_fine": args.fine, "ac2_quick": args.quick,
                           "block": args.ac2_block, "ac2_chunk": args.ac2_chunk}
        else:
            tuned[prec], _ = pac.autotune(sample, 1.4, 0.0, precision=prec)
    log("stress: settings %s" % tuned)
    log("stress: warmup")
    for _ in range(args.warmup):
        pac.compute_device(ptr, n, 1.4, 0.0, **tuned[pac.PREC_STAGED])
    torch.cuda.synchronize()
    pac.memory_stats(reset_peak=True)
    t0 = time.perf_counter()
    r = None
    for _ in range(args.steps):
        r = None
        r = pac.compute_device(ptr, n, 1.4, 0.0, **tuned[pac.PREC_STAGED])
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / args.steps
    _, peak_one, _ = pac.memory_stats(reset_peak=True)
    st = r.stats()
    # per-kernel device times of one more run (hipEvents on the library's stream)
    kst = pac.compute_device(ptr, n, 1.4, 0.0, profile=True,
                             **tuned[pac.PREC_STAGED]).stats()
    log("stress: plain FP64 path")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rp = pac.compute_device(ptr, n, 1.4, 0.0, precision=pac.PREC_PLAIN_FP64,
                            **tuned[pac.PREC_PLAIN_FP64])
    torch.cuda.synchronize()
    t_pl = time.perf_counter() - t0
    L, Lp = r.lists(), rp.lists()
    same_paths = all(np.array_equal(L[d], Lp[d]) for d in (1, 2, 3))
    chunked = None
    if args.chunk > 0:
        # the LD driver (parallelac-largeData <chunk-size>) on the same input
        log("stress: chunked, chunk_balls=%d" % args.chunk)
        kw = dict(tuned[pac.PREC_STAGED], chunk_balls=args.chunk)
        pac.compute_device(ptr, n, 1.4, 0.0, **kw)
        torch.cuda.synchronize()
        pac.memory_stats(reset_peak=True)
        t0 = time.perf_counter()
        rc = pac.compute_device(ptr, n, 1.4, 0.0, **kw)
        torch.cuda.synchronize()
        t_ch = time.perf_counter() - t0
        _, peak_ch, _ = pac.memory_stats(reset_peak=True)
        Lc, stc = rc.lists(), rc.stats()
        c = 512
        while c < args.chunk:
            c *= 2
        chunked = {"chunk_balls": args.chunk, "chunk": c, "chunks": -(-n // c),
                   "ms": t_ch * 1e3,
                   "eq_one_pass": all(np.array_equal(Lc[d], L[d]) for d in (1, 2, 3)),
                   "alpha_tets_vs_one_pass": stc.tet_alpha / max(1, st.tet_alpha),
                   "peak_device_mb": peak_ch / 2**20,
                   "one_pass_peak_device_mb": peak_one / 2**20}
        del rc, Lc
    oracle_ok = None
    cpu = None
    if not args.no_cpu:
        # bounded sample: the oracle decides most lattice simplices exactly
        # (ties, SoS), ~4 ms of CPU per ball there, so the full input would
        # take it hours; the 8000 balls nearest a point on the lattice
        # block's +x face, lattice and Poisson-disk; the GPU runs the same
        # sample
        log("stress: oracle on an 8000-ball sample")
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        threads = cpu_threads(args)
        xyz = balls[:, :3].astype(np.float64)
        ctr = xyz.mean(0)
        half = 0.5 * (0.1 * n * 2.0 ** 3) ** (1.0 / 3.0)
        d2 = ((xyz - (ctr + np.array([half, 0.0, 0.0]))) ** 2).sum(1)
        k = min(n, 8000)
        sample = balls[np.sort(np.argpartition(d2, k - 1)[:k])]
        gs = pac.compute(sample, 1.4, 0.0, **tuned[pac.PREC_STAGED]).lists()
        t0 = time.perf_counter()
        o, ost, _ = oracle.run(sample, 1.4, 0.0, nthreads=threads)
        tc = time.perf_counter() - t0
        oracle_ok = all(np.array_equal(gs[d], o[d]) for d in (1, 2, 3))
        cpu = {"value": (ost["E_cand"] + ost["T_cand"] + ost["K_cand"]) / tc,
               "unit": "classified simplices/s", "cores": threads, "kind": "port",
               "sample": "the %d balls of the stress input nearest the lattice "
                         "block's +x face (GPU lists of the same sample compared), "
                         "%.1f s" % (k, tc), "host": cpu_info(threads)}
    classified = _classified(st)
    line = {
        "metric": "classified simplices/s, 2M-ball degenerate stress, bit-exact",
        "value": classified / t, "unit": "classified simplices/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32->f64->dd->int512", "data": "synthetic",
        "config": {"workload": "stress-2M: Poisson-disk in r=212.2 A, 10% central "
                               "cubic lattice spacing 2.0 A r=1.70, FP32, solvent 1.4, "
                               "alpha 0, one pass", "n_balls": n},
        "bit_exact": {"staged_eq_plain": same_paths,
                      "staged_eq_oracle_on_cpu_sample": oracle_ok},
        "plain_fp64": {"ms_per_step": t_pl * 1e3, "tuned": tuned[pac.PREC_PLAIN_FP64]},
        "tuned": tuned[pac.PREC_STAGED],
        "chunked": chunked,
        "counts": {"classified": classified, "edges": int(st.edges_final),
                   "triangles": int(st.tris_final), "tets": int(st.tets_final)},
        "escalations": _esc(pac, st), "sos_calls": int(st.sos_calls),
        "phase_ms": {pac.PHASE_LABELS[i]: round(st.phase_ms[i], 3) for i in range(8)},
        "kernel_ms_staged": {pac.KERNELS[i]: round(kst.kernel_ms[i], 3) for i in range(8)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
