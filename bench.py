"""Graph500 BFS benchmark on one MI355X: harmonic-mean GTEPS at SCALE 26.

The benchmark unit is the reference's: one direction-optimizing BFS per root
over a set of 64 roots of the Kronecker SCALE-26 graph (edgefactor 16, seed
(2,3)), each run timed around `run_bfs` (mpi/main.cc:610-676).  A step is
one pass over the whole root set (64 BFS runs), so `--steps K` times K x 64
runs and the harmonic mean covers every one of them.  Roots come from the
reference's root rule (find_roots + redraw when a root reaches <= 3 levels,
mpi/main.cc:334-359).  The graph is generated and built on the GPU before
timing; pred stays on the device during the timed BFS (get_pred is outside
the timer, as in the reference).  Every run of the first timed pass is
validated on the device right after it, outside its timer, on that run's own
pred (reference -v 2); its traversed-edge count (TE) is the TEPS numerator.
Later passes must reproduce the first pass's frontier sizes level by level.

The roofline object describes the dominant kernel (the one with the most
device time per BFS in the timed runs, from HIP events on the stream it runs
on): its algorithmic bytes per launch -- counted on the device in an untimed
counting pass over the same roots -- over its mean launch time.  A second
figure does the same for the whole BFS, and the top-down yardstick
4 m_c + 8 (n_c+1) + 12 N (every component edge read once, as a top-down
BFS would) is reported under its own name.

Secondary measurements on the same line: "scale24" (the SCALE-24 drop-in
case, 64 validated roots, with the device time between levels), "batched"
(the bit-parallel multi-source BFS, 64/128/256 roots per launch, traversed
edges/s and its ratio to the plain form) and "serving" (continuous batching
over four resident graphs, SCALE 20/23/25/27, open Poisson arrivals; p50/p99
latency and queries/s beside a one-query-per-launch baseline).

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "graph500-bfs_amd"))

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5,
                    help="timed passes over the root set (one BFS per root each)")
    ap.add_argument("--warmup", type=int, default=1, help="untimed passes before them")
    ap.add_argument("--scale", type=int, default=26)
    ap.add_argument("--edgefactor", type=int, default=16)
    ap.add_argument("--roots", type=int, default=64)
    ap.add_argument("--alpha", type=float, default=10.0)
    ap.add_argument("--beta", type=float, default=14.0)
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="budget for the CPU baseline sample (0 disables)")
    ap.add_argument("--no-validate", action="store_true",
                    help="skip per-root validation (TE then comes from pf_nedge)")
    ap.add_argument("--levels", action="store_true", help="print per-level trace of one root")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched multi-source BFS measurement")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop measurement")
    ap.add_argument("--no-scale24", action="store_true", help="skip the SCALE-24 drop-in leg")
    ap.add_argument("--batched-no-tune", action="store_true",
                    help="batched leg with the default tuning (no autotuning; for profiling)")
    ap.add_argument("--serve-queries", type=int, default=8000)
    ap.add_argument("--serve-qps", type=float, default=2000.0)
    ap.add_argument("--serve-engines", type=int, default=2,
                    help="batched engines per resident graph in the serving loop")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


PF_NEDGE = None


def profiled_traffic(args, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary of this same workload (None for any other
    workload or kernel)."""
    if (args.scale, args.edgefactor, args.alpha, args.beta) != (26, 16, 10.0, 14.0):
        return None
    path = os.path.join(ROOT, "profiles", "bfs_scale26_hbm_traffic.md")
    try:
        for line in open(path):
            f = line.split()
            # "  <kernel>  fetch <MB> MB  write <MB> MB  per launch"
            if len(f) >= 7 and f[0] == kernel and f[1] == "fetch" and f[4] == "write":
                return int((float(f[2]) + float(f[5])) * 1e6)
    except OSError:
        pass
    return None


def pf_nedge(scale):
    global PF_NEDGE
    if PF_NEDGE is None:
        with open(os.path.join(ROOT, "tests", "golden", "reference_counts.json")) as f:
            PF_NEDGE = json.load(f)["pf_nedge"]
    return PF_NEDGE[scale] if scale < len(PF_NEDGE) else None


def cpu_threads():
    """Host threads for the CPU baseline: OMP_NUM_THREADS when set, else the
    cores this process may run on, at most 16."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return max(1, min(16, len(os.sched_getaffinity(0))))


def cpu_baseline(graph, roots, te_by_root, budget_s, args):
    """The reference's CPU path restated (the oracle's C direction-optimizing
    BFS with OpenMP over the host threads, as the reference splits its
    top-down frontier and bottom-up vertex range over threads) on the same
    graph and roots, in order, until the time budget is spent."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import ctypes as C
    import oracle_lib as O
    inf = graph.info
    csr = graph.csr()
    og = O.lib().orc_graph_wrap(
        inf.log_n, inf.n_c, inf.m,
        csr["row_off"].ctypes.data_as(C.POINTER(C.c_int64)),
        csr["col"].ctypes.data_as(C.POINTER(C.c_uint32)),
        csr["deg_new"].ctypes.data_as(C.POINTER(C.c_int64)),
        csr["new2orig"].ctypes.data_as(C.POINTER(C.c_uint32)),
        csr["orig2new"].ctypes.data_as(C.POINTER(C.c_uint32)))
    pred = np.empty(inf.n, np.int64)
    at = np.empty(9, np.int64)
    dirs = np.zeros(256, np.int8)
    nt = cpu_threads()
    times, tes = [], []
    t_start = time.time()
    for r in roots:
        t0 = time.perf_counter()
        O.lib().orc_do_bfs_omp(og, int(r), args.edgefactor, args.alpha, args.beta,
                               pred.ctypes.data_as(C.POINTER(C.c_int64)),
                               at.ctypes.data_as(C.POINTER(C.c_int64)),
                               dirs.ctypes.data_as(C.POINTER(C.c_int8)), 256, nt)
        times.append(time.perf_counter() - t0)
        tes.append(te_by_root[int(r)])
        if time.time() - t_start > budget_s:
            break
    O.lib().orc_graph_free(og)
    hm = len(times) / sum(t / e for t, e in zip(times, tes))
    return {"value": round(hm / 1e9, 3), "unit": "GTEPS", "cores": nt, "kind": "port",
            "label": "restated: the reference's CPU path is not on this host",
            "sample": f"{len(times)} of the {len(roots)} timed roots of the same SCALE "
                      f"{args.scale} graph, in order, direction-optimizing BFS in C with OpenMP "
                      f"over {nt} host threads (oracle/g500_oracle.c orc_do_bfs_omp), "
                      f"{sum(times):.1f} s; TE per root from the GPU run's validation"}


def batched(g500bfs, graph, bfs, args, plain_teps):
    """Batched bit-parallel BFS on the same graph: tuned per batch size on one
    root set, the tuned configuration kept only if it also beats the default
    on a second set, timed on a third; levels spot-checked against the plain
    engine."""
    mb = g500bfs.MultiBfs(graph, max_roots=256)
    tune_roots = graph.find_roots(256, r1=101)
    check_roots = graph.find_roots(256, r1=202)

    def batch_ms(tuning, r):
        mb.set_tuning(**tuning)
        mb.run(r)
        best = float("inf")
        for _ in range(2):
            mb.run(r)
            best = min(best, mb.stats()["gpu_ms"])
        return best

    default = mb.tuning()
    roots = graph.find_roots(256, r1=0)
    out = {}
    for nr in (64, 128, 256):
        mb.set_tuning(**default)
        if args.batched_no_tune:
            tuning = default
        else:
            tuning, tune_ms = mb.autotune(tune_roots[:nr], rounds=1)
            t_tuned = batch_ms(tuning, check_roots[:nr])
            t_default = batch_ms(default, check_roots[:nr])
            log(f"batched {nr}: tuning {tuning} ({tune_ms:.2f} ms per batch of the tuning roots; "
                f"second set: {t_tuned:.2f} ms tuned, {t_default:.2f} ms default)")
            if t_default < t_tuned:
                tuning = default
            mb.set_tuning(**tuning)
        r = roots[:nr]
        mb.run(r)
        host, dev = [], []
        for _ in range(5):
            t0 = time.perf_counter()
            mb.run(r)
            host.append(time.perf_counter() - t0)
            dev.append(mb.stats()["gpu_ms"])
        te = mb.te()
        t = float(np.median(host))
        exact = True
        for i in (0, nr // 2, nr - 1):
            bfs.run(int(r[i]), args.edgefactor, args.alpha, args.beta)
            l16 = bfs.levels()
            exact &= bool(np.array_equal(mb.levels(i),
                                         np.where(l16 == 0xFFFF, 0xFF, l16).astype(np.uint8)))
        teps = float(te.sum()) / t
        out[str(nr)] = {"tuning": tuning,
                        "ms_per_batch": round(t * 1e3, 3),
                        "gpu_ms_per_batch": round(float(np.median(dev)), 3),
                        "roots_per_s": round(nr / t, 1),
                        "traversed_edges_per_s": round(teps, 1),
                        "x_plain_teps": round(teps / plain_teps, 2),
                        "levels_bit_exact_spot_check": exact}
        log(f"batched {nr}: {t * 1e3:.2f} ms, {teps / 1e12:.3f} TTEPS, "
            f"{teps / plain_teps:.2f}x plain, exact={exact}")
    mb.destroy()
    return out


def serving(g500bfs, args):
    """Continuous-batching serving loop over four resident Kronecker graphs
    (SCALE 20, 23, 25, 27; seeds (2+i, 3+i)), open Poisson arrivals."""
    import serve
    graphs, pools = [], []
    for i, scale in enumerate((20, 23, 25, 27)):
        g = g500bfs.Graph.generate(scale, 16, 2 + i, 3 + i).construct()
        graphs.append(g)
        pools.append(g.find_roots(512))
    # two engines per graph: the server overlaps a second batch of a graph
    # only when a full one is waiting, so more engines add no small batches
    srv = serve.Server(graphs, max_batch=256, max_engines_per_graph=args.serve_engines)
    srv.prewarm(args.serve_engines, pools)
    qs = serve.poisson_workload(graphs, pools, args.serve_queries, args.serve_qps, seed=7)
    t0 = time.perf_counter()
    done = serve.run_open_loop(srv, qs)
    wall = time.perf_counter() - t0
    res = serve.latency_summary(done)
    res.update({"offered_qps": args.serve_qps, "wall_s": round(wall, 3),
                "batches": srv.batches,
                "engines": [len(es) for es in srv.engines],
                "engine_gib": round(srv.used / 2**30, 1)})
    srv.close()
    sample = [serve.Query(q.qid, q.gid, q.root, q.targets) for q in done[:200]]
    base = serve.one_per_launch(graphs, sample)
    same = all(a.te == b.te and np.array_equal(a.levels, b.levels)
               for a, b in zip(done[:200], sample))
    res.update({"one_query_per_launch": base, "answers_equal_alone": bool(same),
                "graphs": "Kronecker SCALE 20/23/25/27, edgefactor 16, seeds (2,3)..(5,6)"})
    for k, v in list(res.items()):
        if isinstance(v, float):
            res[k] = round(v, 3)
    log(f"serving: {res}")
    for g in graphs:
        g.free()
    return res


def select_roots(graph, bfs, args):
    """Roots by the reference's rule, redrawn while any reaches <= 3 levels
    (mpi/main.cc:334-359)."""
    r1 = 0
    while True:
        roots = graph.find_roots(args.roots, r1=r1)
        ok = True
        for r in roots:
            at = bfs.run(int(r), args.edgefactor, args.alpha, args.beta)
            if args.scale > 6 and at[8] <= 3:
                ok = False
        if ok:
            return roots
        r1 += 1
        log("picking roots again")


def frontier_trace(bfs):
    return tuple((t["top_down"], t["nq_size"]) for t in bfs.stats()[1])


def scale24_leg(g500bfs, args):
    """The SCALE-24 drop-in case: 64 roots by the reference rule, each timed
    run validated on its own pred; then one untimed pass with level timing for
    the device time between levels."""
    scale = 24
    graph = g500bfs.Graph.generate(scale, args.edgefactor, 2, 3).construct()
    bfs = g500bfs.Bfs(graph)
    sub = argparse.Namespace(**vars(args))
    sub.scale = scale
    roots = select_roots(graph, bfs, sub)
    golden = pf_nedge(scale) if args.edgefactor == 16 else None
    times, gpu, tes, valid = [], [], [], True
    for r in roots:
        r = int(r)
        t0 = time.perf_counter()
        bfs.run(r, args.edgefactor, args.alpha, args.beta)
        times.append(time.perf_counter() - t0)
        gpu.append(bfs.stats(1)[0].gpu_ms)
        ok, te, _ = bfs.validate(r)
        valid &= ok and (golden is None or te == golden)
        tes.append(te)
    times, gpu, tes = np.array(times), np.array(gpu), np.array(tes, float)
    bfs.set_timing(True)
    gaps, nb, levels, miss = 0.0, 0, 0, 0
    for r in roots:
        bfs.run(int(r), args.edgefactor, args.alpha, args.beta)
        st = bfs.stats(1)[0]
        gaps += st.gap_ms
        nb += st.gaps
        levels += st.nlevels
        miss += st.mispredicted
    bfs.set_timing(False)
    out = {"workload": "plain single-root BFS, Kronecker SCALE 24, edgefactor 16, 64 roots",
           "value_gteps": round(len(times) / np.sum(times / tes) / 1e9, 3),
           "gpu_event_gteps": round(len(gpu) / np.sum(gpu * 1e-3 / tes) / 1e9, 3),
           "median_bfs_ms": round(float(np.median(times)) * 1e3, 4),
           "all_roots_valid": bool(valid), "te_per_root": int(tes[0]),
           "levels_per_bfs": round(levels / len(roots), 2),
           "between_steps_us": round(gaps / max(nb, 1) * 1e3, 2),
           "mispredicted_steps_per_bfs": round(miss / len(roots), 2),
           "between_steps_note": "device time from the end of one step's kernels (a level, or "
                                 "the frontier conversion at the bottom-up -> top-down switch) "
                                 "to the start of the next step's, in runs with level timing; "
                                 "the next step is queued before the previous one reports, on "
                                 "the direction the host predicts, and a wrong guess costs "
                                 "one host round trip"}
    log(f"scale24: {out}")
    bfs.destroy()
    graph.free()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): the engine "
              "works on one GPU's graph; sharding over several GPUs is not built.", flush=True)
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0:
        print("--steps must be >= 1 and --warmup >= 0", flush=True)
        sys.exit(2)
    import g500bfs
    g500bfs.load()

    t0 = time.perf_counter()
    graph = g500bfs.Graph.generate(args.scale, args.edgefactor, 2, 3)
    t_gen = time.perf_counter() - t0
    t0 = time.perf_counter()
    graph.construct()
    t_con = time.perf_counter() - t0
    inf = graph.info
    log(f"graph: scale {args.scale} n={inf.n} n_c={inf.n_c} m={inf.m} "
        f"gen {t_gen:.2f}s construct {t_con:.2f}s dev {inf.device_bytes / 2**30:.1f} GiB")

    bfs = g500bfs.Bfs(graph)
    roots = [int(r) for r in select_roots(graph, bfs, args)]
    nr = len(roots)
    for _ in range(args.warmup):
        for r in roots:
            bfs.run(r, args.edgefactor, args.alpha, args.beta)

    # timed: `steps` passes over the root set, one host timer per run
    # (mpi/main.cc:618-626); the first pass validates each run's own pred
    # right after it, outside its timer
    golden = pf_nedge(args.scale) if args.edgefactor == 16 else None
    runs = args.steps * nr
    times = np.zeros(runs)
    gpu_ms = np.zeros(runs)
    gather_ms = np.zeros(runs)
    sweep_ms = np.zeros(runs)
    sweeps = np.zeros(runs, np.int64)
    mispredicted = np.zeros(runs, np.int64)
    te_by_root, trace_by_root = {}, {}
    all_valid, traces_repeat = True, True
    for k in range(runs):
        r = roots[k % nr]
        t0 = time.perf_counter()
        bfs.run(r, args.edgefactor, args.alpha, args.beta)
        times[k] = time.perf_counter() - t0
        st, _ = bfs.stats(1)
        gpu_ms[k], gather_ms[k], sweep_ms[k], sweeps[k] = (st.gpu_ms, st.gather_ms,
                                                           st.sweep_ms, st.sweeps)
        mispredicted[k] = st.mispredicted
        if k < nr:
            trace_by_root[r] = frontier_trace(bfs)
            if args.no_validate:
                te_by_root[r] = golden
            else:
                ok, te, errs = bfs.validate(r)
                all_valid &= ok
                te_by_root[r] = te
                if not ok:
                    log(f"root {r}: validation failed {errs}")
                if golden is not None and te != golden:
                    log(f"root {r}: TE {te} != pf_nedge[{args.scale}] {golden}")
        elif frontier_trace(bfs) != trace_by_root[r]:
            traces_repeat = False
            log(f"root {r}: frontier sizes differ from the first pass")
    tes = np.array([te_by_root[roots[k % nr]] for k in range(runs)], float)
    teps_hm = runs / np.sum(times / tes)
    gpu_teps_hm = runs / np.sum(gpu_ms * 1e-3 / tes)

    # untimed counting pass over the same roots: the bytes each run needs
    bfs.set_counting(True)
    alg, alg_sweep, alg_gather = {}, {}, {}
    for r in roots:
        bfs.run(r, args.edgefactor, args.alpha, args.beta)
        st, _ = bfs.stats(1)
        if frontier_trace(bfs) != trace_by_root[r]:
            traces_repeat = False
        alg[r], alg_sweep[r], alg_gather[r] = (st.alg_bytes, st.sweep_alg_bytes,
                                               st.gather_alg_bytes)
    bfs.set_counting(False)
    alg_runs = np.array([alg[roots[k % nr]] for k in range(runs)], float)
    sweep_b = np.array([alg_sweep[roots[k % nr]] for k in range(runs)], float)
    gather_b = np.array([alg_gather[roots[k % nr]] for k in range(runs)], float)
    mean_gpu_s = float(np.mean(gpu_ms)) * 1e-3

    # dominant kernel: most device time per BFS over the timed runs
    kern = {
        "k_pred_gather_core": dict(ms=gather_ms, launches=np.ones(runs), bytes=gather_b,
                                   per_unit="24 B per vertex with edges (its original ID, "
                                            "new ID, entry read and written)"),
        "k_bu_sweep": dict(ms=sweep_ms, launches=sweeps.astype(float), bytes=sweep_b,
                           per_unit="per launch: old, multi and new bitmap (24 B per 64 "
                                    "vertices), an 8-B first-neighbour pair per unvisited "
                                    "vertex, 8 B per vertex found, 4 B per list entry, and "
                                    "the deferred rows' offsets and entries up to the first "
                                    "visited neighbour (counted on the device)"),
    }
    dom = max(kern, key=lambda k: float(np.sum(kern[k]["ms"])))
    d = kern[dom]
    nl = float(np.sum(d["launches"]))
    launch_ms = float(np.sum(d["ms"])) / nl
    launch_bytes = float(np.sum(d["bytes"])) / nl
    achieved = launch_bytes / (launch_ms * 1e-3) / 1e9
    traffic = profiled_traffic(args, dom)
    bfs_gbs = float(np.mean(alg_runs)) / mean_gpu_s / 1e9

    # the top-down yardstick for one root's BFS (same component for every
    # root)
    r0 = roots[0]
    bfs.run(r0, args.edgefactor, args.alpha, args.beta)
    if args.levels:
        bfs.set_timing(True)
        bfs.run(r0, args.edgefactor, args.alpha, args.beta)
        st, trace = bfs.stats()
        for t in trace:
            log(f"  level {t['level']:2d} {'TD' if t['top_down'] else 'BU'} "
                f"nq={t['nq_size']:>10d} {t['gpu_ms']:.3f} ms")
        bfs.set_timing(False)
        bfs.run(r0, args.edgefactor, args.alpha, args.beta)
    csr = graph.csr()
    lev = bfs.levels()
    reached = lev != 0xFFFF
    nn = csr["orig2new"][reached].astype(np.int64)
    m_c = int((csr["row_off"][nn + 1] - csr["row_off"][nn]).sum())
    n_comp = int(reached.sum())
    b_yard = 4 * m_c + 8 * (n_comp + 1) + 12 * inf.n
    te0 = te_by_root[r0]
    del csr

    cpu = None
    if args.cpu_seconds > 0:
        try:
            cpu = cpu_baseline(graph, roots, te_by_root, args.cpu_seconds, args)
        except Exception as e:  # the baseline is a report, never the product
            log(f"cpu baseline failed: {e!r}")

    bat = None
    if not args.no_batched:
        bat = batched(g500bfs, graph, bfs, args, teps_hm)
    bfs.destroy()
    graph.free()
    s24 = None
    if not args.no_scale24 and args.scale != 24:
        s24 = scale24_leg(g500bfs, args)
    srv = None
    if not args.no_serve:
        srv = serving(g500bfs, args)

    out = {
        "metric": "harmonic-mean GTEPS @SCALE 26, 1 GPU",
        "value": round(teps_hm / 1e9, 3),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(float(np.sum(times)) / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic (Graph500 Kronecker generator, seed (2,3), generated on device)",
        "config": {
            "workload": f"plain single-root direction-optimizing BFS, Kronecker SCALE {args.scale}, "
                        f"edgefactor {args.edgefactor}, {nr} roots; a step is one BFS per root",
            "scale": args.scale, "edgefactor": args.edgefactor, "roots": nr,
            "roots_timed": nr, "bfs_runs_timed": runs,
            "alpha": args.alpha, "beta": args.beta,
            "n": inf.n, "n_c": inf.n_c, "m": inf.m,
            "component_vertices": n_comp, "component_edges": m_c,
            "te_per_root": te0, "all_roots_valid": bool(all_valid),
            "validated": "every run of the first timed pass, on its own pred"
                         if not args.no_validate else "none",
            "later_passes_repeat_frontiers": bool(traces_repeat),
            "gpu_event_hm_gteps": round(gpu_teps_hm / 1e9, 3),
            "median_bfs_ms": round(float(np.median(times)) * 1e3, 4),
            "mean_bfs_gpu_ms": round(mean_gpu_s * 1e3, 4),
            "mispredicted_steps_per_bfs": round(float(np.mean(mispredicted)), 3),
            "generation_s": round(t_gen, 3), "construction_s": round(t_con, 3),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_unit": "HBM bytes per launch (rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0, "
                            "profiles/bfs_scale26_hbm_traffic.md)",
            "alg_bytes_per_launch": int(launch_bytes),
            "alg_bytes_per_unit": d["per_unit"],
            "launch_ms": round(launch_ms, 4),
            "launches_per_bfs": round(nl / runs, 3),
            "kernel_ms_per_bfs": {k: round(float(np.sum(v["ms"])) / runs, 4)
                                  for k, v in kern.items()},
            "bfs": {"alg_bytes": int(np.mean(alg_runs)),
                    "achieved": round(bfs_gbs, 1),
                    "frac": round(bfs_gbs / HBM_PEAK_GBS, 4),
                    "bytes_per_te": round(float(np.mean(alg_runs)) / te0, 3),
                    "note": "bytes every kernel of one BFS needs (counted on the device) "
                            "over the mean BFS device time"},
            "yardstick": {"bytes": int(b_yard), "bytes_per_te": round(b_yard / te0, 3),
                          "ratio": round(b_yard / mean_gpu_s / 1e9 / HBM_PEAK_GBS, 4),
                          "note": "4 m_c + 8 (n_c+1) + 12 N: every component edge read once, "
                                  "as top-down would; the bottom-up levels skip most of them, "
                                  "so this ratio is not bounded by 1"},
        },
        "cpu_baseline": cpu,
        "scale24": s24,
        "batched": bat,
        "serving": srv,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
