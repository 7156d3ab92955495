"""Benchmark: classic label propagation on the RMAT scale-25 graph, one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--scale 25]

A step is one LP iteration (every vertex takes the most frequent label among
its neighbours, ties to the highest label) over the whole graph, with the
CSR and labels resident in HBM before the timed region starts.  Prints one
JSON line: GTEPS per iteration (directed CSR entries / seconds per step),
the roofline of the dominant kernel measured with HIP events over the timed
steps, and the CPU baseline (the reference's OpenMP CPUlp loop restated,
timed on a bounded sample of the same graph on this host).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from glp_amd import rocprof_summary as RS  # noqa: E402  (reads committed rocprof summaries only)

HBM_PEAK_GBS = 8000.0
# calibrated rocprofv3 FETCH_SIZE + WRITE_SIZE per step of each kernel family
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic_rmat_s25_classic.txt")
METRIC = "GTEPS per LP iteration (RMAT-25, 1 GPU); batched graphs/s vs plain; p50/p99"


def log(*a):
    print("[bench %.1fs]" % (time.time() - T0), *a, file=sys.stderr, flush=True)


T0 = time.time()


def tier_bytes(edges, verts, weighted=False):
    """Algorithmic HBM bytes of one tier launch: per edge the neighbour id and
    its label (plus the weight), per vertex the offset, the kept label and
    the written label."""
    return (12 if weighted else 8) * edges + 12 * verts


def cpu_baseline(off, nbr, target_s, threads):
    """The reference CPU loop (tutorial/CPUlp.cuh:28-92 sync_run_1, restated:
    OpenMP over vertices in id order with its default static schedule, a
    std::map count per vertex, live label reads), timed over a strided sample
    of the same graph's vertices from L0: every k-th vertex, k chosen from a
    1/256 pilot so the timed sample takes about target_s.  With degree-sorted
    ids the static schedule gives the first thread the hubs, in a sample as
    in a whole step, so the sample's rate (its edges over its time) is the
    whole step's.  The same sample with dynamic chunks (a balanced schedule
    the reference does not use) is reported beside it."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    n = len(off) - 1
    deg = np.diff(off.astype(np.int64))
    m = int(deg.sum())

    def timed(verts, balanced):
        lab = np.arange(n, dtype=np.int32)
        return O.cpulp_sample(off, nbr, lab, verts, threads, balanced)

    pilot = np.arange(0, n, 256, dtype=np.int32)
    rate = max(int(deg[pilot].sum()), 1) / max(timed(pilot, False), 1e-6)
    stride = int(min(256, max(1, round(m / (rate * target_s)))))
    verts = np.arange(0, n, stride, dtype=np.int32)
    e = int(deg[verts].sum())
    secs = timed(verts, False)
    # the balanced schedule on a sample of about a third of the time budget
    rate_b = max(int(deg[pilot].sum()), 1) / max(timed(pilot, True), 1e-6)
    stride_b = int(min(256, max(1, round(m / (rate_b * target_s / 3)))))
    verts_b = np.arange(0, n, stride_b, dtype=np.int32)
    secs_b = timed(verts_b, True)
    return {"value": e / secs / 1e9, "unit": "GTEPS", "cores": threads, "kind": "port",
            "sample": "one verbatim CPUlp step from L0 over every %d-th vertex (%d vertices, %d "
                      "edges of %d) of the same graph, static OpenMP schedule as the reference, "
                      "%.2f s on %d threads" % (stride, len(verts), e, m, secs, threads),
            "balanced_schedule_value": int(deg[verts_b].sum()) / secs_b / 1e9,
            "balanced_schedule_sample": "every %d-th vertex, dynamic chunks of 64, %.2f s"
                                        % (stride_b, secs_b)}


def bench_batched(args, G):
    """1,024 planted-partition graphs (n log-uniform in [1000, 262144], one
    community per ~1000 vertices, average degree U[4,64], inter-community
    fraction U[.1,.4]), classic LP, 20 steps: the packed batch replayed from
    a HIP graph against one plain launch sequence per graph."""
    threads = min(len(os.sched_getaffinity(0)), 16)
    log("generating %d mix graphs" % args.graphs)
    graphs = []
    for i in range(args.graphs):
        p = G.mix_params(args.seed, i)
        graphs.append(G.Graph.planted(p["n"], p["k"], p["deg_in"], p["deg_out"], seed=p["seed"],
                                      threads=threads))
    n_tot = sum(g.n for g in graphs)
    m_tot = sum(g.m for g in graphs)
    log("mix: %d graphs, %d vertices, %d edges" % (len(graphs), n_tot, m_tot))
    b = G.Batch(graphs, method=G.CLASSIC)
    tuned_ms = b.autotune(3) if not args.no_autotune else None
    chosen = {k: G.tuning_get(k) for k in ("pack_max_deg", "half_max_deg", "half_lanes", "wave_max_deg",
                                           "block_max_deg", "block_threads", "hub_chunk", "part_keys")}
    log(("autotuned: %s" % chosen) if tuned_ms else "not autotuned")
    steps = args.steps
    # plain: one session per graph, stepped one graph after another
    b.run(1, mode=2)
    b.reset()
    plain_ms = b.run(steps, mode=2)
    plain_labels = [b.labels(i) for i in range(len(graphs))]
    log("plain %.1f ms" % plain_ms)
    # batched: the packed batch, steps replayed from a captured HIP graph
    b.run(1, mode=1)
    b.reset()
    batch_ms = b.run(steps, mode=1)
    same = all(np.array_equal(b.labels(i), plain_labels[i]) for i in range(len(graphs)))
    log("batched %.1f ms, labels identical: %s" % (batch_ms, same))
    b.reset()
    eager_ms = b.run(steps, mode=0)
    gps = len(graphs) / (batch_ms / 1e3)
    plain_gps = len(graphs) / (plain_ms / 1e3)
    return {
        "metric": "batched graphs/s vs plain (1,024-graph planted-partition mix, classic LP, %d steps)" % steps,
        "value": round(gps, 2), "unit": "graphs/s", "n_gpus": 1, "steps": steps, "warmup": 1,
        "ms_per_step": round(batch_ms / steps, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": "batched classic LP over %d planted-partition graphs" % len(graphs),
                   "graphs": len(graphs), "n_total": n_tot, "m_total": m_tot, "seed": args.seed,
                   "autotuned": tuned_ms is not None, "tuning": chosen},
        "plain_graphs_per_s": round(plain_gps, 2), "speedup_vs_plain": round(gps / plain_gps, 3),
        "batched_eager_ms": round(eager_ms, 2), "plain_ms": round(plain_ms, 2),
        "batched_ms": round(batch_ms, 2), "labels_identical": bool(same),
        "gteps_batched": round(m_tot * steps / (batch_ms / 1e3) / 1e9, 3),
    }


def bench_served(args, G):
    """Poisson arrivals of weighted classic / layered (gamma 1/4) LP jobs on
    seeded RMAT graphs (scale U{lo..hi}, edge factor U{8..32}, weights
    U[1,8]), 10 steps each, under a 200 GB HBM budget; p50/p99 latency."""
    rng = np.random.default_rng(args.seed)
    threads = min(len(os.sched_getaffinity(0)), 16)
    lo, hi = args.serve_scales
    specs = []
    for i in range(args.jobs):
        scale = int(rng.integers(lo, hi + 1))
        ef = int(rng.integers(8, 33))
        layered = bool(rng.integers(0, 2))
        specs.append((scale, ef, layered))
    log("generating %d job graphs (scales %d..%d)" % (args.jobs, lo, hi))
    graphs = [G.Graph.rmat(s, ef, seed=args.seed * 100003 + i, wmax=8, threads=threads)
              for i, (s, ef, _) in enumerate(specs)]
    params = [G.make_params(method=G.LAYERED if lay else G.CLASSIC, gamma=(1, 4) if lay else (0, 1),
                            weighted=True) for (_, _, lay) in specs]
    m_tot = sum(g.m for g in graphs)
    sv = G.Server(budget_bytes=200 << 30, max_slots=args.slots)
    gaps = rng.exponential(1.0 / args.rate, size=args.jobs)
    t_start = time.perf_counter()
    due = t_start
    ids = []
    for i, g in enumerate(graphs):
        due += gaps[i]
        delay = due - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        ids.append(sv.submit(g, 10, params=params[i]))
    results = [sv.wait(j, g.n) for j, g in zip(ids, graphs)]
    lat = np.array([st.done_ms - st.submit_ms for _, st in results])
    first = min(st.submit_ms for _, st in results)
    last = max(st.done_ms for _, st in results)
    # each job alone must give the same labels as in the mix
    check = list(range(0, args.jobs, max(1, args.jobs // 16)))
    same = all(np.array_equal(results[i][0], G.run(graphs[i], 10, params=params[i])[0]) for i in check)
    sv.close()
    jobs_s = args.jobs / ((last - first) / 1e3)
    return {
        "metric": "served LP jobs/s with p50/p99 latency (Poisson arrivals, 1 GPU)",
        "value": round(jobs_s, 2), "unit": "jobs/s", "n_gpus": 1, "steps": 10, "warmup": 0,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int32", "data": "synthetic",
        "config": {"workload": "%d weighted classic/layered(gamma=1/4) RMAT jobs, scale U{%d..%d}, "
                               "edge factor U{8..32}, weights U[1,8], 10 steps, Poisson rate %.0f/s, "
                               "200 GB budget, %d slots" % (args.jobs, lo, hi, args.rate, args.slots),
                   "jobs": args.jobs, "m_total": int(m_tot)},
        "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "mean_ms": round(float(lat.mean()), 3), "edges_steps_per_s": round(m_tot * 10 / ((last - first) / 1e3), 1),
        "results_equal_alone": bool(same), "checked_jobs": len(check),
    }


RR_TABLE = 1 << 25    # int32 entries: 128 MiB, the size of RMAT-25's label array
RR_READS = 1 << 28


def bench_random_read(args, G):
    """Calibration of the label gather's access pattern: random 4-byte reads
    from a 128 MiB and a 1 GiB table (library kernel k_calib_read)."""
    out = {"metric": "random 4-byte read calibration", "unit": "G reads/s", "reads_per_launch": RR_READS}
    for entries in (RR_TABLE, 1 << 28):
        ms = G.calib_random_read(entries, RR_READS, 5)
        out["table_%dMiB" % (entries * 4 >> 20)] = {"ms": round(ms, 4),
                                                     "G_reads_per_s": round(RR_READS / (ms / 1e3) / 1e9, 2)}
    out["value"] = out["table_%dMiB" % (RR_TABLE * 4 >> 20)]["G_reads_per_s"]
    return out


def bench_copy(args):
    """Calibration for rocprof byte counters: device-to-device copies of a
    known size (the runtime's copy kernel), so FETCH_SIZE / WRITE_SIZE per
    dispatch can be compared with the bytes actually moved."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    size = args.copy_mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(size)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(size)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(size)) == 0
    reps = 5
    hip.hipDeviceSynchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(size), 3) == 0   # device to device
    hip.hipDeviceSynchronize()
    dt = (time.perf_counter() - t0) / reps
    hip.hipFree(a)
    hip.hipFree(b)
    return {"metric": "copy calibration", "value": round(2 * size / dt / 1e9, 1), "unit": "GB/s",
            "bytes_read_per_copy": size, "bytes_written_per_copy": size, "copies": reps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=25)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graph-replay", action="store_true", help="replay steps from a HIP graph")
    ap.add_argument("--gather-steps", type=int, default=3,
                    help="steps of the separate-gather mode timed after the main run (0: skip)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the layered (same graph) and SLPA (LFR-style 2^20) legs")
    ap.add_argument("--frontier", type=int, default=1,
                    help="1: also time the exact-frontier mode on the same steps (reported apart)")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="tuning override (pack_max_deg, wave_max_deg, block_max_deg, hub_chunk, "
                         "part_keys, part_limit, block_threads); repeatable")
    ap.add_argument("--workload", choices=["rmat", "batched", "served", "copy", "random-read"],
                    default="rmat")
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--graphs", type=int, default=1024)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--jobs", type=int, default=200)
    ap.add_argument("--rate", type=float, default=100.0)
    ap.add_argument("--slots", type=int, default=64)
    ap.add_argument("--serve-scales", type=int, nargs=2, default=[12, 17])
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import glp_amd as G
    tuned = {}
    for kv in args.tune:
        k, v = kv.split("=", 1)
        G.tuning_set(k, int(v))
        tuned[k] = int(v)
    if G.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if args.workload == "batched":
        print(json.dumps(bench_batched(args, G)), flush=True)
        return
    if args.workload == "served":
        print(json.dumps(bench_served(args, G)), flush=True)
        return
    if args.workload == "copy":
        print(json.dumps(bench_copy(args)), flush=True)
        return
    if args.workload == "random-read":
        print(json.dumps(bench_random_read(args, G)), flush=True)
        return
    threads = min(len(os.sched_getaffinity(0)), 16)
    log("generating RMAT scale %d ef %d seed %d (%d threads)" % (args.scale, args.edge_factor,
                                                               args.seed, threads))
    g = G.Graph.rmat(args.scale, args.edge_factor, 0.57, 0.19, 0.19, seed=args.seed,
                     degree_sort=True, threads=threads)
    info = g.info()
    n, m = info["n"], info["m"]
    log("graph n=%d m=%d max_degree=%d" % (n, m, info["max_degree"]))

    sess = G.Session(g, G.make_params(method=G.CLASSIC, profile=True))
    log("session ready (upload %.1f ms, plan %.1f ms)" % (sess.timing()["upload_ms"],
                                                          sess.timing()["plan_ms"]))
    sess.step(args.warmup, graph=args.graph_replay)
    sess.sync()
    base = sess.timing()
    log("warmup done")
    t0 = time.perf_counter()
    sess.step(args.steps, graph=args.graph_replay)
    sess.sync()
    t1 = time.perf_counter()
    tim = sess.timing()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    event_ms = (tim["steps_ms"] - base["steps_ms"]) / args.steps
    gteps = m / (ms_per_step / 1e3) / 1e9
    log("timed %d steps: %.3f ms/step (events %.3f ms/step) -> %.2f GTEPS" %
        (args.steps, ms_per_step, event_ms, gteps))
    changes = sess.changes().tolist()

    # per-kernel durations over the timed steps (events recorded between launches)
    tiers = {}
    for name in ("packed", "wave", "block", "hub", "aux", "gather"):
        launches = tim["tier_launches"][name] - base["tier_launches"][name]
        if launches <= 0:
            continue
        ms = (tim["tier_ms"][name] - base["tier_ms"][name]) / launches
        tiers[name] = {"ms": ms, "edges": tim["tier_edges"].get(name, 0),
                       "vertices": tim["tier_vertices"].get(name, 0)}
    kernel_names = {"packed": "k_packed", "wave": "k_wave", "block": "k_block (all degree classes)",
                    "hub": "k_hub_agg", "aux": "k_hub_scan+k_hub_scatter+k_hub_part+k_hub_final"}
    families = {"packed": ["k_packed"], "wave": ["k_wave"], "block": ["k_block"],
                "hub": ["k_hub_agg"]}
    dom = max((k for k in tiers if k in families), key=lambda k: tiers[k]["ms"], default=None)
    roofline = None
    wl_key = "rmat-s%d-ef%d-seed%d-classic" % (args.scale, args.edge_factor, args.seed)
    if dom is not None:
        b = tier_bytes(tiers[dom]["edges"], tiers[dom]["vertices"])
        achieved = b / (tiers[dom]["ms"] / 1e3) / 1e9
        traffic = RS.traffic_for(TRAFFIC_FILE, wl_key, "main", families[dom])
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": None if traffic is None else int(traffic),
                    "kernel": kernel_names[dom], "kernel_ms": round(tiers[dom]["ms"], 4),
                    "algorithmic_bytes": int(b)}
        # the ceiling of the label gather's access pattern on this device:
        # random 4-byte reads of a label-array-sized table (each one a 128-byte
        # line fetch when it misses L2)
        rr_ms = G.calib_random_read(RR_TABLE, RR_READS, 5)
        rr = RR_READS / (rr_ms / 1e3) / 1e9
        roofline["random_read_ceiling"] = {"G_reads_per_s": round(rr, 2),
                                           "GBs_in_128B_lines": round(rr * 128, 1),
                                           "table_MiB": RR_TABLE * 4 >> 20}
        if traffic is not None:
            # measured HBM bytes (rocprof, calibrated) over this run's duration
            tg = traffic / (tiers[dom]["ms"] / 1e3) / 1e9
            roofline["traffic_GBs"] = round(tg, 1)
            roofline["traffic_frac"] = round(tg / HBM_PEAK_GBS, 4)
            roofline["traffic_frac_of_random_read_ceiling"] = round(tg / (rr * 128), 4)
    whole = tier_bytes(m, n)
    step_roofline = {"algorithmic_bytes": int(whole),
                     "achieved_GBs": round(whole / (ms_per_step / 1e3) / 1e9, 1),
                     "frac": round(whole / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
    dense_labels = sess.labels()
    sess.free()

    # the exact frontier (classic only): same steps, vertices whose neighbours
    # all kept their labels are skipped; labels must equal the dense run's
    frontier = None
    if args.frontier:
        fs = G.Session(g, G.make_params(method=G.CLASSIC, frontier=True))
        fs.step(args.warmup)
        fs.sync()
        t0 = time.perf_counter()
        fs.step(args.steps)
        fs.sync()
        fms = (time.perf_counter() - t0) * 1e3 / args.steps
        same = bool(np.array_equal(fs.labels(), dense_labels))
        frontier = {"ms_per_step": round(fms, 4), "edges_per_s_equiv_G": round(m / (fms / 1e3) / 1e9, 2),
                    "labels_identical_to_dense": same}
        log("frontier %.3f ms/step, labels identical: %s" % (fms, same))
        fs.free()
        # the separate gather with the frontier: the labels of sources that
        # changed in the previous step are pushed to their in-edges once few
        # change (frontier 2: the two-step lookahead, changes against two
        # steps back into alternating adj buffers)
        for mode, key in ((1, "gather"), (2, "gather_lookahead")):
            fg = G.Session(g, G.make_params(method=G.CLASSIC, frontier=mode, gather=True))
            fg.step(args.warmup)
            fg.sync()
            t0 = time.perf_counter()
            fg.step(args.steps)
            fg.sync()
            gms_f = (time.perf_counter() - t0) * 1e3 / args.steps
            same_g = bool(np.array_equal(fg.labels(), dense_labels))
            frontier[key + "_ms_per_step"] = round(gms_f, 4)
            frontier[key + "_labels_identical_to_dense"] = same_g
            log("frontier %s %.3f ms/step, labels identical: %s" % (key, gms_f, same_g))
            fg.free()

    # the other methods of the same configurations: layered LP on this graph
    # (gamma 0, 1/4, 1; same warmup and timed steps) and SLPA (T=20, r=0.1)
    # on the LFR-style graph of 2^20 vertices
    extra = None
    if not args.no_extra:
        extra = {"layered": {}}
        for gn, gd in ((0, 1), (1, 4), (1, 1)):
            ls = G.Session(g, G.make_params(method=G.LAYERED, gamma=(gn, gd)))
            ls.step(args.warmup)
            ls.sync()
            t0 = time.perf_counter()
            ls.step(args.steps)
            ls.sync()
            lms = (time.perf_counter() - t0) * 1e3 / args.steps
            extra["layered"]["%d/%d" % (gn, gd)] = {"ms_per_step": round(lms, 4),
                                                    "GTEPS": round(m / (lms / 1e3) / 1e9, 3)}
            log("layered gamma=%d/%d: %.3f ms/step" % (gn, gd, lms))
            ls.free()
        lg = G.Graph.lfr(1 << 20, seed=args.seed)
        ss = G.Session(lg, G.make_params(method=G.SLPA, slpa_T=20, slpa_r=(1, 10), seed=args.seed))
        t0 = time.perf_counter()
        ss.step(20)
        ss.sync()
        sms = (time.perf_counter() - t0) * 1e3 / 20
        t1 = time.perf_counter()
        dom, ns, _ = ss.slpa_post(8)
        pms = (time.perf_counter() - t1) * 1e3
        extra["slpa"] = {"graph": "LFR-style n=2^20 (tau1 2.5, degrees [8,512], tau2 1.5, sizes "
                                  "[20,5000], mu 0.3), m=%d" % lg.m,
                         "steps": 20, "ms_per_step": round(sms, 4),
                         "GTEPS": round(lg.m / (sms / 1e3) / 1e9, 3), "post_ms": round(pms, 2),
                         "mean_communities_per_vertex": round(float(ns.mean()), 3)}
        log("slpa LFR-1M: %.3f ms/step, post %.1f ms" % (sms, pms))
        ss.free()
        lg.free()

    # the separate gather pass (LoadNeighbor as its own kernel): 12 B per edge
    gather = None
    if args.gather_steps > 0:
        gs = G.Session(g, G.make_params(method=G.CLASSIC, gather=True, profile=True))
        gs.step(1)
        gs.sync()
        g0 = gs.timing()
        gs.step(args.gather_steps)
        gs.sync()
        g1 = gs.timing()
        n_l = g1["tier_launches"]["gather"] - g0["tier_launches"]["gather"]
        gms = (g1["tier_ms"]["gather"] - g0["tier_ms"]["gather"]) / max(n_l, 1)
        gt = RS.traffic_for(TRAFFIC_FILE, wl_key, "gather", ["k_gather"])
        gather = {"kernel": "k_gather", "kernel_ms": round(gms, 4), "algorithmic_bytes": 12 * m,
                  "traffic": None if gt is None else int(gt),
                  "achieved_GBs": round(12 * m / (gms / 1e3) / 1e9, 1),
                  "frac": round(12 * m / (gms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                  "step_ms": round((g1["steps_ms"] - g0["steps_ms"]) / args.gather_steps, 4)}
        if gt is not None:
            gather["traffic_GBs"] = round(gt / (gms / 1e3) / 1e9, 1)
            gather["traffic_frac"] = round(gt / (gms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)
        log("gather kernel %.3f ms -> %.1f GB/s" % (gms, gather["achieved_GBs"]))
        gs.free()

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (%d threads, ~%.0f s)" % (threads, args.cpu_seconds))
        off, nbr, _ = g.csr()
        cpu = cpu_baseline(off, nbr, args.cpu_seconds, threads)
        log("cpu baseline %.4f GTEPS" % cpu["value"])

    out = {
        "metric": METRIC, "value": round(gteps, 3), "unit": "GTEPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "classic LP (method 1), RMAT scale %d, edge factor %d, "
                               "(a,b,c,d)=(0.57,0.19,0.19,0.05), seed %d, symmetrized, deduplicated, "
                               "degree-sorted ids, L0[v]=v" % (args.scale, args.edge_factor, args.seed),
                   "n": n, "m": m, "max_degree": info["max_degree"],
                   "graph_replay": bool(args.graph_replay), "tuning": tuned or "default"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "step_roofline": step_roofline,
        "gather": gather,
        "frontier": frontier,
        "other_methods": extra,
        "tiers_ms": {k: round(v["ms"], 4) for k, v in tiers.items()},
        "changes_per_step": changes[-args.steps:],
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
