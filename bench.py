"""Benchmark: label propagation on one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--scale 25] [...]

The headline is classic LP on the RMAT scale-25 graph: a step is one LP
iteration (every vertex takes the most frequent label among its neighbours,
ties to the highest label) over the whole graph, with the CSR and labels
resident in HBM before the timed region starts; `value` is GTEPS per
iteration (directed CSR entries / seconds per step) over K steps after W
warmup steps.  The same JSON line carries the rest of the metric:
  * from_L0_10_steps: the first ten steps from L0[v] = v (the RMAT-25
    classic configuration as specified), and layered LP's (gamma 0, 1/4, 1);
  * parity_sample: steps 1 and 10 from L0 checked against the oracle on
    every hub and a strided sample of every degree tier of RMAT-25;
  * batched: 1,024 planted-partition graphs, 20 steps, the packed batch
    replayed from a HIP graph against one plain launch sequence per graph;
  * served: 20,000 Poisson-arrival jobs (weighted classic / layered 1/4,
    RMAT scale 12-22, edge factor 8-32, 10 steps) under a 200 GB budget,
    jobs/s and p50/p99 latency;
  * the roofline of the dominant kernel (HIP events over the timed steps)
    and the CPU baseline (the reference's OpenMP loops restated, timed on
    this host on bounded samples).
"""
import argparse
import json
import os
import platform
import subprocess
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from glp_amd import rocprof_summary as RS  # noqa: E402  (reads committed rocprof summaries only)

HBM_PEAK_GBS = 8000.0
# calibrated rocprofv3 FETCH_SIZE + WRITE_SIZE per step of each kernel family
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_traffic_rmat_s25_classic.txt")
TRAFFIC_SOURCE = ("profiles/pmc_traffic_rmat_s25_classic.txt: rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 "
                  "per step (copy-kernel calibration), divided by this run's kernel time")
TRAFFIC_FILE_LAYERED = os.path.join(ROOT, "profiles", "pmc_traffic_rmat_s25_layered_g1_4.txt")
METRIC = "GTEPS per LP iteration (RMAT-25, 1 GPU); batched graphs/s vs plain; p50/p99"
TIER_FAMILIES = {"packed": ["k_packed"], "wave": ["k_wave", "k_half"], "block": ["k_block"],
                 "hub": ["k_hub_agg"]}
KERNEL_NAMES = {"packed": "k_packed", "wave": "k_wave + k_half", "block": "k_block (all degree classes)",
                "hub": "k_hub_agg", "aux": "k_hub_part+k_hub_final"}


def log(*a):
    print("[bench %.1fs]" % (time.time() - T0), *a, file=sys.stderr, flush=True)


T0 = time.time()


def host_threads():
    """CPU threads this process may use: its affinity, at most 16 (one GPU's
    share of the host)."""
    return min(len(os.sched_getaffinity(0)), 16)


def tier_bytes(edges, verts, per_edge=8, per_vertex=12):
    """Algorithmic HBM bytes of one tier launch: per edge the neighbour id and
    its label (classic: 8 B; layered adds the volume of its label: 12 B),
    per vertex the offset, the kept label and the written label (layered
    adds the volume increment: 16 B)."""
    return per_edge * edges + per_vertex * verts


def tiers_over(base, tim, steps):
    """Per-tier HIP-event time per launch over the steps since `base`."""
    tiers = {}
    for name in ("packed", "wave", "block", "hub", "aux", "gather"):
        launches = tim["tier_launches"][name] - base["tier_launches"][name]
        if launches <= 0:
            continue
        tiers[name] = {"ms": (tim["tier_ms"][name] - base["tier_ms"][name]) / launches,
                       "edges": tim["tier_edges"].get(name, 0), "vertices": tim["tier_vertices"].get(name, 0)}
    return tiers


def dominant_roofline(tiers, per_edge, per_vertex):
    dom = max((k for k in tiers if k in TIER_FAMILIES), key=lambda k: tiers[k]["ms"], default=None)
    if dom is None:
        return None, None
    b = tier_bytes(tiers[dom]["edges"], tiers[dom]["vertices"], per_edge, per_vertex)
    achieved = b / (tiers[dom]["ms"] / 1e3) / 1e9
    return dom, {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": round(achieved / HBM_PEAK_GBS, 4), "kernel": KERNEL_NAMES[dom],
                 "kernel_ms": round(tiers[dom]["ms"], 4), "algorithmic_bytes": int(b),
                 "bytes_per_edge": per_edge, "bytes_per_vertex": per_vertex}


# ------------------------------------------------------------ CPU baseline ----

def cpu_model():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return platform.processor() or "unknown"


def cpu_c2_sample(O, off, nbr, target_s, threads):
    """One verbatim CPUlp step (tutorial/CPUlp.cuh:28-92 sync_run_1, restated:
    OpenMP over vertices in id order with its default static schedule, a
    std::map count per vertex, live label reads) from L0 over every k-th
    vertex of the graph, k chosen from a 1/256 pilot so the sample takes
    about target_s.  With degree-sorted ids the static schedule gives the
    first thread the hubs, in a sample as in a whole step, so the sample's
    rate (its edges over its time) is the whole step's."""
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
    out = {"value": e / secs / 1e9, "unit": "GTEPS", "threads": threads,
           "sample": "one verbatim CPUlp step from L0 over every %d-th vertex (%d vertices, %d edges of %d), "
                     "static OpenMP schedule as the reference, %.2f s" % (stride, len(verts), e, m, secs)}
    # the same sample with dynamic chunks (a balanced schedule the reference
    # does not use), on about a third of the time budget
    rate_b = max(int(deg[pilot].sum()), 1) / max(timed(pilot, True), 1e-6)
    stride_b = int(min(256, max(1, round(m / (rate_b * target_s / 3)))))
    verts_b = np.arange(0, n, stride_b, dtype=np.int32)
    secs_b = timed(verts_b, True)
    out["balanced_schedule_value"] = int(deg[verts_b].sum()) / secs_b / 1e9
    out["balanced_schedule_sample"] = "every %d-th vertex, dynamic chunks of 64, %.2f s" % (stride_b, secs_b)
    return out


def cpu_layered_sample(O, off, nbr, gamma, target_s, threads):
    """One step of the reference's layered_run (tutorial/CPUlp.cuh:169-262,
    restated: std::map count, running score (1+gamma)C - gamma V(l), live
    label reads) from L0 on a strided sample, as cpu_c2_sample."""
    n = len(off) - 1
    deg = np.diff(off.astype(np.int64))
    m = int(deg.sum())
    p, q = gamma

    def timed(verts):
        lab = np.arange(n, dtype=np.int32)
        vr = np.ones(n, dtype=np.int32)
        vw = np.zeros(n, dtype=np.int32)
        return O.cpulp_layered_sample(off, nbr, lab, vr, vw, verts, threads, p, q)

    pilot = np.arange(0, n, 256, dtype=np.int32)
    rate = max(int(deg[pilot].sum()), 1) / max(timed(pilot), 1e-6)
    stride = int(min(256, max(1, round(m / (rate * target_s)))))
    verts = np.arange(0, n, stride, dtype=np.int32)
    e = int(deg[verts].sum())
    secs = timed(verts)
    return {"value": e / secs / 1e9, "unit": "GTEPS", "threads": threads,
            "sample": "every %d-th vertex (%d edges), %.2f s" % (stride, e, secs)}


def cpu_c1(O, G, args, threads):
    """The 65,536-vertex planted partition (64 communities of 1,024, expected
    12 intra + 4 inter neighbours), all 20 steps of the CPUlp loop in its
    verbatim mode (live reads, first maximum; timing only) and its sync mode
    (RS tie rule), the sync labels compared with the GPU's."""
    g = G.Graph.planted(65536, 64, 12.0, 4.0, seed=args.seed)
    off, nbr, _ = g.csr()
    m = len(nbr)
    gpu_labels, _ = G.run(g, 20)
    out = {"graph": "planted partition n=65536, 64 communities, m=%d" % m, "steps": 20}
    for t in sorted({20, threads}):
        _, secs_v = O.cpulp_run(off, nbr, 20, t, 0)
        sync_labels, secs_s = O.cpulp_run(off, nbr, 20, t, 1)
        out["threads_%d" % t] = {"verbatim_ms_per_step": round(secs_v * 1e3 / 20, 3),
                                 "verbatim_GTEPS": round(m * 20 / secs_v / 1e9, 4),
                                 "sync_ms_per_step": round(secs_s * 1e3 / 20, 3),
                                 "sync_labels_equal_gpu": bool(np.array_equal(sync_labels, gpu_labels))}
    return out


def cpu_baseline(O, G, args, off, nbr):
    threads = host_threads()
    out = {"cores": threads, "kind": "port", "cpu_model": cpu_model(), "nproc": os.cpu_count(),
           "affinity": len(os.sched_getaffinity(0))}
    log("cpu baseline: C2 sample, %d threads (~%.0f s)" % (threads, args.cpu_seconds))
    main = cpu_c2_sample(O, off, nbr, args.cpu_seconds, threads)
    out.update({"value": main["value"], "unit": "GTEPS", "sample": main["sample"],
                "balanced_schedule_value": main["balanced_schedule_value"],
                "balanced_schedule_sample": main["balanced_schedule_sample"]})
    # tutorial/CPUlp.cuh:50 hard-codes num_threads(20)
    log("cpu baseline: C2 sample, 20 threads (~%.0f s)" % (args.cpu_seconds / 2))
    r20 = cpu_c2_sample(O, off, nbr, args.cpu_seconds / 2, 20)
    out["reference_threads_20"] = {"value": r20["value"], "sample": r20["sample"]}
    log("cpu baseline: C1 20 steps, verbatim and sync")
    out["c1"] = cpu_c1(O, G, args, threads)
    out["layered"] = {}
    for gam in ((0, 1), (1, 4), (1, 1)):
        log("cpu baseline: layered gamma=%d/%d sample" % gam)
        out["layered"]["%d/%d" % gam] = cpu_layered_sample(O, off, nbr, gam, args.cpu_seconds / 3, threads)
    return out


# ------------------------------------------------------------ RMAT-25 legs ----

def rmat_main(args, G, g, n, m):
    """The headline: W warmup then K timed classic steps, per-tier events."""
    sess = G.Session(g, G.make_params(method=G.CLASSIC, profile=True))
    log("session ready (upload %.1f ms, plan %.1f ms)" % (sess.timing()["upload_ms"], sess.timing()["plan_ms"]))
    sess.step(args.warmup, graph=args.graph_replay)
    sess.sync()
    base = sess.timing()
    t0 = time.perf_counter()
    sess.step(args.steps, graph=args.graph_replay)
    sess.sync()
    t1 = time.perf_counter()
    tim = sess.timing()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    event_ms = (tim["steps_ms"] - base["steps_ms"]) / args.steps
    log("timed %d steps: %.3f ms/step (events %.3f ms/step) -> %.2f GTEPS" %
        (args.steps, ms_per_step, event_ms, m / (ms_per_step / 1e3) / 1e9))
    tiers = tiers_over(base, tim, args.steps)
    changes = sess.changes().tolist()
    labels = sess.labels()
    sess.free()
    return ms_per_step, tiers, changes, labels


def layered_legs(args, G, g, n, m):
    """Layered LP on the same graph, gamma 0, 1/4, 1 (dyadic): the first ten
    steps from L0 and the same W + K steps as the headline, with the
    dominant kernel's roofline (12 B per edge, 16 B per vertex)."""
    out = {}
    for gn, gd in ((0, 1), (1, 4), (1, 1)):
        ls = G.Session(g, G.make_params(method=G.LAYERED, gamma=(gn, gd), profile=True))
        t0 = time.perf_counter()
        ls.step(10)
        ls.sync()
        l0_ms = (time.perf_counter() - t0) * 1e3 / 10
        ls.reset()
        ls.step(args.warmup)
        ls.sync()
        base = ls.timing()
        t0 = time.perf_counter()
        ls.step(args.steps)
        ls.sync()
        lms = (time.perf_counter() - t0) * 1e3 / args.steps
        tiers = tiers_over(base, ls.timing(), args.steps)
        dom, roof = dominant_roofline(tiers, 12, 16)
        if roof is not None and (gn, gd) == (1, 4):
            # rocprof FETCH_SIZE + WRITE_SIZE of the same kernels at gamma 1/4
            t = RS.traffic_for(TRAFFIC_FILE_LAYERED, "rmat-s%d-ef%d-seed%d-layered-1/4" % (
                args.scale, args.edge_factor, args.seed), "main", TIER_FAMILIES[dom])
            if t is not None:
                roof["traffic"] = int(t)
                roof["traffic_GBs"] = round(t / (tiers[dom]["ms"] / 1e3) / 1e9, 1)
                roof["traffic_source"] = "profiles/pmc_traffic_rmat_s25_layered_g1_4.txt (calibrated as the classic file)"
        out["%d/%d" % (gn, gd)] = {
            "ms_per_step": round(lms, 4), "GTEPS": round(m / (lms / 1e3) / 1e9, 3),
            "from_L0_10_steps_ms_per_step": round(l0_ms, 4),
            "from_L0_10_steps_GTEPS": round(m / (l0_ms / 1e3) / 1e9, 3),
            "roofline": roof, "tiers_ms": {k: round(v["ms"], 4) for k, v in tiers.items()},
            "changes_last_step": int(ls.changes()[-1])}
        log("layered gamma=%d/%d: %.3f ms/step (from L0: %.3f)" % (gn, gd, lms, l0_ms))
        ls.free()
    return out


def counting_legs(args, G, g, m):
    """Count-min (3 x 200), frequent and buffer (16 slots) on the same graph:
    three steps from L0, timed (labels stay diverse: the dearest steps)."""
    out = {}
    for name, kw in (("countmin", {"method": G.COUNTMIN, "cm_rows": 3, "cm_width": 200}),
                     ("frequent", {"method": G.FREQUENT, "table_cap": 16}),
                     ("buffer", {"method": G.BUFFER, "table_cap": 16})):
        s = G.Session(g, G.make_params(**kw))
        s.step(1)
        s.sync()
        s.reset()
        t0 = time.perf_counter()
        s.step(3)
        s.sync()
        ms = (time.perf_counter() - t0) * 1e3 / 3
        out[name] = {"ms_per_step": round(ms, 3), "GTEPS": round(m / (ms / 1e3) / 1e9, 3),
                     "steps": "1-3 from L0",
                     "shape": {k: v for k, v in kw.items() if k != "method"}}
        log("%s: %.3f ms/step" % (name, ms))
        s.free()
    return out


def from_l0_and_parity(args, G, O, g, off, nbr, n, m):
    """The first ten classic steps from L0, timed, and the oracle's check of
    steps 1 and 10 on every hub (degree > the block tier's bound) and a
    strided sample of every other tier, each from the GPU's labels of the
    step before."""
    s = G.Session(g, G.make_params(method=G.CLASSIC))
    t0 = time.perf_counter()
    s.step(10)
    s.sync()
    l0_ms = (time.perf_counter() - t0) * 1e3 / 10
    changes = s.changes().tolist()
    s.reset()
    deg = np.diff(off.astype(np.int64))
    bmax = G.tuning_get("block_max_deg")
    tiers = {"hub": deg > bmax, "block": (deg > G.tuning_get("wave_max_deg")) & (deg <= bmax),
             "wave": (deg > G.tuning_get("pack_max_deg")) & (deg <= G.tuning_get("wave_max_deg")),
             "packed": (deg > 0) & (deg <= G.tuning_get("pack_max_deg")), "isolated": deg == 0}
    verts = []
    per_tier = {}
    for name, mask in tiers.items():
        ids = np.flatnonzero(mask).astype(np.int32)
        if name != "hub":
            ids = ids[::max(1, len(ids) // args.parity_per_tier)]
        per_tier[name] = int(len(ids))
        verts.append(ids)
    verts = np.concatenate(verts)
    checks = {}
    prev = np.arange(n, dtype=np.int32)
    threads = host_threads()
    done = 0
    for step in (1, 10):
        if step - 1 > done:
            s.step(step - 1 - done)
            prev = s.labels()
            done = step - 1
        s.step(1)
        got = s.labels()
        done = step
        t0 = time.perf_counter()
        want = O.classic_subset(off, nbr, prev, verts, threads=threads)
        bad = int(np.count_nonzero(got[verts] != want))
        checks["step_%d" % step] = {"mismatches": bad, "oracle_s": round(time.perf_counter() - t0, 2)}
        log("parity sample step %d: %d vertices, %d mismatches" % (step, len(verts), bad))
        prev = got
    s.free()
    sample_edges = int(deg[verts].sum())
    return ({"ms_per_step": round(l0_ms, 4), "GTEPS": round(m / (l0_ms / 1e3) / 1e9, 3),
             "changes_per_step": changes[-10:]},
            {"vertices": int(len(verts)), "edges": sample_edges, "per_tier": per_tier,
             "mismatches": sum(c["mismatches"] for c in checks.values()), "steps": checks,
             "rule": "every hub (degree > %d) and about %d vertices of each other tier; the GPU's labels of "
                     "step t against the oracle's step from the GPU's labels of step t-1" %
                     (bmax, args.parity_per_tier)})


def gather_leg(args, G, g, m, wl_key):
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
    out = {"kernel": "k_gather", "kernel_ms": round(gms, 4), "algorithmic_bytes": 12 * m,
           "traffic": None if gt is None else int(gt),
           "achieved_GBs": round(12 * m / (gms / 1e3) / 1e9, 1),
           "frac": round(12 * m / (gms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
           "step_ms": round((g1["steps_ms"] - g0["steps_ms"]) / args.gather_steps, 4)}
    if gt is not None:
        out["traffic_GBs"] = round(gt / (gms / 1e3) / 1e9, 1)
        out["traffic_frac"] = round(gt / (gms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)
        out["traffic_source"] = TRAFFIC_SOURCE
    log("gather kernel %.3f ms -> %.1f GB/s" % (gms, out["achieved_GBs"]))
    gs.free()
    return out


def frontier_leg(args, G, g, m, dense_labels):
    """The exact frontier (classic only): same steps, vertices whose
    neighbours all kept their labels are skipped; labels must equal the
    dense run's.  Then the separate gather with the frontier (labels of the
    previous step's changed sources pushed to their in-edges; mode 2 adds
    the two-step lookahead)."""
    fs = G.Session(g, G.make_params(method=G.CLASSIC, frontier=True))
    fs.step(args.warmup)
    fs.sync()
    t0 = time.perf_counter()
    fs.step(args.steps)
    fs.sync()
    fms = (time.perf_counter() - t0) * 1e3 / args.steps
    same = bool(np.array_equal(fs.labels(), dense_labels))
    out = {"ms_per_step": round(fms, 4), "edges_per_s_equiv_G": round(m / (fms / 1e3) / 1e9, 2),
           "labels_identical_to_dense": same}
    log("frontier %.3f ms/step, labels identical: %s" % (fms, same))
    fs.free()
    for mode, key in ((1, "gather"), (2, "gather_lookahead")):
        fg = G.Session(g, G.make_params(method=G.CLASSIC, frontier=mode, gather=True))
        fg.step(args.warmup)
        fg.sync()
        t0 = time.perf_counter()
        fg.step(args.steps)
        fg.sync()
        gms_f = (time.perf_counter() - t0) * 1e3 / args.steps
        out[key + "_ms_per_step"] = round(gms_f, 4)
        out[key + "_labels_identical_to_dense"] = bool(np.array_equal(fg.labels(), dense_labels))
        log("frontier %s %.3f ms/step" % (key, gms_f))
        fg.free()
    return out


def slpa_leg(args, G):
    lg = G.Graph.lfr(1 << 20, seed=args.seed)
    ss = G.Session(lg, G.make_params(method=G.SLPA, slpa_T=20, slpa_r=(1, 10), seed=args.seed))
    ss.step(20)  # untimed pass: the kernels' first launches
    ss.sync()
    ss.reset()
    t0 = time.perf_counter()
    ss.step(20)
    ss.sync()
    sms = (time.perf_counter() - t0) * 1e3 / 20
    t1 = time.perf_counter()
    _, ns, _ = ss.slpa_post(8)
    pms = (time.perf_counter() - t1) * 1e3
    out = {"graph": "LFR-style n=2^20 (tau1 2.5, degrees [8,512], tau2 1.5, sizes [20,5000], mu 0.3), m=%d" % lg.m,
           "steps": 20, "ms_per_step": round(sms, 4), "GTEPS": round(lg.m / (sms / 1e3) / 1e9, 3),
           "post_ms": round(pms, 2), "mean_communities_per_vertex": round(float(ns.mean()), 3),
           "algorithmic_bytes_per_step": 8 * lg.m + 16 * lg.n,
           "step_roofline_frac": round((8 * lg.m + 16 * lg.n) / (sms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
    log("slpa LFR-1M: %.3f ms/step, post %.1f ms" % (sms, pms))
    ss.free()
    lg.free()
    return out


# ----------------------------------------------------------------- batched ----

def bench_batched(args, G):
    """1,024 planted-partition graphs (n log-uniform in [1000, 262144], one
    community per ~1000 vertices, average degree U[4,64], inter-community
    fraction U[.1,.4]), classic LP, 20 steps: the packed batch replayed from
    a HIP graph against one plain launch sequence per graph."""
    threads = host_threads()
    steps = args.batch_steps
    log("generating %d mix graphs" % args.graphs)
    graphs = []
    for i in range(args.graphs):
        p = G.mix_params(args.seed, i)
        graphs.append(G.Graph.planted(p["n"], p["k"], p["deg_in"], p["deg_out"], seed=p["seed"],
                                      threads=threads))
    n_tot = sum(g.n for g in graphs)
    m_tot = sum(g.m for g in graphs)
    log("mix: %d graphs, %d vertices, %d edges" % (len(graphs), n_tot, m_tot))
    saved = {k: G.tuning_get(k) for k in G.TUNING_KEYS}
    b = G.Batch(graphs, method=G.CLASSIC)
    tuned_ms = b.autotune(3) if not args.no_autotune else None
    chosen = {k: G.tuning_get(k) for k in ("pack_max_deg", "half_max_deg", "half_lanes", "wave_max_deg",
                                           "block_max_deg", "block_threads", "hub_chunk", "part_keys",
                                           "fused_small")}
    log(("autotuned: %s" % chosen) if tuned_ms else "not autotuned")
    plain_ms, plain_labels = None, None
    if not args.batched_only:
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
    same = None
    if plain_labels is not None:
        same = all(np.array_equal(b.labels(i), plain_labels[i]) for i in range(len(graphs)))
    log("batched %.1f ms, labels identical: %s" % (batch_ms, same))
    b.reset()
    eager_ms = b.run(steps, mode=0)
    b.free()
    for k, v in saved.items():
        G.tuning_set(k, v)
    gps = len(graphs) / (batch_ms / 1e3)
    plain_gps = len(graphs) / (plain_ms / 1e3) if plain_ms else None
    return {
        "metric": "batched graphs/s vs plain (1,024-graph planted-partition mix, classic LP, %d steps)" % steps,
        "value": round(gps, 2), "unit": "graphs/s", "steps": steps,
        "ms_per_step": round(batch_ms / steps, 4),
        "config": {"workload": "batched classic LP over %d planted-partition graphs (n log-uniform in "
                               "[1000, 262144], ~1000 vertices per community, degree U[4,64], "
                               "inter fraction U[.1,.4]), seed %d" % (len(graphs), args.seed),
                   "graphs": len(graphs), "n_total": n_tot, "m_total": m_tot,
                   "autotuned": tuned_ms is not None, "tuning": chosen},
        "plain_graphs_per_s": round(plain_gps, 2) if plain_gps else None,
        "speedup_vs_plain": round(gps / plain_gps, 3) if plain_gps else None,
        "batched_eager_ms": round(eager_ms, 2), "plain_ms": round(plain_ms, 2) if plain_ms else None,
        "batched_ms": round(batch_ms, 2), "labels_identical": same,
        "gteps_batched": round(m_tot * steps / (batch_ms / 1e3) / 1e9, 3),
        # the union step's algorithmic bytes (8 B per edge, 12 B per vertex)
        # over its replayed step time
        "step_roofline": {"algorithmic_bytes": int(8 * m_tot + 12 * n_tot),
                          "achieved_GBs": round((8 * m_tot + 12 * n_tot) / (batch_ms / steps / 1e3) / 1e9, 1),
                          "frac": round((8 * m_tot + 12 * n_tot) / (batch_ms / steps / 1e3) / 1e9 / HBM_PEAK_GBS, 4)},
    }


# ------------------------------------------------------------------ served ----

def bench_served(args, G, O):
    """Poisson arrivals of weighted classic / layered (gamma 1/4) LP jobs,
    10 steps each, under a 200 GB HBM budget.  Job i draws its scale from
    U{12..22}, its edge factor from U{8..32} and its method 50/50; its graph
    is the seeded RMAT graph of that (scale, edge factor) (a = .57, b = c =
    .19, weights U[1,8]), one per pair, all 275 generated in HBM before the
    stream starts and resident through it.  Latency runs from submit to the
    labels on the host."""
    lo, hi = args.serve_scales
    rng = np.random.default_rng(args.seed)
    sv = G.Server(budget_bytes=200 << 30, max_slots=args.slots)
    t0 = time.perf_counter()
    gid = {}
    for scale in range(lo, hi + 1):
        for ef in range(8, 33):
            gid[(scale, ef)] = sv.add_rmat(scale, ef, seed=scale * 1000 + ef, wmax=8)
    pool_s = time.perf_counter() - t0
    info = {k: sv.graph_info(v) for k, v in gid.items()}
    pool_bytes = sum(v["bytes"] for v in info.values())
    log("graph pool: %d graphs, %.1f GB, %.1f s" % (len(gid), pool_bytes / 1e9, pool_s))
    spec = [(int(rng.integers(lo, hi + 1)), int(rng.integers(8, 33)), bool(rng.integers(0, 2)))
            for _ in range(args.jobs)]
    p_lay = G.make_params(method=G.LAYERED, gamma=(1, 4), weighted=True)
    p_cls = G.make_params(method=G.CLASSIC, weighted=True)
    gaps = rng.exponential(1.0 / args.rate, size=args.jobs)
    check = set(range(0, args.jobs, max(1, args.jobs // args.serve_checks)))
    ids = [None] * args.jobs
    stats = [None] * args.jobs
    kept = {}
    submitted = threading.Semaphore(0)

    def waiter():
        # waits for the jobs in submission order and forgets each one, so the
        # host holds the labels of the checked jobs only
        for i in range(args.jobs):
            submitted.acquire()
            n = info[spec[i][:2]]["n"]
            labels, st = sv.wait(ids[i], n)
            stats[i] = (st.submit_ms, st.admit_ms, st.start_ms, st.done_ms, st.state)
            if i in check:
                kept[i] = labels
            sv.release(ids[i])

    th = threading.Thread(target=waiter)
    th.start()
    t_start = time.perf_counter()
    due = t_start
    for i, (scale, ef, lay) in enumerate(spec):
        due += gaps[i]
        delay = due - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        ids[i] = sv.submit_resident(gid[(scale, ef)], 10, params=p_lay if lay else p_cls)
        submitted.release()
        if (i + 1) % 5000 == 0:
            log("served: %d jobs submitted" % (i + 1))
    th.join()
    wall = time.perf_counter() - t_start
    st = np.array(stats)
    ok = bool((st[:, 4] == 2).all())
    lat = st[:, 3] - st[:, 0]
    queue = st[:, 1] - st[:, 0]
    start = st[:, 2] - st[:, 1]
    run = st[:, 3] - st[:, 2]
    span_s = (st[:, 3].max() - st[:, 0].min()) / 1e3
    m_tot = sum(info[s[:2]]["m"] for s in spec)
    # each checked job alone (on the idle server) and, for graphs of scale
    # <= 14, against the oracle
    same_alone, oracle_ok, oracle_n = True, True, 0
    for i in sorted(check):
        scale, ef, lay = spec[i]
        g = gid[(scale, ef)]
        n = info[(scale, ef)]["n"]
        alone, _ = sv.wait(sv.submit_resident(g, 10, params=p_lay if lay else p_cls), n)
        same_alone &= bool(np.array_equal(alone, kept[i]))
        if scale <= 14:
            off, nbr, w = sv.graph_csr(g, weights=True)
            want = (O.layered(off, nbr, 10, 1, 4, w=w)[0][-1] if lay else O.classic(off, nbr, 10, w=w)[-1])
            oracle_ok &= bool(np.array_equal(want, kept[i]))
            oracle_n += 1
    sv.close()
    pct = lambda a, q: round(float(np.percentile(a, q)), 3)  # noqa: E731
    out = {
        "metric": "served LP jobs/s with p50/p99 latency (Poisson arrivals, 1 GPU)",
        "value": round(args.jobs / span_s, 2), "unit": "jobs/s", "steps": 10,
        "config": {"workload": "%d weighted classic/layered(gamma=1/4) RMAT jobs 50/50, scale U{%d..%d}, "
                               "edge factor U{8..32}, (a,b,c,d)=(.57,.19,.19,.05), weights U[1,8], 10 steps, "
                               "Poisson rate %.0f/s, 200 GB budget, %d slots; graph pool: one seeded RMAT graph "
                               "per (scale, edge factor), generated in HBM before the stream and resident"
                               % (args.jobs, lo, hi, args.rate, args.slots),
                   "jobs": args.jobs, "rate_per_s": args.rate, "slots": args.slots,
                   "pool_graphs": len(gid), "pool_bytes": int(pool_bytes), "pool_build_s": round(pool_s, 2),
                   "m_total": int(m_tot)},
        "all_done": ok, "wall_s": round(wall, 2),
        "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99), "mean_ms": round(float(lat.mean()), 3),
        "phases_ms": {"queue_p50": pct(queue, 50), "queue_p99": pct(queue, 99),
                      "admit_to_start_p50": pct(start, 50), "admit_to_start_p99": pct(start, 99),
                      "run_p50": pct(run, 50), "run_p99": pct(run, 99)},
        "edge_steps_per_s_G": round(m_tot * 10 / span_s / 1e9, 2),
        "results_equal_alone": same_alone, "checked_jobs": len(check),
        "oracle_checked_jobs": oracle_n, "oracle_equal": oracle_ok,
    }
    log("served: %.1f jobs/s, p50 %.1f ms, p99 %.1f ms, equal alone %s, oracle %s (%d)" % (
        out["value"], out["p50_ms"], out["p99_ms"], same_alone, oracle_ok, oracle_n))
    return out


# ------------------------------------------------------------ calibrations ----

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


# -------------------------------------------------------------------- main ----

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=25)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--method", choices=["classic", "layered", "counting"], default="classic",
                    help="layered: the headline leg runs layered LP with --gamma; counting: only the count-min, "
                         "frequent and buffer legs (profiling runs)")
    ap.add_argument("--gamma", default="1/4", help="layered gamma p/q (q a power of two)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graph-replay", action="store_true", help="replay steps from a HIP graph")
    ap.add_argument("--gather-steps", type=int, default=3,
                    help="steps of the separate-gather mode timed after the main run (0: skip)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the layered (same graph) and SLPA (LFR-style 2^20) legs")
    ap.add_argument("--frontier", type=int, default=1,
                    help="1: also time the exact-frontier mode on the same steps (reported apart)")
    ap.add_argument("--no-parity", action="store_true", help="skip the from-L0 leg and its oracle sample")
    ap.add_argument("--parity-per-tier", type=int, default=20000)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="tuning override (pack_max_deg, wave_max_deg, block_max_deg, hub_chunk, "
                         "part_keys, part_limit, block_threads); repeatable")
    ap.add_argument("--workload", choices=["rmat", "batched", "served", "slpa", "copy", "random-read"],
                    default="rmat")
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--graphs", type=int, default=1024)
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--batched-only", action="store_true",
                    help="batched workload without the plain per-graph comparison (for profiling)")
    ap.add_argument("--jobs", type=int, default=20000)
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--slots", type=int, default=64)
    ap.add_argument("--serve-scales", type=int, nargs=2, default=[12, 22])
    ap.add_argument("--serve-checks", type=int, default=16)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import glp_amd as G
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O  # the checker and the CPU baseline's loops
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
        print(json.dumps(bench_served(args, G, O)), flush=True)
        return
    if args.workload == "slpa":
        out = slpa_leg(args, G)
        out.update({"metric": "GTEPS per SLPA step (LFR-style 2^20, T=20)", "value": out["GTEPS"],
                    "unit": "GTEPS"})
        print(json.dumps(out), flush=True)
        return
    if args.workload == "copy":
        print(json.dumps(bench_copy(args)), flush=True)
        return
    if args.workload == "random-read":
        print(json.dumps(bench_random_read(args, G)), flush=True)
        return
    threads = host_threads()
    log("generating RMAT scale %d ef %d seed %d (%d threads)" % (args.scale, args.edge_factor, args.seed, threads))
    g = G.Graph.rmat(args.scale, args.edge_factor, 0.57, 0.19, 0.19, seed=args.seed, degree_sort=True,
                     threads=threads)
    info = g.info()
    n, m = info["n"], info["m"]
    log("graph n=%d m=%d max_degree=%d" % (n, m, info["max_degree"]))
    wl_key = "rmat-s%d-ef%d-seed%d-classic" % (args.scale, args.edge_factor, args.seed)

    if args.method == "counting":
        print(json.dumps({"counting": counting_legs(args, G, g, m)}), flush=True)
        return

    if args.method == "layered":
        # profiling form: the headline leg runs layered LP (gamma p/q)
        gp, gq = (int(x) for x in args.gamma.split("/"))
        ls = G.Session(g, G.make_params(method=G.LAYERED, gamma=(gp, gq), profile=True))
        ls.step(args.warmup)
        ls.sync()
        base = ls.timing()
        t0 = time.perf_counter()
        ls.step(args.steps)
        ls.sync()
        lms = (time.perf_counter() - t0) * 1e3 / args.steps
        tiers = tiers_over(base, ls.timing(), args.steps)
        _, roof = dominant_roofline(tiers, 12, 16)
        ls.free()
        print(json.dumps({"metric": "GTEPS per layered LP iteration (RMAT-%d, gamma %d/%d)" % (args.scale, gp, gq),
                          "value": round(m / (lms / 1e3) / 1e9, 3), "unit": "GTEPS", "ms_per_step": round(lms, 4),
                          "steps": args.steps, "warmup": args.warmup, "roofline": roof,
                          "tiers_ms": {k: round(v["ms"], 4) for k, v in tiers.items()}}), flush=True)
        return

    ms_per_step, tiers, changes, dense_labels = rmat_main(args, G, g, n, m)
    gteps = m / (ms_per_step / 1e3) / 1e9
    dom, roofline = dominant_roofline(tiers, 8, 12)
    if roofline is not None:
        traffic = RS.traffic_for(TRAFFIC_FILE, wl_key, "main", TIER_FAMILIES[dom])
        roofline["traffic"] = None if traffic is None else int(traffic)
        # the ceiling of the label gather's access pattern on this device:
        # random 4-byte reads of a label-array-sized table (each one a 128-byte
        # line fetch when it misses L2)
        rr_ms = G.calib_random_read(RR_TABLE, RR_READS, 5)
        rr = RR_READS / (rr_ms / 1e3) / 1e9
        roofline["random_read_ceiling"] = {"G_reads_per_s": round(rr, 2), "GBs_in_128B_lines": round(rr * 128, 1),
                                           "table_MiB": RR_TABLE * 4 >> 20}
        if traffic is not None:
            tg = traffic / (tiers[dom]["ms"] / 1e3) / 1e9
            roofline["traffic_GBs"] = round(tg, 1)
            roofline["traffic_frac"] = round(tg / HBM_PEAK_GBS, 4)
            roofline["traffic_frac_of_random_read_ceiling"] = round(tg / (rr * 128), 4)
            roofline["traffic_source"] = TRAFFIC_SOURCE
    whole = tier_bytes(m, n)
    step_roofline = {"algorithmic_bytes": int(whole), "achieved_GBs": round(whole / (ms_per_step / 1e3) / 1e9, 1),
                     "frac": round(whole / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}

    from_l0, parity = None, None
    off = nbr = None
    if not args.no_parity or not args.no_cpu:
        off, nbr, _ = g.csr()
    if not args.no_parity:
        from_l0, parity = from_l0_and_parity(args, G, O, g, off, nbr, n, m)
    frontier = frontier_leg(args, G, g, m, dense_labels) if args.frontier else None
    extra = None
    if not args.no_extra:
        extra = {"layered": layered_legs(args, G, g, n, m), "slpa": slpa_leg(args, G),
                 "counting": counting_legs(args, G, g, m)}
    gather = gather_leg(args, G, g, m, wl_key) if args.gather_steps > 0 else None
    cpu = cpu_baseline(O, G, args, off, nbr) if not args.no_cpu else None
    del off, nbr
    g.free()
    batched = bench_batched(args, G) if not args.no_batched else None
    served = bench_served(args, G, O) if not args.no_served else None

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
        "from_L0_10_steps": from_l0,
        "parity_sample": parity,
        "batched": batched,
        "served": served,
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
