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

HBM_PEAK_GBS = 8000.0
METRIC = "GTEPS per LP iteration (RMAT-25, 1 GPU); batched graphs/s vs plain; p50/p99"


def log(*a):
    print("[bench %.1fs]" % (time.time() - T0), *a, file=sys.stderr, flush=True)


T0 = time.time()


def tier_bytes(edges, verts, weighted=False):
    """Algorithmic HBM bytes of one tier launch: per edge the neighbour id and
    its label (plus the weight), per vertex the offset, the kept label and
    the written label."""
    return (12 if weighted else 8) * edges + 12 * verts


def cpu_baseline(off, nbr, target_s, threads, max_steps):
    """The reference CPU loop (tutorial/CPUlp.cuh:28-92 sync_run_1, restated:
    OpenMP over vertices, a std::map count per vertex, live label reads) run
    for whole steps over the same graph from L0, as many as fit in target_s
    (at least one, at most max_steps)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    m = int(off[-1])
    _, secs1 = O.cpulp_run(off, nbr, 1, threads, 0)
    steps = int(max(1, min(max_steps, target_s // max(secs1, 1e-9))))
    secs = secs1
    if steps > 1:
        _, secs = O.cpulp_run(off, nbr, steps, threads, 0)
    return {"value": steps * m / secs / 1e9, "unit": "GTEPS", "cores": threads, "kind": "port",
            "sample": "the first %d step(s) of verbatim CPUlp on the same graph (all %d vertices, "
                      "%d edges per step), %.2f s; a separate first step took %.2f s"
                      % (steps, len(off) - 1, m, secs, secs1)}


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
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import glp_amd as G
    if G.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
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
                    "hub": "k_hub_count", "aux": "k_hub_reduce+k_hub_final"}
    dom = max((k for k in tiers if k in kernel_names), key=lambda k: tiers[k]["ms"], default=None)
    roofline = None
    if dom is not None and dom != "aux":
        b = tier_bytes(tiers[dom]["edges"], tiers[dom]["vertices"])
        achieved = b / (tiers[dom]["ms"] / 1e3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                    "kernel": kernel_names[dom], "kernel_ms": round(tiers[dom]["ms"], 4),
                    "algorithmic_bytes": int(b)}
    whole = tier_bytes(m, n)
    step_roofline = {"algorithmic_bytes": int(whole),
                     "achieved_GBs": round(whole / (ms_per_step / 1e3) / 1e9, 1),
                     "frac": round(whole / (ms_per_step / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
    sess.free()

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (%d threads, ~%.0f s)" % (threads, args.cpu_seconds))
        off, nbr, _ = g.csr()
        cpu = cpu_baseline(off, nbr, args.cpu_seconds, threads, args.steps)
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
                   "graph_replay": bool(args.graph_replay)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "step_roofline": step_roofline,
        "tiers_ms": {k: round(v["ms"], 4) for k, v in tiers.items()},
        "changes_per_step": changes[-args.steps:],
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
