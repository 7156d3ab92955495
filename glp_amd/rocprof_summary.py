"""Text summaries of rocprofv3 output for the LP kernels.

    python -m glp_amd.rocprof_summary stats DIR --command "..." > profiles/X.txt
    python -m glp_amd.rocprof_summary calibrate FETCH_DIR WRITE_DIR --bytes N --command "..."
    python -m glp_amd.rocprof_summary traffic FETCH_DIR WRITE_DIR --calibration FILE
            --warmup W --steps K [--gather-steps G] --workload KEY --command "..."
    python -m glp_amd.rocprof_summary sizes DIR --warmup W --steps K [--gather-steps G] --command "..."
    python -m glp_amd.rocprof_summary steps DIR --steps N --command "..."
    python -m glp_amd.rocprof_summary counters DIR [DIR ...] [--match NAME] --command "..."

`stats` condenses a --kernel-trace --stats run.  `calibrate` compares the
FETCH_SIZE / WRITE_SIZE that rocprofv3 reports for device-to-device copies of
a known size with the bytes the copies move and writes the two scale factors.
`traffic` splits a --pmc run of `bench.py` into LP steps (every step ends
with one k_end_step dispatch) and reports, per kernel family, the calibrated
HBM bytes per step over the timed steps; `traffic_for` reads that file back.
`steps` lists a --kernel-trace run's kernel time per LP step and family.
`counters` sums any --pmc counters per kernel instance.
"""
import argparse
import csv
import glob
import os
import sys
from collections import defaultdict


def family(kernel_name):
    """Short kernel name: 'void glp::k_block<...>(...)' -> 'k_block'."""
    s = kernel_name.strip().replace("(anonymous namespace)::", "")
    if s.startswith("void "):
        s = s[5:]
    s = s.split("(")[0].split("<")[0]
    return s.split("::")[-1]


def _csvs(d, suffix):
    return sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True))


def read_counter(d, counter):
    """Per dispatch (id, family, value) of one counter, summed over rows."""
    acc = {}
    for f in _csvs(d, "counter_collection.csv"):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                if r["Counter_Name"] != counter:
                    continue
                k = int(r["Dispatch_Id"])
                prev = acc.get(k, (family(r["Kernel_Name"]), 0.0))
                acc[k] = (prev[0], prev[1] + float(r["Counter_Value"]))
    return [(k, acc[k][0], acc[k][1]) for k in sorted(acc)]


def split_steps(rows):
    """Groups dispatches into steps closed by k_end_step; dispatches after the
    last k_end_step are dropped."""
    steps, cur = [], []
    for row in rows:
        cur.append(row)
        if row[1] == "k_end_step":
            steps.append(cur)
            cur = []
    return steps


def cmd_stats(a):
    files = _csvs(a.dir, "kernel_stats.csv")
    if not files:
        sys.exit("no kernel_stats.csv under %s" % a.dir)
    rows = list(csv.DictReader(open(files[0])))
    out = ["# rocprofv3 kernel statistics", "# command: %s" % a.command, "#",
           "%-16s %7s %12s %12s %12s %7s" % ("kernel", "calls", "avg_ms", "min_ms", "max_ms", "pct")]
    agg = defaultdict(lambda: [0, 0.0, float("inf"), 0.0, 0.0])
    for r in rows:
        e = agg[family(r["Name"])]
        c = int(r["Calls"])
        e[0] += c
        e[1] += float(r["TotalDurationNs"])
        e[2] = min(e[2], float(r["MinNs"]))
        e[3] = max(e[3], float(r["MaxNs"]))
        e[4] += float(r["Percentage"])
    for k, e in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        out.append("%-16s %7d %12.4f %12.4f %12.4f %7.2f" % (k, e[0], e[1] / e[0] / 1e6, e[2] / 1e6,
                                                          e[3] / 1e6, e[4]))
    out.append("#")
    out.append("# full kernel names")
    for r in rows:
        out.append("#   %s calls=%s avg_ns=%s" % (r["Name"][:160], r["Calls"], r["AverageNs"]))
    print("\n".join(out))


def cmd_calibrate(a):
    res = []
    for d, counter in ((a.fetch_dir, "FETCH_SIZE"), (a.write_dir, "WRITE_SIZE")):
        vals = [v for _, fam, v in read_counter(d, counter) if "copyBuffer" in fam]
        if not vals:
            sys.exit("no copy dispatches with %s under %s" % (counter, d))
        kib = sum(vals) / len(vals)
        res.append((counter, kib, a.bytes / (kib * 1024.0), len(vals)))
    out = ["# rocprofv3 byte-counter calibration: device-to-device copies of %d bytes" % a.bytes,
           "# command: %s" % a.command, "#",
           "# counter     KiB/dispatch  dispatches  scale (bytes moved / counter bytes)"]
    for counter, kib, scale, nd in res:
        out.append("# %-11s %13.1f %11d %8.4f" % (counter, kib, nd, scale))
    out.append("fetch_scale %.4f" % res[0][2])
    out.append("write_scale %.4f" % res[1][2])
    print("\n".join(out))


def read_calibration(path):
    sc = {"fetch_scale": 1.0, "write_scale": 1.0}
    with open(path) as fh:
        for line in fh:
            p = line.split()
            if len(p) == 2 and p[0] in sc:
                sc[p[0]] = float(p[1])
    return sc


def cmd_traffic(a):
    sc = read_calibration(a.calibration)
    per = {}
    for d, counter, scale in ((a.fetch_dir, "FETCH_SIZE", sc["fetch_scale"]),
                              (a.write_dir, "WRITE_SIZE", sc["write_scale"])):
        steps = split_steps(read_counter(d, counter))
        need = a.warmup + a.steps + (1 + a.gather_steps if a.gather_steps else 0)
        if len(steps) < need:
            sys.exit("%s: %d steps found, %d expected" % (counter, len(steps), need))
        groups = {"main": steps[a.warmup:a.warmup + a.steps]}
        if a.gather_steps:
            g0 = a.warmup + a.steps + 1
            groups["gather"] = steps[g0:g0 + a.gather_steps]
        for gname, ss in groups.items():
            for st in ss:
                for _, fam, v in st:
                    e = per.setdefault((gname, fam), {"FETCH_SIZE": 0.0, "WRITE_SIZE": 0.0, "launches": 0})
                    e[counter] += v * 1024.0 * scale / len(ss)
                    if counter == "FETCH_SIZE":
                        e["launches"] += 1.0 / len(ss)
    out = ["# HBM traffic per LP step from rocprofv3 FETCH_SIZE and WRITE_SIZE, averaged over the",
           "# timed steps (%d warmup + %d timed; separate-gather session: %d timed after 1)"
           % (a.warmup, a.steps, a.gather_steps),
           "# command: %s" % a.command,
           "# scales: FETCH_SIZE x %.4f, WRITE_SIZE x %.4f (device-to-device copy calibration)"
           % (sc["fetch_scale"], sc["write_scale"]), "#",
           "# session kernel        launches/step   fetch_bytes    write_bytes    traffic_bytes"]
    lines = []
    for (gname, fam), e in sorted(per.items()):
        t = e["FETCH_SIZE"] + e["WRITE_SIZE"]
        out.append("# %-7s %-16s %8.1f %15.0f %14.0f %16.0f" % (gname, fam, e["launches"], e["FETCH_SIZE"],
                                                                e["WRITE_SIZE"], t))
        lines.append("traffic %s %s %.0f" % (gname, fam, t))
    print("\n".join(out + ["workload %s" % a.workload] + lines))


SIZE_COUNTERS = ("TCC_EA0_RDREQ_32B_sum", "TCC_EA0_RDREQ_64B_sum", "TCC_EA0_RDREQ_128B_sum",
                 "TCC_BUBBLE_sum")


def cmd_sizes(a):
    """Read bytes per step from the L2 request-size counters (32/64/128 B),
    next to TCC_BUBBLE (the 128 B count FETCH_SIZE's formula uses)."""
    per = defaultdict(lambda: defaultdict(float))
    for c in SIZE_COUNTERS:
        steps = split_steps(read_counter(a.dir, c))
        groups = {"main": steps[a.warmup:a.warmup + a.steps]}
        if a.gather_steps:
            g0 = a.warmup + a.steps + 1
            groups["gather"] = steps[g0:g0 + a.gather_steps]
        for gname, ss in groups.items():
            for st in ss:
                for _, fam, v in st:
                    per[(gname, fam)][c] += v / len(ss)
    out = ["# L2-to-memory read requests by size per LP step (rocprofv3 TCC_EA0_RDREQ_{32B,64B,128B},",
           "# TCC_BUBBLE), averaged over the timed steps, and the read bytes they imply",
           "# command: %s" % a.command, "#",
           "# session kernel            req_32B      req_64B     req_128B  TCC_BUBBLE    read_bytes"]
    for (gname, fam), v in sorted(per.items()):
        b = 32 * v[SIZE_COUNTERS[0]] + 64 * v[SIZE_COUNTERS[1]] + 128 * v[SIZE_COUNTERS[2]]
        if b < 1e6:
            continue
        out.append("# %-7s %-16s %10.0f %12.0f %12.0f %11.0f %13.0f" % (
            gname, fam, v[SIZE_COUNTERS[0]], v[SIZE_COUNTERS[1]], v[SIZE_COUNTERS[2]], v[SIZE_COUNTERS[3]], b))
    print("\n".join(out))


def traffic_for(path, workload, session, families):
    """Sum of the per-step traffic of the given kernel families measured on
    this workload, or None."""
    if not os.path.exists(path):
        return None
    tot, found, same = 0.0, False, False
    with open(path) as fh:
        for line in fh:
            p = line.split()
            if len(p) == 2 and p[0] == "workload":
                same = p[1] == workload
            if same and len(p) == 4 and p[0] == "traffic" and p[1] == session and p[2] in families:
                tot += float(p[3])
                found = True
    return tot if found else None


def cmd_steps(a):
    """Per-step kernel time (ms per family) of the first LP session in a
    --kernel-trace run: dispatches in start order, steps closed by
    k_end_step, the first `steps` steps."""
    files = _csvs(a.dir, "kernel_trace.csv")
    if not files:
        sys.exit("no kernel_trace.csv under %s" % a.dir)
    rows = []
    for f in files:
        with open(f) as fh:
            for r in csv.DictReader(fh):
                t0, t1 = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
                rows.append((t0, family(r["Kernel_Name"]), (t1 - t0) / 1e6))
    rows.sort()
    steps = split_steps([(t, fam, ms) for t, fam, ms in rows])[: a.steps]
    fams = defaultdict(float)
    for st in steps:
        for _, fam, ms in st:
            if fam != "k_end_step":
                fams[fam] += ms
    order = [f for f, _ in sorted(fams.items(), key=lambda kv: -kv[1])][: a.columns]
    out = ["# kernel time per LP step (ms, from the trace's start/end timestamps)",
           "# command: %s" % a.command, "#",
           "step " + " ".join("%13s" % f for f in order) + "        total"]
    for i, st in enumerate(steps):
        per = defaultdict(float)
        for _, fam, ms in st:
            per[fam] += ms
        out.append("%4d " % (i + 1) + " ".join("%13.3f" % per.get(f, 0.0) for f in order) +
                   " %12.3f" % sum(per.values()))
    print("\n".join(out))


def instance(kernel_name):
    """Kernel family with its template arguments: 'k_half<..., 4, 16, 4>'."""
    s = kernel_name.strip().replace("(anonymous namespace)::", "")
    if s.startswith("void "):
        s = s[5:]
    s = s.split("(")[0]
    head, _, tmpl = s.partition("<")
    name = head.split("::")[-1]
    return name + ("<" + tmpl.replace("glp::", "") if tmpl else "")


def cmd_counters(a):
    """Per kernel instance, each counter summed over its dispatches (several
    --pmc runs of one command may be given), plus derived ratios."""
    acc = defaultdict(lambda: defaultdict(float))
    calls = defaultdict(set)
    for d in a.dirs:
        for f in _csvs(d, "counter_collection.csv"):
            with open(f) as fh:
                for r in csv.DictReader(fh):
                    k = instance(r["Kernel_Name"])
                    acc[k][r["Counter_Name"]] += float(r["Counter_Value"])
                    calls[k].add((d, r["Dispatch_Id"]))
    names = sorted({c for v in acc.values() for c in v})
    out = ["# rocprofv3 counters per kernel instance (sums over dispatches)",
           "# command: %s" % a.command, "#"]
    for k in sorted(acc, key=lambda k: -acc[k].get("SQ_WAVE_CYCLES", 0.0)):
        if a.match and not any(m in k for m in a.match):
            continue
        v = acc[k]
        out.append("%s  (dispatches %d)" % (k, len(calls[k]) // max(len(a.dirs), 1)))
        out.append("    " + "  ".join("%s=%.4g" % (c, v[c]) for c in names if c in v))
        if "TCC_HIT_sum" in v and v.get("TCC_HIT_sum", 0) + v.get("TCC_MISS_sum", 0) > 0:
            out.append("    L2 hit rate %.3f" % (v["TCC_HIT_sum"] / (v["TCC_HIT_sum"] + v["TCC_MISS_sum"])))
        if v.get("SQ_WAVE_CYCLES"):
            for c in ("SQ_ACTIVE_INST_ANY", "SQ_WAIT_INST_ANY"):
                if c in v:
                    out.append("    %s / SQ_WAVE_CYCLES %.3f" % (c, v[c] / v["SQ_WAVE_CYCLES"]))
        if v.get("SQ_INSTS_LDS"):
            out.append("    SQ_LDS_BANK_CONFLICT / SQ_INSTS_LDS %.2f" % (v.get("SQ_LDS_BANK_CONFLICT", 0) / v["SQ_INSTS_LDS"]))
    print("\n".join(out))


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--command", required=True)
    c = sub.add_parser("calibrate")
    c.add_argument("fetch_dir")
    c.add_argument("write_dir")
    c.add_argument("--bytes", type=int, required=True)
    c.add_argument("--command", required=True)
    t = sub.add_parser("traffic")
    t.add_argument("fetch_dir")
    t.add_argument("write_dir")
    t.add_argument("--calibration", required=True)
    t.add_argument("--warmup", type=int, required=True)
    t.add_argument("--steps", type=int, required=True)
    t.add_argument("--gather-steps", type=int, default=0)
    t.add_argument("--workload", required=True, help="key, e.g. rmat-s25-ef16-seed1-classic")
    t.add_argument("--command", required=True)
    z = sub.add_parser("sizes")
    z.add_argument("dir")
    z.add_argument("--warmup", type=int, required=True)
    z.add_argument("--steps", type=int, required=True)
    z.add_argument("--gather-steps", type=int, default=0)
    z.add_argument("--command", required=True)
    p = sub.add_parser("steps")
    p.add_argument("dir")
    p.add_argument("--steps", type=int, required=True, help="steps of the first session to list")
    p.add_argument("--columns", type=int, default=9, help="kernel families shown (by total time)")
    p.add_argument("--command", required=True)
    n = sub.add_parser("counters")
    n.add_argument("dirs", nargs="+")
    n.add_argument("--match", action="append", default=[], help="keep kernels whose name holds this")
    n.add_argument("--command", required=True)
    a = ap.parse_args()
    {"stats": cmd_stats, "calibrate": cmd_calibrate, "traffic": cmd_traffic, "sizes": cmd_sizes,
     "steps": cmd_steps, "counters": cmd_counters}[a.cmd](a)


if __name__ == "__main__":
    main()
