"""The rocprofv3 summary tool on synthetic counter and trace files: step
splitting at k_end_step, warmup/timed/gather windows, the copy calibration
scales, and the traffic file read back by bench.py."""
import csv
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
from glp_amd import rocprof_summary as RS  # noqa: E402


def write_counters(d, counter, dispatches):
    """dispatches: list of (kernel name, value in KiB), in dispatch order."""
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "run_counter_collection.csv"), "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        for i, (name, v) in enumerate(dispatches):
            w.writerow([i + 1, name, counter, v])


def run_tool(*args):
    r = subprocess.run([sys.executable, "-m", "glp_amd.rocprof_summary", *args], cwd=ROOT,
                       capture_output=True, text=True, check=True)
    return r.stdout


def step(kib_block, kib_hub):
    return [("void glp::k_block<H>(a)", kib_block), ("void glp::k_hub_agg<H>(a)", kib_hub),
            ("glp::k_end_step(x)", 1.0)]


def test_family_names():
    assert RS.family("void glp::k_block<glp::ClassicHooks, false, false, 256>(glp::StepArgs)") == "k_block"
    assert RS.family("glp::k_end_step(unsigned long long*)") == "k_end_step"
    assert RS.family("__amd_rocclr_copyBuffer") == "__amd_rocclr_copyBuffer"


def test_split_steps_drops_the_unfinished_tail():
    rows = [(1, "k_a", 1.0), (2, "k_end_step", 0.0), (3, "k_a", 2.0), (4, "k_end_step", 0.0),
            (5, "k_a", 9.0)]
    steps = RS.split_steps(rows)
    assert [len(s) for s in steps] == [2, 2]


def test_calibrate_and_traffic_windows(tmp_path):
    # calibration: copies of 1 MiB read as FETCH 512 KiB and WRITE 1024 KiB
    fc, wc = str(tmp_path / "cf"), str(tmp_path / "cw")
    write_counters(fc, "FETCH_SIZE", [("__amd_rocclr_copyBuffer", 512.0)] * 3)
    write_counters(wc, "WRITE_SIZE", [("__amd_rocclr_copyBuffer", 1024.0)] * 3)
    cal = run_tool("calibrate", fc, wc, "--bytes", str(1 << 20), "--command", "copy")
    calf = tmp_path / "cal.txt"
    calf.write_text(cal)
    assert RS.read_calibration(str(calf)) == {"fetch_scale": 2.0, "write_scale": 1.0}
    # main session: 1 warmup step (ignored), 2 timed steps; then one untimed
    # gather-session step and 1 timed gather step
    main = step(999.0, 999.0) + step(10.0, 4.0) + step(30.0, 8.0)
    gather = step(777.0, 777.0) + step(50.0, 2.0)
    fd, wd = str(tmp_path / "f"), str(tmp_path / "w")
    write_counters(fd, "FETCH_SIZE", main + gather)
    write_counters(wd, "WRITE_SIZE", [(n, 1.0) for n, _ in main + gather])
    out = run_tool("traffic", fd, wd, "--calibration", str(calf), "--warmup", "1", "--steps", "2",
                   "--gather-steps", "1", "--workload", "toy", "--command", "bench")
    tf = tmp_path / "traffic.txt"
    tf.write_text(out)
    # per timed step: fetch = mean KiB x 1024 x 2, write = 1 KiB x 1024 x 1
    want_block = (10.0 + 30.0) / 2 * 1024 * 2 + 1024
    want_hub = (4.0 + 8.0) / 2 * 1024 * 2 + 1024
    assert RS.traffic_for(str(tf), "toy", "main", ["k_block"]) == pytest.approx(want_block)
    assert RS.traffic_for(str(tf), "toy", "main", ["k_block", "k_hub_agg"]) == pytest.approx(want_block + want_hub)
    assert RS.traffic_for(str(tf), "toy", "gather", ["k_block"]) == pytest.approx(50.0 * 2048 + 1024)
    assert RS.traffic_for(str(tf), "other", "main", ["k_block"]) is None
    assert RS.traffic_for(str(tmp_path / "missing.txt"), "toy", "main", ["k_block"]) is None


def test_steps_table(tmp_path):
    d = tmp_path / "trace"
    d.mkdir()
    rows = []
    t = 0
    for ms_block, ms_hub in ((2.0, 1.0), (3.0, 0.5)):
        for name, ms in (("void glp::k_block<H>(a)", ms_block), ("void glp::k_hub_agg<H>(a)", ms_hub),
                         ("glp::k_end_step(x)", 0.001)):
            rows.append((t, t + int(ms * 1e6), name))
            t += int(ms * 1e6) + 1000
    with open(d / "run_kernel_trace.csv", "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        for t0, t1, name in rows:
            w.writerow([name, t0, t1])
    out = run_tool("steps", str(d), "--steps", "2", "--command", "trace").splitlines()
    header = [ln for ln in out if ln.startswith("step")][0].split()
    assert header[1:3] == ["k_block", "k_hub_agg"]
    body = [ln.split() for ln in out if ln[:1] == " " or ln[:1].isdigit()]
    assert [float(x) for x in body[0][1:3]] == [2.0, 1.0]
    assert [float(x) for x in body[1][1:3]] == [3.0, 0.5]
    assert float(body[1][-1]) == pytest.approx(3.501)


def test_instance_names_and_anonymous_namespace():
    assert RS.instance("void glp::k_half<glp::ClassicHooks, false, false, 4, 16, 4>(glp::StepArgs)") == \
        "k_half<ClassicHooks, false, false, 4, 16, 4>"
    assert RS.instance("glp::k_end_step(unsigned long long*)") == "k_end_step"
    assert RS.family("glp::(anonymous namespace)::k_key_offsets(unsigned long long const*, long)") == \
        "k_key_offsets"


def test_counters_sums_per_instance(tmp_path):
    a, b = str(tmp_path / "a"), str(tmp_path / "b")
    write_counters(a, "SQ_INSTS_LDS", [("void glp::k_half<H, 4>(x)", 10.0), ("void glp::k_half<H, 4>(x)", 30.0),
                                       ("void glp::k_packed<H, 4>(x)", 5.0)])
    write_counters(b, "TCC_HIT_sum", [("void glp::k_half<H, 4>(x)", 3.0), ("void glp::k_half<H, 4>(x)", 3.0),
                                      ("void glp::k_packed<H, 4>(x)", 1.0)])
    # the miss counter in a third file of the same run as b
    with open(os.path.join(b, "run_counter_collection.csv"), "a", newline="") as fh:
        w = csv.writer(fh)
        w.writerow([1, "void glp::k_half<H, 4>(x)", "TCC_MISS_sum", 2.0])
        w.writerow([2, "void glp::k_half<H, 4>(x)", "TCC_MISS_sum", 2.0])
    out = run_tool("counters", a, b, "--match", "k_half", "--command", "cmd")
    assert "k_packed" not in out
    assert "k_half<H, 4>  (dispatches 2)" in out
    assert "SQ_INSTS_LDS=40" in out and "TCC_HIT_sum=6" in out
    assert "L2 hit rate 0.600" in out
