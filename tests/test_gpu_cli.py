"""The GPULP command line and the autotuner on the GPU: argument forms of
main.cu / README.md, the appended log line, and labels against the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

import glp_amd as G
import oracle_lib as O

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(G.LIB_PATH), "bin")


def gpulp(*args, cwd=None):
    r = subprocess.run([os.path.join(BIN, "GPULP")] + [str(a) for a in args], capture_output=True,
                       text=True, cwd=cwd, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


@pytest.fixture(scope="module")
def graph_file(tmp_path_factory):
    d = tmp_path_factory.mktemp("cli")
    g = G.Graph.rmat(12, 16, seed=41)
    p = str(d / "g.bin")
    g.save_bin(p)
    return p


def test_positional_form_and_log_line(graph_file, tmp_path):
    out = tmp_path / "out.log"
    lab = tmp_path / "labels.bin"
    stdout = gpulp(1, 6, graph_file, out, 0, "labels=%s" % lab)
    assert "success run" in stdout
    line = out.read_text().strip().splitlines()[-1]
    assert line.startswith("graph: %s method: 1 iteration: 6 time:  load graph to GPU: " % graph_file)
    for field in ("other proprocess:", "label_load:", "LP:", "tiny_speed:", "huge_speed:",
                  "tiny_distrbution:", "huge_distribution:"):
        assert field in line
    off, nbr, _ = O.load_bin(graph_file)
    labels = np.fromfile(lab, dtype=np.int32)
    assert np.array_equal(labels, O.classic(off, nbr, 5)[-1])   # niter=6 -> 5 steps


def test_key_value_form_and_other_methods(graph_file, tmp_path):
    out = tmp_path / "kv.log"
    lab = tmp_path / "l.bin"
    gpulp("method=1", "iter=4", "input=%s" % graph_file, "output=%s" % out, "ifdirect=0", "ifweighted=0",
          "labels=%s" % lab)
    off, nbr, _ = O.load_bin(graph_file)
    assert np.array_equal(np.fromfile(lab, dtype=np.int32), O.classic(off, nbr, 3)[-1])
    gpulp(2, 3, graph_file, out, 0, "gamma=1/4", "labels=%s" % lab)
    assert np.array_equal(np.fromfile(lab, dtype=np.int32), O.layered(off, nbr, 2, 1, 4)[0][-1])
    gpulp(2, 2, graph_file, out, 0)          # the 10-round delta schedule
    gpulp(3, 5, graph_file, out, 0, "seed=9", "labels=%s" % lab)
    mem = O.slpa(off, nbr, seed=9, T=4)
    dom, _, _ = O.slpa_post(mem, 1, 10, 1)
    assert np.array_equal(np.fromfile(lab, dtype=np.int32), dom)
    gpulp(5, 3, graph_file, out, 0, "cm=3,64,1", "labels=%s" % lab)
    prev = np.arange(len(off) - 1, dtype=np.int32)
    for _ in range(2):
        prev = O.countmin_step(off, nbr, prev, 3, 64, 1)
    assert np.array_equal(np.fromfile(lab, dtype=np.int32), prev)
    for method, buf in ((6, 0), (7, 1)):   # frequent and buffer, 5 slots
        gpulp(method, 3, graph_file, out, 0, "cap=5", "labels=%s" % lab)
        prev = np.arange(len(off) - 1, dtype=np.int32)
        for _ in range(2):
            prev = O.table_step(off, nbr, prev, 5, buf)
        assert np.array_equal(np.fromfile(lab, dtype=np.int32), prev), "method %d" % method
    lines = out.read_text().strip().splitlines()
    assert re.match(r"graph: .* method: 2 iteration: 3 time: [0-9.e-]+$", lines[1])


def test_replay_and_user_method(graph_file, tmp_path):
    lab = tmp_path / "r.bin"
    gpulp(4, 5, graph_file, tmp_path / "o.log", 0, "replay=1", "labels=%s" % lab)
    off, nbr, _ = O.load_bin(graph_file)
    assert np.array_equal(np.fromfile(lab, dtype=np.int32), O.classic(off, nbr, 4)[-1])


def test_autotune_keeps_results(tmp_path):
    g = G.Graph.rmat(13, 16, seed=42)
    off, nbr, _ = g.csr()
    saved = {k: G.tuning_get(k) for k in G.TUNING_KEYS}
    try:
        p = G.make_params()
        best = G.autotune(g, p, 3, str(tmp_path / "t.tuning"))
        assert best > 0
        text = (tmp_path / "t.tuning").read_text()
        written = dict(line.split("=") for line in text.split())
        # every key is written, with the tuning the autotuner left in force
        assert sorted(written) == sorted(G.TUNING_KEYS)
        assert all(int(written[k]) == G.tuning_get(k) for k in written)
        labels, _ = G.run(g, 4)
        assert np.array_equal(labels, O.classic(off, nbr, 4)[-1])
    finally:
        for k, v in saved.items():
            G.tuning_set(k, v)


def test_seeded_method_with_seed_file(graph_file, tmp_path):
    """Method 8 reads a seed file (one vertex id per line, as
    layeredLprop::assign_seeds_for_device does) and runs niter-1 steps."""
    off, nbr, _ = O.load_bin(graph_file)
    n = len(off) - 1
    seeds = np.zeros(n, dtype=np.uint8)
    ids = np.random.default_rng(7).choice(n, size=40, replace=False)
    seeds[ids] = 1
    sf = tmp_path / "seeds.txt"
    sf.write_text("".join("%d\n" % i for i in ids))
    out = tmp_path / "s.log"
    lab = tmp_path / "s.bin"
    gpulp(8, 5, graph_file, out, 0, "seeds=%s" % sf, "labels=%s" % lab)
    assert np.array_equal(np.fromfile(lab, dtype=np.int32), O.seeded(off, nbr, 4, seeds)[-1])
    assert "method: 8 iteration: 5 time: " in out.read_text()
    r = subprocess.run([os.path.join(BIN, "GPULP"), "8", "3", graph_file, str(out), "0"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "seeds" in r.stderr
