"""Host-side checks that need no GPU: the C ABI exports, the .bin reader
against the reference's own loader, the tools against the reference's own
tool binaries (built from its sources into oracle/_ref), generators, and
error paths."""
import os
import re
import struct
import subprocess

import numpy as np
import pytest

import glp_amd as G
import oracle_lib as O

ROOT = O.ROOT
GOLD = os.path.join(ROOT, "tests", "golden")
REF = O.REF_DIR


def declared_symbols():
    names = set()
    for h in ("glp.h", "glp_tools.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(glp_[a-z0-9_]+)\s*\(", src))
    return names


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", G.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (glp_[a-z0-9_]+)", out))
    missing = declared_symbols() - exported
    assert not missing, sorted(missing)


def test_version_and_params():
    assert "gfx950" in G.version()
    p = G.make_params()
    assert p.method == G.CLASSIC and p.gamma_den == 1 and list(p.reserved) == [0] * 5


def write_pairs(path, pairs, extra=b""):
    with open(path, "wb") as f:
        for u, v in pairs:
            f.write(struct.pack("<ii", u, v))
        f.write(extra)


def ref_load(path, directed=False):
    out = path + ".csr"
    subprocess.run([os.path.join(REF, "ref_loadbin"), path, str(int(directed)), out], check=True,
                   capture_output=True)
    raw = np.fromfile(out, dtype=np.int32)
    n, m = int(raw[0]), int(raw[1])
    return raw[2:3 + n], raw[3 + n:3 + n + m]


needs_ref = pytest.mark.skipif(not os.path.exists(os.path.join(REF, "ref_loadbin")),
                               reason="oracle/_ref not built (reference tree absent)")


@needs_ref
@pytest.mark.parametrize("case", ["small", "dups_loops_partial", "random"])
def test_load_bin_matches_reference_loader(tmp_path, case):
    p = str(tmp_path / ("g_%s.bin" % case))
    if case == "small":
        write_pairs(p, [(0, 1), (1, 2), (2, 0), (3, 4)])
    elif case == "dups_loops_partial":
        write_pairs(p, [(0, 1), (0, 1), (2, 2), (5, 1), (1, 0), (7, 3)], extra=b"\x01\x02\x03")
    else:
        rng = np.random.default_rng(1)
        write_pairs(p, [tuple(x) for x in rng.integers(0, 500, size=(3000, 2))])
    roff, rnbr = ref_load(p)
    ooff, onbr, _ = O.load_bin(p)
    g = G.Graph.load_bin(p)
    goff, gnbr, _ = g.csr()
    assert np.array_equal(roff, ooff) and np.array_equal(rnbr, onbr)
    assert np.array_equal(roff, goff) and np.array_equal(rnbr, gnbr)


def test_load_bin_weighted_and_directed(tmp_path):
    p = str(tmp_path / "w.bin")
    recs = [(0, 1, 5), (1, 2, 7), (2, 2, 9), (3, 0, 1)]
    with open(p, "wb") as f:
        for r in recs:
            f.write(struct.pack("<iii", *r))
    g = G.Graph.load_bin(p, weighted=True)
    off, nbr, w = g.csr()
    ooff, onbr, ow = O.load_bin(p, weighted=True)
    assert np.array_equal(off, ooff) and np.array_equal(nbr, onbr) and np.array_equal(w, ow)
    assert w.tolist() == [5, 1, 5, 7, 7, 1]
    d = G.Graph.load_bin(p, directed=True, weighted=True)
    # directed: n = max id + 1 (the reference's max id is one short)
    assert d.n == 4 and d.m == 3


def test_load_bin_errors(tmp_path):
    with pytest.raises(G.GLPError) as e:
        G.Graph.load_bin(str(tmp_path / "missing.bin"))
    assert e.value.code == -2
    p = str(tmp_path / "neg.bin")
    write_pairs(p, [(0, 1), (-3, 2)])
    with pytest.raises(G.GLPError) as e:
        G.Graph.load_bin(p)
    assert e.value.code == -3


def test_from_csr_validation():
    with pytest.raises(G.GLPError):
        G.Graph.from_csr([0, 2, 1], [1, 0])               # not monotone
    with pytest.raises(G.GLPError):
        G.Graph.from_csr([0, 1], [5])                     # neighbour out of range
    g = G.Graph.from_csr([0, 1, 2], [1, 0], weights=[3, 3])
    assert g.info() == {"n": 2, "m": 2, "max_degree": 1, "weighted": True}


def test_save_bin_roundtrip(tmp_path):
    g = G.Graph.rmat(10, 8, seed=2)
    p = str(tmp_path / "r.bin")
    g.save_bin(p)
    h = G.Graph.load_bin(p)
    a, b = g.csr(), h.csr()
    # same edge sets per vertex (the reader keeps file order, the generator ascending)
    assert np.array_equal(a[0][:h.n + 1], b[0])
    for v in range(h.n):
        assert sorted(a[1][a[0][v]:a[0][v + 1]]) == sorted(b[1][b[0][v]:b[0][v + 1]])


@needs_ref
def test_edgelist2bin_matches_reference(tmp_path):
    txt = tmp_path / "edges.txt"
    rng = np.random.default_rng(4)
    lines = ["# comment line"]
    for u, v in rng.integers(0, 300, size=(800, 2)):
        lines.append("%d\t%d" % (u + 1000, v + 1000) if rng.random() < 0.5 else "%d %d" % (u + 1000, v + 1000))
    lines.append("17 17")
    txt.write_text("\n".join(lines) + "\n")
    subprocess.run([os.path.join(REF, "edgelist2bin"), str(txt)], cwd=tmp_path, check=True,
                   capture_output=True)
    ref = (tmp_path / "edges.bin").read_bytes()
    out = tmp_path / "mine.bin"
    G.edgelist2bin(str(txt), str(out))
    assert out.read_bytes() == ref


@needs_ref
def test_normalize_matches_reference(tmp_path):
    src = tmp_path / "g.bin"
    rng = np.random.default_rng(6)
    pairs = [tuple(x) for x in rng.integers(0, 400, size=(2500, 2))] + [(35, 3), (9, 9), (-4, 7)]
    write_pairs(str(src), pairs)
    subprocess.run([os.path.join(REF, "normalize"), str(src)], cwd=tmp_path, check=True,
                   capture_output=True)
    ref = (tmp_path / "sorted_g.bin").read_bytes()
    out = tmp_path / "mine.bin"
    G.normalize(str(src), str(out))
    assert out.read_bytes() == ref


def test_direct_edgelist2bin_hand_case(tmp_path):
    """Directed writer: self-loops dropped, duplicates kept, u -> v only,
    vertices by non-increasing out-degree (a short list sorts stably),
    neighbours ascending by original id; the id cut-off keeps an edge only
    when an endpoint is below it."""
    src = tmp_path / "d.bin"
    write_pairs(str(src), [(0, 1), (0, 2), (2, 1), (1, 1), (3, 0), (0, 1)])
    out = tmp_path / "o.bin"
    G.direct_edgelist2bin(str(src), str(out))
    got = np.fromfile(str(out), dtype=np.int32).reshape(-1, 2).tolist()
    # order 0 (3 out-edges), 2 (1), 3 (1), 1 (0): new ids 0, 1, 2, 3
    assert got == [[0, 3], [0, 3], [0, 1], [1, 3], [2, 0]]
    G.direct_edgelist2bin(str(src), str(out), id_cutoff=1)   # drops (2,1)
    got = np.fromfile(str(out), dtype=np.int32).reshape(-1, 2).tolist()
    # order 0 (3), 3 (1), 1 (0), 2 (0): new ids 0, 1, 2, 3
    assert got == [[0, 2], [0, 2], [0, 3], [1, 0]]
    g = G.Graph.load_bin(str(out), directed=True)
    off, nbr, _ = g.csr()
    assert off.tolist() == [0, 3, 4, 4, 4] and nbr.tolist() == [2, 2, 3, 0]


@needs_ref
def test_direct_edgelist2bin_matches_reference(tmp_path):
    src = tmp_path / "g.bin"
    rng = np.random.default_rng(8)
    pairs = [tuple(x) for x in rng.integers(0, 300, size=(3000, 2))]
    cut = 5013066
    pairs += [(cut + 1, cut + 2), (cut - 1, cut + 3), (cut + 4, 5), (7, 7), (12, 12)]
    pairs += pairs[:200]                                    # duplicates
    write_pairs(str(src), pairs)
    subprocess.run([os.path.join(REF, "direct_edgelist2bin"), str(src)], cwd=tmp_path, check=True,
                   capture_output=True)
    ref = (tmp_path / "sorted_direct_g.bin").read_bytes()
    out = tmp_path / "mine.bin"
    G.direct_edgelist2bin(str(src), str(out))
    assert out.read_bytes() == ref
    # the CLI program writes the reference's output name
    sub = tmp_path / "cli"
    sub.mkdir()
    subprocess.run([os.path.join(ROOT, "glp_amd", "bin", "direct_edgelist2bin"), str(src)], cwd=sub,
                   check=True, capture_output=True)
    assert (sub / "sorted_direct_g.bin").read_bytes() == ref


def test_generators_deterministic_and_thread_independent():
    a = G.Graph.rmat(12, 16, seed=5, threads=1).csr()
    b = G.Graph.rmat(12, 16, seed=5, threads=4).csr()
    assert all(np.array_equal(x, y) for x, y in zip(a[:2], b[:2]))
    c = G.Graph.rmat(12, 16, seed=6).csr()
    assert not np.array_equal(a[1][:1000], c[1][:1000])
    p1 = G.Graph.planted(4096, 4, 12, 4, seed=3, threads=1).csr()
    p2 = G.Graph.planted(4096, 4, 12, 4, seed=3, threads=3).csr()
    assert all(np.array_equal(x, y) for x, y in zip(p1[:2], p2[:2]))


def test_planted_partition_shape():
    g, comm = G.Graph.planted(65536, 64, 12.0, 4.0, seed=1, with_communities=True)
    off, nbr, _ = g.csr()
    deg = np.diff(off)
    src = np.repeat(np.arange(len(deg)), deg)
    intra = (comm[src] == comm[nbr]).mean()
    assert abs(deg.mean() - 16) < 0.2                 # expected degree 12 + 4
    assert abs(intra - 0.75) < 0.01                   # 12 of 16 inside
    assert not np.any(src == nbr)                     # no self-loops
    pairs = src.astype(np.int64) * len(deg) + nbr
    assert len(np.unique(pairs)) == len(pairs)        # no duplicates


def test_rmat_shape():
    g = G.Graph.rmat(14, 16, seed=1)
    off, nbr, _ = g.csr()
    deg = np.diff(off)
    assert np.all(deg[:-1] >= deg[1:])                # degree-sorted ids
    src = np.repeat(np.arange(len(deg)), deg)
    assert not np.any(src == nbr)
    key = np.minimum(src, nbr).astype(np.int64) * len(deg) + np.maximum(src, nbr)
    assert np.all(np.bincount(np.unique(key, return_counts=True)[1]) [2:3] >= 0)
    # symmetric: every (u,v) has its (v,u)
    fwd = np.sort(src.astype(np.int64) * len(deg) + nbr)
    bwd = np.sort(nbr.astype(np.int64) * len(deg) + src)
    assert np.array_equal(fwd, bwd)


def test_lfr_shape():
    g, comm = G.Graph.lfr(1 << 14, seed=2, with_communities=True)
    off, nbr, _ = g.csr()
    deg = np.diff(off)
    src = np.repeat(np.arange(len(deg)), deg)
    mix = (comm[src] != comm[nbr]).mean()
    assert 0.2 < mix < 0.4                            # mu = 0.3
    assert deg.max() <= 512


def test_mix_params_ranges():
    for i in range(50):
        p = G.mix_params(7, i)
        assert 1000 <= p["n"] <= 262144
        assert p["k"] == (p["n"] + 999) // 1000
        d = p["deg_in"] + p["deg_out"]
        assert 4 <= d <= 64 and 0.1 <= p["deg_out"] / d <= 0.4
    assert G.mix_params(7, 3) == G.mix_params(7, 3)


def test_stream_matches_oracle_and_kat():
    s = G.stream_u32(0, 0, 0, 4)
    assert s[0] == 0x7B1DCDAF
    for seed, t in [(1, 1), (99, 5), (2 ** 20 + 3, 17)]:
        got = G.stream_u32(seed, t, 100, 50)
        want = [O.stream_u32(seed, t, 100 + i) for i in range(50)]
        assert got.tolist() == want


def test_session_without_device_fails_cleanly():
    if G.device_count() > 0:
        pytest.skip("a device is visible")
    g = G.Graph.rmat(8, 4, seed=1)
    with pytest.raises(G.GLPError) as e:
        G.run(g, 1)
    assert e.value.code == -10


@pytest.mark.parametrize("entries,reads", [(3, 1024), (0, 1024), (1 << 32, 1024), (1024, 0)])
def test_calib_random_read_rejects_bad_arguments(entries, reads):
    """The random-read calibration validates its arguments before touching
    the device: the table must be a power of two <= 2^31 entries."""
    with pytest.raises(G.GLPError) as e:
        G.calib_random_read(entries, reads, 1)
    assert e.value.code == -1


def test_graph_symmetry(tmp_path):
    """The frontier marks a changed vertex's out-neighbours, so it needs every
    edge u->v to have some v->u; loaded undirected graphs and the generators'
    graphs are symmetric by construction, others are checked (as sets)."""
    path = G.Graph.from_csr([0, 1, 2, 2], [1, 2])          # 0->1->2
    assert not path.symmetric()
    multi = G.Graph.from_csr([0, 2, 3], [1, 1, 0])          # 0->1 twice, 1->0 once
    assert multi.symmetric()
    assert not G.Graph.from_csr([0, 2, 3, 4], [1, 2, 0, 1]).symmetric()   # 0->2 without 2->0
    f = tmp_path / "g.bin"
    np.array([0, 1, 1, 2, 3, 1], dtype=np.int32).tofile(f)
    assert G.Graph.load_bin(str(f)).symmetric()
    assert not G.Graph.load_bin(str(f), directed=True).symmetric()
    np.array([0, 1, 1, 0], dtype=np.int32).tofile(f)       # both directions listed
    assert G.Graph.load_bin(str(f), directed=True).symmetric()
    g = G.Graph.rmat(10, 8, seed=3)
    assert g.symmetric()
    off, nbr, _ = g.csr()
    assert G.Graph.from_csr(off, nbr).symmetric()           # checked, not assumed
    assert g.degree_sorted()[0].symmetric()
