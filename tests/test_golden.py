"""Committed golden fixtures (tests/golden): the oracle must keep producing
them (CPU), and the GPU path must match them bit for bit (gpu)."""
import os

import numpy as np
import pytest

import glp_amd as G
import oracle_lib as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixture(name, weighted=False):
    path = os.path.join(GOLD, name + ".bin")
    with np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False) as z:
        data = {k: z[k] for k in z.files}
    return path, O.load_bin(path, weighted=weighted), data


def test_oracle_reproduces_fixtures():
    path, (off, nbr, _), d = fixture("planted_4096")
    assert np.array_equal(O.classic(off, nbr, 20), d["classic"])
    path, (off, nbr, _), d = fixture("rmat_s11")
    assert np.array_equal(O.classic(off, nbr, 8), d["classic"])
    for key in ("layered_0_1", "layered_1_4", "layered_1_1"):
        p, q = int(key.split("_")[1]), int(key.split("_")[2])
        assert np.array_equal(O.layered(off, nbr, 6, p, q)[0], d[key])
    path, (off, nbr, w), d = fixture("rmat_s10_w", weighted=True)
    assert np.array_equal(O.classic(off, nbr, 6, w=w), d["classic"])
    path, (off, nbr, _), d = fixture("lfr_2048")
    assert np.array_equal(O.slpa(off, nbr, seed=7, T=10), d["slpa_mem"])


def test_library_loader_matches_fixture_csr():
    for name, wt in (("planted_4096", False), ("rmat_s11", False), ("rmat_s10_w", True), ("lfr_2048", False)):
        path, (off, nbr, w), _ = fixture(name, wt)
        goff, gnbr, gw = G.Graph.load_bin(path, weighted=wt).csr()
        assert np.array_equal(off, goff) and np.array_equal(nbr, gnbr)
        if wt:
            assert np.array_equal(w, gw)


@pytest.mark.gpu
def test_gpu_matches_fixtures():
    path, _, d = fixture("planted_4096")
    _, dump, _ = G.run(G.Graph.load_bin(path), 20, dump=True)
    assert np.array_equal(dump, d["classic"])
    path, _, d = fixture("rmat_s11")
    g = G.Graph.load_bin(path)
    _, dump, _ = G.run(g, 8, dump=True)
    assert np.array_equal(dump, d["classic"])
    for key in ("layered_0_1", "layered_1_4", "layered_1_1"):
        p, q = int(key.split("_")[1]), int(key.split("_")[2])
        _, dump, _ = G.run(g, 6, dump=True, method=G.LAYERED, gamma=(p, q))
        assert np.array_equal(dump, d[key]), key
    path, _, d = fixture("rmat_s10_w", weighted=True)
    gw = G.Graph.load_bin(path, weighted=True)
    _, dump, _ = G.run(gw, 6, dump=True, weighted=True)
    assert np.array_equal(dump, d["classic"])
    _, dump, _ = G.run(gw, 5, dump=True, weighted=True, method=G.LAYERED, gamma=(1, 4))
    assert np.array_equal(dump, d["layered_1_4"])
    path, _, d = fixture("lfr_2048")
    with G.Session(G.Graph.load_bin(path), method=G.SLPA, seed=7, slpa_T=10) as s:
        s.step(10)
        assert np.array_equal(s.slpa_memory(11), d["slpa_mem"])
        dom, ns, sets = s.slpa_post(8)
    assert np.array_equal(dom, d["dominant"]) and np.array_equal(ns, d["nsets"])
    assert np.array_equal(sets, d["sets"])
