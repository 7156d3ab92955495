"""The allocation-free entry point (glp_run_noalloc) against glp_run, two
calls in a row (the second on a smaller graph, so stale contents of the
static array from the first call would show), bit-exact at every step."""
import numpy as np
import pytest

import glp_amd as G

pytestmark = pytest.mark.gpu


def check(g, steps, **kw):
    labels, dump, _ = G.run(g, steps, dump=True, **kw)
    got, gdump = G.run_noalloc(g, steps, dump=True, **kw)
    assert np.array_equal(gdump, dump)
    assert np.array_equal(got, labels)


@pytest.mark.parametrize("kw", [{}, {"method": G.LAYERED, "gamma": (1, 4), "weighted": True}])
def test_noalloc_matches_run_twice(kw):
    big = G.Graph.rmat(16, 16, seed=51, wmax=8)
    small = G.Graph.rmat(11, 8, seed=52, wmax=8)
    check(big, 6, **kw)
    check(small, 4, **kw)


def test_noalloc_hub_tier_and_limits():
    # small thresholds put the high-degree vertices in the hub tier, whose
    # scratch then lives in the static array too
    with G.tuning(pack_max_deg=8, wave_max_deg=16, block_max_deg=32, hub_chunk=256, part_keys=256):
        g = G.Graph.rmat(13, 16, seed=53)
        _, t = G.run(g, 1)
        assert t["tier_vertices"]["hub"] > 0
        check(g, 5)
        check(G.Graph.rmat(9, 8, seed=54), 3)
    too_big = G.Graph.rmat(21, 16, seed=55)   # CSR alone exceeds GLP_NOALLOC_BYTES
    with pytest.raises(G.GLPError) as e:
        G.run_noalloc(too_big, 1)
    assert e.value.code == -6
    with pytest.raises(G.GLPError) as e:
        G.run_noalloc(G.Graph.rmat(8, 8, seed=1), 1, method=G.SLPA)
    assert e.value.code == -7
