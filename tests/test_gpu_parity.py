"""GPU parity: the HIP path (libglp.so through its C ABI) against the CPU
oracle, bit-exact at every step."""
import numpy as np
import pytest

import glp_amd as G
import oracle_lib as O

pytestmark = pytest.mark.gpu


def gpu_steps(graph, steps, **kw):
    _, dump, _ = G.run(graph, steps, dump=True, **kw)
    return dump


def hand_graph(edges, n, weights=None):
    off, nbr, w = O.csr_from_edges(n, edges, weights)
    return G.Graph.from_csr(off, nbr, w), off, nbr, w


def assert_steps_equal(got, want):
    for t in range(len(want)):
        bad = np.flatnonzero(got[t] != want[t])
        assert len(bad) == 0, "step %d: %d labels differ, first v=%d gpu=%d oracle=%d" % (
            t + 1, len(bad), bad[0], got[t][bad[0]], want[t][bad[0]])


# ------------------------------------------------------------------ classic ----

def test_classic_closed_forms():
    g, off, nbr, _ = hand_graph([(0, 1), (1, 2), (0, 2), (3, 4), (4, 5), (3, 5), (2, 3)], 7)
    got = gpu_steps(g, 2)
    assert got[0].tolist() == [2, 2, 3, 5, 5, 4, 6]
    assert got[1].tolist() == [3, 3, 2, 5, 5, 5, 6]
    star, *_ = hand_graph([(0, i) for i in range(1, 6)], 6)
    got = gpu_steps(star, 2)
    assert got[0].tolist() == [5, 0, 0, 0, 0, 0] and got[1].tolist() == [0, 5, 5, 5, 5, 5]


def test_classic_planted_65536_twenty_steps():
    g = G.Graph.planted(65536, 64, 12.0, 4.0, seed=1)
    off, nbr, _ = g.csr()
    got = gpu_steps(g, 20)
    assert_steps_equal(got, O.classic(off, nbr, 20))


@pytest.mark.parametrize("scale,ef,sort", [(14, 16, True), (16, 16, True), (15, 8, False)])
def test_classic_rmat_all_tiers(scale, ef, sort):
    g = G.Graph.rmat(scale, ef, seed=scale * 7 + ef, degree_sort=sort)
    off, nbr, _ = g.csr()
    labels, dump, timing = G.run(g, 6, dump=True)
    assert_steps_equal(dump, O.classic(off, nbr, 6))
    if scale == 16:  # max degree ~1e4: packed, wave, block and hub tiers all run
        for tier in ("packed", "wave", "block", "hub"):
            assert timing["tier_vertices"][tier] > 0, tier


def tier_boundary_graph(seed=0):
    """Centres with degrees on every tier and hub-chunk boundary, attached to a
    ring of small communities so neighbour labels repeat and tie."""
    rng = np.random.default_rng(seed)
    degs = [1, 2, 31, 32, 33, 63, 64, 65, 255, 256, 257, 511, 512, 513, 2047, 2048, 2049,
            4095, 4096, 4097, 8191, 8192, 8193, 12289]
    base = len(degs)
    pool = 20000
    n = base + pool
    edges = []
    for i, d in enumerate(degs):
        nb = rng.choice(pool, size=d, replace=False) + base
        edges += [(i, int(x)) for x in nb]
    # background: pool vertices in communities of 50 with sparse ties
    for c in range(pool // 50):
        members = base + c * 50 + np.arange(50)
        for _ in range(120):
            a, b = rng.choice(members, size=2, replace=False)
            edges.append((int(a), int(b)))
    return n, edges


def test_classic_tier_boundaries():
    n, edges = tier_boundary_graph()
    g, off, nbr, _ = hand_graph(edges, n)
    assert_steps_equal(gpu_steps(g, 5), O.classic(off, nbr, 5))


@pytest.mark.parametrize("half,pack,lanes", [(0, 32, 16), (48, 32, 16), (64, 32, 16), (64, 0, 16),
                                             (48, 32, 32), (64, 32, 32), (64, 0, 32)])
def test_half_wave_tier(half, pack, lanes):
    """Degrees in (pack_max_deg, half_max_deg] run 64/lanes per wavefront
    (16 or 32 lanes per vertex); half_max_deg=0 sends them through the
    one-vertex wave path.  pack_max_deg=0 puts degrees 1..64 in the half
    tier.  Classic unweighted and weighted, and layered, must match the
    oracle either way; a planted-partition graph run to convergence takes
    the majority shortcut and the lane merging."""
    with G.tuning(pack_max_deg=pack, half_max_deg=half, half_lanes=lanes):
        gp = G.Graph.planted(6000, 6, 28.0, 8.0, seed=35)
        offp, nbrp, _ = gp.csr()
        _, dumpp, _ = G.run(gp, 12, dump=True)
        assert_steps_equal(dumpp, O.classic(offp, nbrp, 12))
        g = G.Graph.rmat(14, 16, seed=33)
        off, nbr, _ = g.csr()
        _, dump, timing = G.run(g, 6, dump=True)
        assert_steps_equal(dump, O.classic(off, nbr, 6))
        gw = G.Graph.rmat(13, 16, seed=34, wmax=8)
        offw, nbrw, w = gw.csr()
        _, dumpw, _ = G.run(gw, 5, dump=True, weighted=True)
        assert_steps_equal(dumpw, O.classic(offw, nbrw, 5, w=w))
        _, duml, _ = G.run(gw, 4, dump=True, method=G.LAYERED, gamma=(1, 4), weighted=True)
        assert_steps_equal(duml, O.layered(offw, nbrw, 4, 1, 4, w=w)[0])
        n, edges = tier_boundary_graph(seed=3)
        gb, offb, nbrb, _ = hand_graph(edges, n)
        assert_steps_equal(gpu_steps(gb, 4), O.classic(offb, nbrb, 4))


@pytest.mark.parametrize("limit", [0, 8])
def test_hub_partition_and_halving(limit):
    """Tiny thresholds push most vertices through the hub passes; part_limit=8
    forces nearly every hash range to overflow and be halved."""
    with G.tuning(pack_max_deg=8, wave_max_deg=16, block_max_deg=32, hub_chunk=256,
                  part_keys=256, part_limit=limit):
        g = G.Graph.rmat(14, 16, seed=21)
        off, nbr, _ = g.csr()
        labels, dump, timing = G.run(g, 5, dump=True)
    assert timing["tier_vertices"]["hub"] > 1000
    assert_steps_equal(dump, O.classic(off, nbr, 5))


def test_hub_weighted_small_tiers():
    with G.tuning(pack_max_deg=4, wave_max_deg=8, block_max_deg=16, hub_chunk=256, part_keys=256):
        g = G.Graph.rmat(13, 16, seed=22, wmax=8)
        off, nbr, w = g.csr()
        got = gpu_steps(g, 4, weighted=True)
    assert_steps_equal(got, O.classic(off, nbr, 4, w=w))


def test_classic_multigraph_and_isolated():
    rng = np.random.default_rng(1)
    n = 3000
    e = rng.integers(0, n // 2, size=(20000, 2))          # upper half isolated
    e = np.concatenate([e, e[:5000]])                      # duplicates
    g, off, nbr, _ = hand_graph([tuple(x) for x in e], n)
    assert_steps_equal(gpu_steps(g, 4), O.classic(off, nbr, 4))


def test_classic_weighted():
    g = G.Graph.rmat(13, 16, seed=5, wmax=8)
    off, nbr, w = g.csr()
    assert_steps_equal(gpu_steps(g, 5, weighted=True), O.classic(off, nbr, 5, w=w))


def test_neighbour_order_invariance():
    g = G.Graph.rmat(13, 16, seed=9)
    off, nbr, _ = g.csr()
    rng = np.random.default_rng(0)
    nbr2 = nbr.copy()
    for v in range(len(off) - 1):
        rng.shuffle(nbr2[off[v]:off[v + 1]])
    g2 = G.Graph.from_csr(off, nbr2)
    assert np.array_equal(gpu_steps(g, 4), gpu_steps(g2, 4))


@pytest.mark.parametrize("kw", [{}, {"gather": True}, {"frontier": True},
                                {"frontier": True, "gather": True}])
def test_empty_and_edgeless(kw):
    g = G.Graph.from_csr(np.zeros(1, np.int32), np.zeros(0, np.int32))
    labels, _ = G.run(g, 3, **kw)
    assert labels.shape == (0,)
    g = G.Graph.from_csr(np.zeros(11, np.int32), np.zeros(0, np.int32))
    labels, _ = G.run(g, 3, **kw)
    assert labels.tolist() == list(range(10))


def test_separate_gather_matches_fused():
    g = G.Graph.rmat(15, 16, seed=3)
    off, nbr, _ = g.csr()
    want = O.classic(off, nbr, 5)
    assert_steps_equal(gpu_steps(g, 5, gather=True), want)


def test_graph_replay_matches_eager_and_changes():
    g = G.Graph.rmat(14, 16, seed=4)
    off, nbr, _ = g.csr()
    want = O.classic(off, nbr, 7)
    with G.Session(g) as s:
        s.step(3)
        s.sync()
        assert np.array_equal(s.labels(), want[2])
        for k in range(4):                      # replays with eager syncs between
            s.step(1, graph=True)
            s.sync()
            assert np.array_equal(s.labels(), want[3 + k])
        ch = s.changes()
    prev = np.vstack([np.arange(len(off) - 1, dtype=np.int32), want[:-1]])
    assert ch.tolist() == [int((want[t] != prev[t]).sum()) for t in range(7)]


@pytest.mark.parametrize("scale,tiny", [(14, False), (15, False), (13, True)])
def test_frontier_equals_dense(scale, tiny):
    """The frontier recomputes only vertices with a changed neighbour once a
    step changes at most n/32 labels; every step must equal the dense run."""
    kw = dict(pack_max_deg=8, wave_max_deg=16, block_max_deg=32, hub_chunk=256,
              part_keys=256) if tiny else {}
    with G.tuning(**kw):
        g = G.Graph.rmat(scale, 16, seed=50 + scale)
        off, nbr, _ = g.csr()
        want = O.classic(off, nbr, 14)
        _, dump, _ = G.run(g, 14, dump=True, frontier=True)
        assert_steps_equal(dump, want)
        with G.Session(g, frontier=True) as s:
            s.step(2)
            s.step(12, graph=True)
            assert np.array_equal(s.labels(), want[-1])
            ch = s.changes()
    prev = np.vstack([np.arange(len(off) - 1, dtype=np.int32), want[:-1]])
    assert ch.tolist() == [int((want[t] != prev[t]).sum()) for t in range(14)]
    assert ch[-1] < (len(off) - 1) // 32   # the frontier was in use at the end


def _push_steps(off, nbr, want, back=1):
    """Steps (1-based) whose frontier gather takes the push form: once adj
    has been gathered `back` times, when the sources whose label in the step
    before differs from their label `back` steps earlier have at most m/8
    in-edges (the gather's own rule; back = 2 is the two-step lookahead)."""
    n, m = len(off) - 1, len(nbr)
    indeg = np.bincount(nbr, minlength=n)
    labs = [np.arange(n, dtype=np.int32)] + list(want)
    return [t for t in range(back + 1, len(labs))
            if int(indeg[labs[t - 1] != labs[t - 1 - back]].sum()) <= m // 8]


@pytest.mark.parametrize("scale,tiny,weighted,frontier", [(14, False, False, 1), (13, True, False, 1),
                                                          (13, False, True, 1), (16, False, False, 1),
                                                          (14, False, False, 2), (13, True, False, 2),
                                                          (13, False, True, 2), (16, False, False, 2)])
def test_frontier_gather_equals_dense(scale, tiny, weighted, frontier):
    """Separate gather with the frontier: after a step whose changed sources
    have few in-edges the gather pushes their labels through the inverse
    index, otherwise it rewrites every edge; every step must equal the
    oracle, also under graph replay and across a checkpoint (which forces a
    full gather).  Sources with more in-edges than one push entry holds
    (256) occur at every scale here; at scale 16 some have thousands.
    frontier=2 adds the two-step lookahead (two adj buffers, changes
    against two steps back)."""
    kw = dict(pack_max_deg=8, wave_max_deg=16, block_max_deg=32, hub_chunk=256,
              part_keys=256) if tiny else {}
    with G.tuning(**kw):
        g = G.Graph.rmat(scale, 16, seed=70 + scale, wmax=8 if weighted else 0)
        off, nbr, w = g.csr()
        want = O.classic(off, nbr, 16, w=w if weighted else None)
        _, dump, _ = G.run(g, 16, dump=True, frontier=frontier, gather=True, weighted=weighted)
        assert_steps_equal(dump, want)
        with G.Session(g, frontier=frontier, gather=True, weighted=weighted) as s:
            s.step(3)
            s.step(8, graph=True)
            assert np.array_equal(s.labels(), want[10])
            state = s.save()
        with G.Session(g, frontier=frontier, gather=True, weighted=weighted) as s2:
            s2.load(state)
            s2.step(5)
            assert np.array_equal(s2.labels(), want[15])
    push = _push_steps(off, nbr, want, back=frontier)
    assert len(push) >= 4 and push[-1] == 16, push   # the push form was in use
    assert int(np.diff(off).max()) > 256


def test_frontier_needs_symmetric_graph():
    """A changed vertex marks its out-neighbours, which are all the vertices
    that read its label only when every edge has its reverse: the frontier
    refuses an asymmetric graph and runs on a symmetric one built from
    caller arrays (checked, not assumed)."""
    g = G.Graph.from_csr([0, 1, 2, 2], [1, 2])   # 0->1->2
    with pytest.raises(G.GLPError):
        G.run(g, 2, frontier=True)
    with pytest.raises(G.GLPError):
        G.run(g, 2, frontier=True, gather=True)
    off, nbr, _ = G.Graph.rmat(12, 8, seed=61).csr()
    gs = G.Graph.from_csr(off, nbr)
    _, dump, _ = G.run(gs, 10, dump=True, frontier=True)
    assert_steps_equal(dump, O.classic(off, nbr, 10))


def test_frontier_rejected_for_layered():
    g = G.Graph.rmat(10, 8, seed=1)
    with pytest.raises(G.GLPError):
        G.run(g, 2, method=G.LAYERED, gamma=(1, 4), frontier=True)


def test_checkpoint_resume():
    g = G.Graph.rmat(13, 16, seed=6)
    off, nbr, _ = g.csr()
    want = O.classic(off, nbr, 6)
    with G.Session(g) as s:
        s.step(3)
        state = s.save()
    with G.Session(g) as s2:
        s2.load(state)
        s2.step(3)
        assert np.array_equal(s2.labels(), want[5])


def test_profile_mode_matches():
    g = G.Graph.rmat(14, 16, seed=8)
    off, nbr, _ = g.csr()
    labels, t = G.run(g, 3, profile=True)
    assert np.array_equal(labels, O.classic(off, nbr, 3)[-1])
    assert t["tier_launches"]["packed"] == 3
