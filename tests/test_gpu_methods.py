"""GPU parity of the non-classic methods (layered, SLPA, count-min, frequent,
buffer, user hooks) and of the batched and served forms, against the oracle
and against solo runs."""
import numpy as np
import pytest

import glp_amd as G
import oracle_lib as O

pytestmark = pytest.mark.gpu


def assert_steps_equal(got, want, what=""):
    for t in range(len(want)):
        bad = np.flatnonzero(got[t] != want[t])
        assert len(bad) == 0, "%s step %d: %d differ, first v=%d gpu=%d oracle=%d" % (
            what, t + 1, len(bad), bad[0], got[t][bad[0]], want[t][bad[0]])


@pytest.fixture(scope="module")
def rmat14():
    g = G.Graph.rmat(14, 16, seed=31)
    return g, g.csr()


@pytest.fixture(scope="module")
def rmat13w():
    g = G.Graph.rmat(13, 16, seed=32, wmax=8)
    return g, g.csr()


# ------------------------------------------------------------------ layered ----

@pytest.mark.parametrize("gamma", [(0, 1), (1, 4), (1, 1)])
def test_layered_matches_oracle(rmat14, gamma):
    g, (off, nbr, _) = rmat14
    _, dump, _ = G.run(g, 6, dump=True, method=G.LAYERED, gamma=gamma)
    want, vols = O.layered(off, nbr, 6, gamma[0], gamma[1])
    assert_steps_equal(dump, want, "layered%s" % (gamma,))
    with G.Session(g, method=G.LAYERED, gamma=gamma) as s:
        s.step(6)
        assert np.array_equal(s.volume(), vols[-1])


def test_layered_weighted(rmat13w):
    g, (off, nbr, w) = rmat13w
    _, dump, _ = G.run(g, 5, dump=True, method=G.LAYERED, gamma=(1, 4), weighted=True)
    want, _ = O.layered(off, nbr, 5, 1, 4, w=w)
    assert_steps_equal(dump, want, "layered weighted")


def test_layered_hub_path():
    with G.tuning(pack_max_deg=8, wave_max_deg=16, block_max_deg=32, hub_chunk=256, part_keys=256):
        g = G.Graph.rmat(13, 16, seed=33)
        off, nbr, _ = g.csr()
        _, dump, _ = G.run(g, 4, dump=True, method=G.LAYERED, gamma=(1, 1))
    assert_steps_equal(dump, O.layered(off, nbr, 4, 1, 1)[0], "layered hubs")


def _layered_switch_oracle(off, nbr, plan, w=None):
    """Layered steps with gamma changing between them: plan = [(steps, p, q)]."""
    n = len(off) - 1
    L = np.arange(n, dtype=np.int32)
    V = np.ones(n, dtype=np.int32)
    offc, nbrc, wc = O._csr(off, nbr, w)
    out = []
    for steps, p, q in plan:
        for _ in range(steps):
            nl = np.zeros(max(n, 1), dtype=np.int32)
            nv = np.zeros(max(n, 1), dtype=np.int32)
            O._o.orc_layered_step(n, O._i(offc), O._i(nbrc), O._u(wc), p, q, O._i(L), O._i(V), O._i(nl),
                                  O._i(nv))
            L, V = nl[:n], nv[:n]
            out.append(L)
    return np.array(out)


@pytest.mark.parametrize("gamma", [(1, 4), (3, 2)])
def test_layered_slot_scores_and_volume_reads_agree(rmat14, gamma):
    """gamma > 0: the fused steps score in the table slots from (label,
    volume) records; the gathered form reads each candidate's volume.  Both
    equal the oracle at every step."""
    g, (off, nbr, _) = rmat14
    want = O.layered(off, nbr, 5, gamma[0], gamma[1])[0]
    _, dump, _ = G.run(g, 5, dump=True, method=G.LAYERED, gamma=gamma)
    assert_steps_equal(dump, want, "slot-scored")
    _, dump_g, _ = G.run(g, 5, dump=True, method=G.LAYERED, gamma=gamma, gather=True)
    assert_steps_equal(dump_g, want, "gathered")


def test_layered_slot_scores_every_tier_weighted():
    """Tiny thresholds put vertices in every tier (hub chunks count
    frequencies, every other table scores); weights up to 8."""
    with G.tuning(pack_max_deg=8, half_max_deg=12, wave_max_deg=16, block_max_deg=32, hub_chunk=256,
                  part_keys=256):
        g = G.Graph.rmat(13, 16, seed=35, wmax=8)
        off, nbr, w = g.csr()
        _, dump, t = G.run(g, 4, dump=True, method=G.LAYERED, gamma=(1, 2), weighted=True)
    assert min(t["tier_vertices"][k] for k in ("packed", "wave", "block", "hub")) > 0
    assert_steps_equal(dump, O.layered(off, nbr, 4, 1, 2, w=w)[0], "slot-scored tiers")


def test_layered_set_gamma_switches_scoring(rmat14):
    """A session that starts at gamma 0 (frequency tables) and moves to 1/4
    (slot scores) and back matches the oracle run with the same schedule."""
    g, (off, nbr, _) = rmat14
    want = _layered_switch_oracle(off, nbr, [(2, 0, 1), (3, 1, 4), (2, 0, 1)])
    with G.Session(g, method=G.LAYERED, gamma=(0, 1)) as s:
        s.step(2)
        s.sync()
        got = [s.labels()]
        s.set_gamma(1, 4)
        s.step(3)
        got.append(s.labels())
        s.set_gamma(0, 1)
        s.step(2)
        got.append(s.labels())
    assert np.array_equal(got[0], want[1]) and np.array_equal(got[1], want[4]) and np.array_equal(got[2], want[6])


def test_layered_range_check():
    g = G.Graph.rmat(10, 8, seed=1)
    with pytest.raises(G.GLPError):
        G.run(g, 1, method=G.LAYERED, gamma=(1, 3))   # denominator not a power of two


# --------------------------------------------------------------------- SLPA ----

@pytest.mark.parametrize("T", [1, 7])
def test_slpa_memory_and_post(rmat14, T):
    g, (off, nbr, _) = rmat14
    mem_want = O.slpa(off, nbr, seed=11, T=T)
    with G.Session(g, method=G.SLPA, seed=11, slpa_T=T, slpa_r=(1, 10)) as s:
        s.step(T)
        mem = s.slpa_memory(T + 1)
        dom, ns, sets = s.slpa_post(kmax=8)
    assert_steps_equal(mem, mem_want, "slpa memory")
    dw, nw, sw = O.slpa_post(mem_want, 1, 10, 8)
    assert np.array_equal(dom, dw) and np.array_equal(ns, nw) and np.array_equal(sets, sw)
    labels, _ = G.run(g, T, method=G.SLPA, seed=11)
    assert np.array_equal(labels, dw)


def test_slpa_weighted(rmat13w):
    """Weighted SLPA: listeners sum the shown labels' edge weights."""
    g, (off, nbr, w) = rmat13w
    with G.Session(g, method=G.SLPA, seed=5, slpa_T=6, weighted=True) as s:
        s.step(6)
        mem = s.slpa_memory(7)
    assert_steps_equal(mem, O.slpa(off, nbr, seed=5, T=6, w=w), "slpa weighted")


def test_slpa_hand_trace():
    """The hand-traced three-step SLPA case of the oracle's tests (seed 3)."""
    g = G.Graph.from_edges(5, [(0, 1), (1, 2), (2, 3), (3, 4), (1, 3)])
    with G.Session(g, method=G.SLPA, seed=3, slpa_T=3) as s:
        s.step(3)
        mem = s.slpa_memory(4)
    assert mem.tolist() == [[0, 1, 2, 3, 4], [1, 3, 3, 4, 3], [1, 3, 3, 3, 3], [3, 4, 4, 3, 4]]


def test_slpa_lfr_graph():
    g = G.Graph.lfr(1 << 13, seed=5)
    off, nbr, _ = g.csr()
    with G.Session(g, method=G.SLPA, seed=3, slpa_T=10) as s:
        s.step(10)
        mem = s.slpa_memory(11)
    assert_steps_equal(mem, O.slpa(off, nbr, seed=3, T=10), "slpa lfr")


# ------------------------------------------------------- approximate counting ----

@pytest.mark.parametrize("rows,width,aliased", [(1, 64, 0), (3, 200, 1), (3, 200, 0), (4, 50, 1),
                                                (5, 500, 0), (2, 7, 0)])
def test_countmin_matches_oracle(rmat14, rows, width, aliased):
    g, (off, nbr, _) = rmat14
    _, dump, _ = G.run(g, 3, dump=True, method=G.COUNTMIN, cm_rows=rows, cm_width=width,
                       cm_aliased=aliased)
    prev = np.arange(len(off) - 1, dtype=np.int32)
    for t in range(3):
        prev = O.countmin_step(off, nbr, prev, rows, width, aliased)
        assert np.array_equal(dump[t], prev), "step %d" % (t + 1)


@pytest.mark.parametrize("a,b,ca,rows,aliased,want", [
    # the oracle's hand-derived width-4 collision cases (test_oracle.py)
    (1, 4, 2, 3, 0, 1), (1, 4, 2, 1, 0, 4),
    (2, 17, 2, 3, 0, 2), (2, 17, 2, 3, 1, 17), (2, 17, 2, 4, 1, 17),
    (1, 15, 1, 3, 0, 15), (1, 15, 1, 3, 1, 1), (1, 15, 1, 4, 1, 15)])
def test_countmin_hand_collisions(a, b, ca, rows, aliased, want):
    """A star centre 0 sees leaf a on ca parallel edges and leaf b on one; at
    step 1 (labels = ids) its count-min winner at width 4 is argued by hand
    from the row hashes' cells."""
    g = G.Graph.from_edges(b + 1, [(0, a)] * ca + [(0, b)])
    labels, _ = G.run(g, 1, method=G.COUNTMIN, cm_rows=rows, cm_width=4, cm_aliased=aliased)
    assert labels[0] == want
    off, nbr, _ = g.csr()
    assert np.array_equal(labels, O.countmin_step(off, nbr, np.arange(b + 1, dtype=np.int32), rows, 4,
                                                  aliased))


def test_countmin_largest_sketch():
    """rows x width = 5 x 4096 (80 KB of LDS, the largest accepted shape), in
    the one-wavefront and the 1024-thread (degree > 1024) launches."""
    g = G.Graph.rmat(13, 16, seed=34)
    off, nbr, _ = g.csr()
    assert np.diff(off).max() > 1024
    labels, _ = G.run(g, 2, method=G.COUNTMIN, cm_rows=5, cm_width=4096)
    prev = np.arange(len(off) - 1, dtype=np.int32)
    for _ in range(2):
        prev = O.countmin_step(off, nbr, prev, 5, 4096, 0)
    assert np.array_equal(labels, prev)
    with pytest.raises(G.GLPError):
        G.run(g, 1, method=G.COUNTMIN, cm_rows=5, cm_width=4097)


def countmin_hub_graph(weighted, seed=5, degs=(8192, 8193, 16384, 20001, 3000)):
    """Hubs of 8192 (one chunk), 8193, 16384 and 20001 edges (a part-filled
    last chunk) on a pool of communities, so that labels repeat among a hub's
    neighbours from step 2 on."""
    rng = np.random.default_rng(seed)
    degs = list(degs)
    base, pool = len(degs), 24000
    edges = []
    for i, d in enumerate(degs):
        edges += [(i, int(x)) for x in rng.choice(pool, size=d, replace=False) + base]
    for c in range(pool // 40):
        members = base + c * 40 + np.arange(40)
        for _ in range(90):
            a, b = rng.choice(members, size=2, replace=False)
            edges.append((int(a), int(b)))
    wts = rng.integers(1, 100, size=len(edges)).tolist() if weighted else None
    return O.csr_from_edges(base + pool, edges, wts)


@pytest.mark.parametrize("rows,width,aliased,weighted", [(3, 200, 1, False), (4, 50, 1, True), (1, 64, 0, False),
                                                         (5, 4096, 0, False), (2, 7, 0, True)])
def test_countmin_split_hubs(rows, width, aliased, weighted):
    """Vertices above 8192 edges are counted in 8192-edge chunks on separate
    workgroups whose sketches add up in HBM; the labels equal the oracle's,
    which counts each vertex in one pass."""
    off, nbr, w = countmin_hub_graph(weighted)
    g = G.Graph.from_csr(off, nbr, w)
    assert np.diff(off)[:5].tolist() == [8192, 8193, 16384, 20001, 3000]
    _, dump, _ = G.run(g, 3, dump=True, method=G.COUNTMIN, cm_rows=rows, cm_width=width, cm_aliased=aliased,
                       weighted=weighted)
    prev = np.arange(len(off) - 1, dtype=np.int32)
    for t in range(3):
        prev = O.countmin_step(off, nbr, prev, rows, width, aliased, w=w if weighted else None)
        assert np.array_equal(dump[t], prev), "step %d" % (t + 1)


TIER_EDGES = (16, 17, 32, 33, 1024, 1025, 4096, 4097, 8192, 8193)


@pytest.mark.parametrize("method", [G.COUNTMIN, G.FREQUENT, G.BUFFER])
def test_counting_tier_boundaries(method):
    """Centres with degrees on each side of every counting tier's bound
    (count-min 16 / 1024 / 4096 / 8192, tables 32): a thread's keys held in
    registers must cover its tier's largest degree exactly."""
    off, nbr, _ = countmin_hub_graph(False, seed=9, degs=TIER_EDGES)
    assert np.diff(off)[:len(TIER_EDGES)].tolist() == list(TIER_EDGES)
    g = G.Graph.from_csr(off, nbr)
    kw = dict(cm_rows=4, cm_width=50, cm_aliased=1) if method == G.COUNTMIN else dict(table_cap=8)
    _, dump, _ = G.run(g, 3, dump=True, method=method, **kw)
    prev = np.arange(len(off) - 1, dtype=np.int32)
    for t in range(3):
        if method == G.COUNTMIN:
            prev = O.countmin_step(off, nbr, prev, 4, 50, 1)
        else:
            prev = O.table_step(off, nbr, prev, 8, int(method == G.BUFFER))
        assert np.array_equal(dump[t], prev), "step %d" % (t + 1)


@pytest.mark.parametrize("method", [G.COUNTMIN, G.FREQUENT, G.BUFFER])
def test_counting_tiny_and_edgeless_graphs(method):
    """A single vertex, edge-less graphs, one edge and a 4-cycle: the
    counting methods' vertex lists are empty or hold only isolated vertices,
    and every label matches the oracle (an edge-less vertex keeps its own)."""
    kw = dict(cm_rows=3, cm_width=200, cm_aliased=1) if method == G.COUNTMIN else dict(table_cap=4)
    for n, edges in ((1, []), (3, []), (2, [(0, 1)]), (4, [(0, 1), (1, 2), (2, 3), (3, 0)])):
        g = G.Graph.from_edges(n, edges)
        off, nbr, _ = g.csr()
        labels, _ = G.run(g, 3, method=method, **kw)
        prev = np.arange(n, dtype=np.int32)
        for _ in range(3):
            if method == G.COUNTMIN:
                prev = O.countmin_step(off, nbr, prev, 3, 200, 1)
            else:
                prev = O.table_step(off, nbr, prev, 4, int(method == G.BUFFER))
        assert np.array_equal(labels, prev), (n, edges)


def test_countmin_weighted(rmat13w):
    g, (off, nbr, w) = rmat13w
    labels, _ = G.run(g, 2, method=G.COUNTMIN, cm_rows=3, cm_width=100, cm_aliased=1, weighted=True)
    prev = np.arange(len(off) - 1, dtype=np.int32)
    for _ in range(2):
        prev = O.countmin_step(off, nbr, prev, 3, 100, 1, w=w)
    assert np.array_equal(labels, prev)


@pytest.mark.parametrize("method,buffer", [(G.FREQUENT, 0), (G.BUFFER, 1)])
@pytest.mark.parametrize("cap", [1, 4, 16, 64])
def test_tables_match_oracle(rmat14, method, buffer, cap):
    g, (off, nbr, _) = rmat14
    _, dump, _ = G.run(g, 3, dump=True, method=method, table_cap=cap)
    prev = np.arange(len(off) - 1, dtype=np.int32)
    for t in range(3):
        prev = O.table_step(off, nbr, prev, cap, buffer)
        assert np.array_equal(dump[t], prev), "step %d" % (t + 1)


@pytest.mark.parametrize("cap", [3, 16, 64])
def test_tables_weighted(rmat13w, cap):
    g, (off, nbr, w) = rmat13w
    for method, buf in ((G.FREQUENT, 0), (G.BUFFER, 1)):
        labels, _ = G.run(g, 2, method=method, table_cap=cap, weighted=True)
        prev = np.arange(len(off) - 1, dtype=np.int32)
        for _ in range(2):
            prev = O.table_step(off, nbr, prev, cap, buf, w=w)
        assert np.array_equal(labels, prev)


# --------------------------------------------------------------- user hooks ----

def test_user_hooks_are_classic(rmat14):
    g, (off, nbr, _) = rmat14
    _, dump, _ = G.run(g, 5, dump=True, method=G.USER)
    assert_steps_equal(dump, O.classic(off, nbr, 5), "user hooks")
    _, dump2, _ = G.run(g, 5, dump=True, method=G.USER, gather=True)
    assert np.array_equal(dump, dump2)


# ------------------------------------------------------------------- seeded ----

@pytest.mark.parametrize("frac,gather", [(0.05, False), (0.002, False), (0.05, True)])
def test_seeded_matches_oracle(rmat14, frac, gather):
    g, (off, nbr, _) = rmat14
    seeds = (np.random.default_rng(41).random(len(off) - 1) < frac).astype(np.uint8)
    _, dump, _ = G.run(g, 8, dump=True, method=G.SEEDED, seeds=seeds, gather=gather)
    assert_steps_equal(dump, O.seeded(off, nbr, 8, seeds), "seeded")


def test_seeded_weighted_and_every_tier(rmat13w):
    g, (off, nbr, w) = rmat13w
    seeds = (np.random.default_rng(42).random(len(off) - 1) < 0.03).astype(np.uint8)
    _, dump, _ = G.run(g, 6, dump=True, method=G.SEEDED, seeds=seeds, weighted=True)
    assert_steps_equal(dump, O.seeded(off, nbr, 6, seeds, w=w), "seeded weighted")
    # tiny thresholds: most vertices (and some seeds) in the block and hub tiers
    with G.tuning(pack_max_deg=4, half_max_deg=8, wave_max_deg=16, block_max_deg=32, hub_chunk=256,
                  part_keys=256):
        gt = G.Graph.rmat(14, 16, seed=43)
        offt, nbrt, _ = gt.csr()
        deg = np.diff(offt)
        seeds = (np.random.default_rng(44).random(len(deg)) < 0.02).astype(np.uint8)
        seeds[np.argsort(-deg)[:3]] = 1                       # hubs among the seeds
        _, dump, t = G.run(gt, 6, dump=True, method=G.SEEDED, seeds=seeds)
    assert t["tier_vertices"]["hub"] > 0
    assert_steps_equal(dump, O.seeded(offt, nbrt, 6, seeds), "seeded tiers")


def test_seeded_session_replay_checkpoint_and_errors(rmat14):
    g, (off, nbr, _) = rmat14
    n = len(off) - 1
    seeds = (np.random.default_rng(45).random(n) < 0.01).astype(np.uint8)
    want = O.seeded(off, nbr, 9, seeds)
    with G.Session(g, method=G.SEEDED) as s:
        with pytest.raises(G.GLPError):
            s.step(1)                                        # no seeds yet
        s.set_seeds(seeds)
        assert np.array_equal(s.labels(), O.seeded_labels0(seeds))
        s.step(2)
        s.step(3, graph=True)
        assert np.array_equal(s.labels(), want[4])
        state = s.save()
        with pytest.raises(G.GLPError):
            s.set_seeds(seeds[:-1])                          # wrong length
    with G.Session(g, method=G.SEEDED) as s2:
        s2.set_seeds(seeds)
        s2.load(state)
        s2.step(4)
        assert np.array_equal(s2.labels(), want[8])
    with G.Session(g) as s3:
        with pytest.raises(G.GLPError):
            s3.set_seeds(seeds)                              # not a seeded session
    with pytest.raises(G.GLPError):
        G.run(g, 2, method=G.SEEDED, seeds=seeds, frontier=True)
    with pytest.raises(G.GLPError):
        G.Batch(mix_graphs(2, 1), method=G.SEEDED)


# --------------------------------------------------- persistent small tiers ----

@pytest.mark.parametrize("method,kw", [
    (G.CLASSIC, {}), (G.CLASSIC, {"weighted": True}), (G.CLASSIC, {"gather": True}),
    (G.LAYERED, {"gamma": (0, 1)}), (G.LAYERED, {"gamma": (1, 4), "weighted": True}),
    (G.SLPA, {"seed": 11}), (G.USER, {}), (G.SEEDED, {"seeds": "5%"})])
def test_fused_small_tiers_equal_separate_launches(method, kw):
    """fused_small=1 (packed, half and wave tiers in one persistent launch
    whose waves take units from a counter) gives every step's labels of one
    launch per tier (fused_small=0), with every tier populated."""
    kw = dict(kw)
    dumps = []
    for fused in (0, 1):
        with G.tuning(fused_small=fused):
            g = G.Graph.rmat(13, 16, seed=32, wmax=8)   # the tuning is read at the graph's upload
            if isinstance(kw.get("seeds"), str):
                kw["seeds"] = (np.random.default_rng(43).random(g.n) < 0.05).astype(np.uint8)
            _, dump, t = G.run(g, 5, dump=True, method=method, **kw)
            assert all(t["tier_vertices"][k] > 0 for k in ("packed", "wave", "block"))
            dumps.append(dump)
    assert_steps_equal(dumps[1], dumps[0], "fused small tiers")


# ------------------------------------------------------------------ batched ----

def mix_graphs(count, seed, nmax=6000):
    gs = []
    for i in range(count):
        p = G.mix_params(seed, i, nmin=200, nmax=nmax)
        gs.append(G.Graph.planted(p["n"], p["k"], p["deg_in"], p["deg_out"], seed=p["seed"]))
    return gs


def oracle_labels(g, steps, method, kw, weighted=False):
    off, nbr, w = g.csr()
    w = w if weighted else None
    if method == G.CLASSIC:
        return O.classic(off, nbr, steps, w=w)[-1]
    if method == G.LAYERED:
        return O.layered(off, nbr, steps, kw["gamma"][0], kw["gamma"][1], w=w)[0][-1]
    prev = np.arange(len(off) - 1, dtype=np.int32)
    for _ in range(steps):
        prev = O.table_step(off, nbr, prev, kw["table_cap"], method == G.BUFFER, w=w)
    return prev


def oracle_labels_csr(off, nbr, w, steps, method, kw):
    if method == G.CLASSIC:
        return O.classic(off, nbr, steps, w=w)[-1]
    return O.layered(off, nbr, steps, kw["gamma"][0], kw["gamma"][1], w=w)[0][-1]


# small thresholds put the RMAT graph's high-degree vertices in the hub tier
# (degree > 32) and the planted graphs' in the block tier
HUB_TUNING = dict(pack_max_deg=8, wave_max_deg=16, block_max_deg=32, hub_chunk=256, part_keys=256)


@pytest.mark.parametrize("method,kw", [(G.CLASSIC, {}), (G.LAYERED, {"gamma": (1, 4)}),
                                       (G.FREQUENT, {"table_cap": 8})])
@pytest.mark.parametrize("tuned", [False, True, "fused"])
def test_batch_equals_alone_and_oracle(method, kw, tuned):
    """Every graph of a batch (union stepped in one launch per tier, its HIP
    graph replay, and the plain per-graph form) matches the oracle and the
    graph run alone.  "fused": the packed and half tiers in one persistent
    launch."""
    tuning = {False: {}, True: HUB_TUNING, "fused": {"fused_small": 1}}[tuned]
    with G.tuning(**tuning):
        gs = mix_graphs(24, 5)
        gs.append(G.Graph.rmat(12, 16, seed=2))     # max degree 1339: hub tier when tuned
        if tuned is True:
            _, t = G.run(gs[-1], 1, method=method, **kw)
            assert method == G.FREQUENT or t["tier_vertices"]["hub"] > 0
        want = [oracle_labels(g, 6, method, kw) for g in gs]
        b = G.Batch(gs, method=method, **kw)
        for mode in (0, 1, 2):
            b.reset()
            b.run(6, mode=mode)
            for i, g in enumerate(gs):
                assert np.array_equal(b.labels(i), want[i]), "mode %d graph %d" % (mode, i)
        for i, g in enumerate(gs[:4] + gs[-1:]):
            alone, _ = G.run(g, 6, method=method, **kw)
            assert np.array_equal(alone, want[gs.index(g)])
        b.free()


@pytest.mark.parametrize("method,kw", [(G.CLASSIC, {}), (G.LAYERED, {"gamma": (1, 4)}),
                                       (G.BUFFER, {"table_cap": 3})])
def test_batch_with_tiny_and_edgeless_graphs(method, kw):
    """Single-vertex and edge-less graphs next to ordinary ones: every mode
    matches the oracle (an edge-less vertex keeps its label)."""
    gs = [G.Graph.from_edges(1, []), G.Graph.rmat(9, 8, seed=4), G.Graph.from_edges(3, []),
          G.Graph.from_edges(2, [(0, 1)]), G.Graph.from_edges(4, [(0, 1), (1, 2), (2, 3), (3, 0)])]
    want = [oracle_labels(g, 4, method, kw) for g in gs]
    b = G.Batch(gs, method=method, **kw)
    for mode in (0, 1, 2):
        b.reset()
        b.run(4, mode=mode)
        for i in range(len(gs)):
            assert np.array_equal(b.labels(i), want[i]), "mode %d graph %d" % (mode, i)
    b.free()


def test_batch_rejects_stream_methods():
    gs = mix_graphs(2, 1)
    with pytest.raises(G.GLPError):
        G.Batch(gs, method=G.SLPA)


# ------------------------------------------------------------------- served ----

@pytest.mark.parametrize("tuned", [False, True])
def test_server_results_equal_alone_in_the_mix(tuned):
    """Weighted classic and layered (gamma 1/4 and 0) jobs served together
    under a small budget: each result matches the oracle and the job run
    alone.  Tuned: the folded hub launches run (every graph has degrees >
    32)."""
    rng = np.random.default_rng(3)
    graphs, params, steps = [], [], []
    with G.tuning(**(HUB_TUNING if tuned else {})):
        for i in range(12):
            scale = int(rng.integers(8, 13))
            g = G.Graph.rmat(scale, int(rng.integers(8, 17)), seed=100 + i, wmax=8)
            layered = i % 2 == 1   # gamma 1/4 (slot-score table), or 0 (layered-hook table)
            p = G.make_params(method=G.LAYERED if layered else G.CLASSIC,
                              gamma=(1, 4) if i % 4 == 1 else (0, 1), weighted=True)
            graphs.append(g)
            params.append(p)
            steps.append(int(rng.integers(3, 9)))
        if tuned:
            assert all(g.info()["max_degree"] > HUB_TUNING["block_max_deg"] for g in graphs)
        sv = G.Server(budget_bytes=64 << 20, max_slots=4)   # small budget: jobs queue for memory
        ids = [sv.submit(g, s, params=p) for g, s, p in zip(graphs, steps, params)]
        for jid, g, s, p in zip(ids, graphs, steps, params):
            got, st = sv.wait(jid, g.n)
            assert st.state == 2 and st.steps == s
            kw = {"gamma": (p.gamma_num, p.gamma_den)}
            want = oracle_labels(g, s, G.LAYERED if p.method == G.LAYERED else G.CLASSIC, kw, weighted=True)
            assert np.array_equal(got, want), "job %d" % jid
            alone, _ = G.run(g, s, params=p)
            assert np.array_equal(got, alone)
        sv.close()


@pytest.mark.parametrize("scale,ef,seed,wmax", [(8, 8, 3, 0), (11, 16, 5, 8), (14, 32, 7, 8)])
def test_server_device_rmat_equals_host_generator(scale, ef, seed, wmax):
    """add_rmat builds Graph.rmat(..., degree_sort=False)'s CSR in HBM: the
    same offsets, neighbours and weights."""
    sv = G.Server(budget_bytes=1 << 30, max_slots=2)
    gid = sv.add_rmat(scale, ef, seed=seed, wmax=wmax)
    off, nbr, w = sv.graph_csr(gid, weights=wmax > 0)
    host = G.Graph.rmat(scale, ef, seed=seed, wmax=wmax, degree_sort=False)
    hoff, hnbr, hw = host.csr()
    assert np.array_equal(off, hoff) and np.array_equal(nbr, hnbr)
    if wmax:
        assert np.array_equal(w, hw)
    inf = sv.graph_info(gid)
    assert inf["n"] == host.n and inf["m"] == host.m and inf["max_degree"] == host.info()["max_degree"]
    sv.close()


@pytest.mark.parametrize("tuned", [False, True])
def test_server_resident_graphs_match_oracle(tuned):
    """Jobs on device-generated resident graphs (classic and layered 1/4,
    weighted), more jobs than slots and memory: every result equals the
    oracle on the graph's CSR and the same job on a second, idle server."""
    rng = np.random.default_rng(8)
    with G.tuning(**(HUB_TUNING if tuned else {})):
        sv = G.Server(budget_bytes=96 << 20, max_slots=3)
        gids = [sv.add_rmat(int(rng.integers(8, 13)), int(rng.integers(8, 33)), seed=200 + i, wmax=8)
                for i in range(4)]
        csr = {g: sv.graph_csr(g, weights=True) for g in gids}
        jobs = []
        for i in range(14):
            g = gids[int(rng.integers(0, len(gids)))]
            layered = bool(rng.integers(0, 2))
            p = G.make_params(method=G.LAYERED if layered else G.CLASSIC, gamma=(1, 4) if layered else (0, 1),
                              weighted=True)
            steps = int(rng.integers(1, 11))
            jobs.append((sv.submit_resident(g, steps, params=p), g, p, steps))
        solo = G.Server(budget_bytes=1 << 30, max_slots=1)
        for jid, g, p, steps in jobs:
            n = csr[g][0].shape[0] - 1
            got, st = sv.wait(jid, n)
            assert st.state == 2 and st.steps == steps
            assert st.submit_ms <= st.admit_ms <= st.start_ms <= st.done_ms
            off, nbr, w = csr[g]
            kw = {"gamma": (p.gamma_num, p.gamma_den)}
            want = oracle_labels_csr(off, nbr, w, steps, G.LAYERED if p.method == G.LAYERED else G.CLASSIC, kw)
            assert np.array_equal(got, want), "job %d" % jid
            alone, _ = solo.wait(solo.submit(G.Graph.from_csr(off, nbr, w), steps, params=p), n)
            assert np.array_equal(got, alone)
            sv.release(jid)
            with pytest.raises(G.GLPError):
                sv.poll(jid)
        solo.close()
        sv.close()


def test_server_graph_removal_and_budget():
    """A removed graph takes no new jobs but finishes its queued ones and
    then returns its bytes; a job that cannot fit next to the resident
    graphs fails instead of waiting, and fits once the graph is removed."""
    host = G.Graph.rmat(12, 16, seed=1, degree_sort=False)
    off, nbr, _ = host.csr()
    probe = G.Server(budget_bytes=1 << 30, max_slots=1)
    gb = probe.graph_info(probe.add_rmat(12, 16, seed=1))["bytes"]
    probe.close()
    # room for the resident graph and half as much again: a private copy of
    # the same graph (its bytes plus the job's state) never fits beside it
    sv = G.Server(budget_bytes=int(gb * 1.5), max_slots=2)
    g1 = sv.add_rmat(12, 16, seed=1)
    jid = sv.submit_resident(g1, 5)
    jid2 = sv.submit(host, 3)
    got, st = sv.wait(jid, host.n)
    assert st.state == 2 and np.array_equal(got, O.classic(off, nbr, 5)[-1])
    with pytest.raises(G.GLPError) as e:
        sv.wait(jid2, host.n)
    assert e.value.code == -8
    sv.remove_graph(g1)
    with pytest.raises(G.GLPError):
        sv.submit_resident(g1, 1)
    with pytest.raises(G.GLPError):
        sv.graph_info(g1)
    got, st = sv.wait(sv.submit(host, 3), host.n)
    assert st.state == 2 and np.array_equal(got, O.classic(off, nbr, 3)[-1])
    sv.close()


def test_server_budget_and_cancel():
    sv = G.Server(budget_bytes=1 << 20, max_slots=2)
    big = G.Graph.rmat(14, 16, seed=1)
    with pytest.raises(G.GLPError) as e:
        sv.submit(big, 3)
    assert e.value.code == -8
    small = G.Graph.rmat(8, 8, seed=2)
    a = sv.submit(small, 200)
    b = sv.submit(small, 2)
    sv.cancel(a)
    got, st = sv.wait(b, small.n)
    alone, _ = G.run(small, 2)
    assert np.array_equal(got, alone)
    assert sv.poll(a).state == 3
    sv.close()


def test_server_zero_steps_and_tiny_graphs():
    """Jobs of 0 steps end with L_0; single-vertex, two-vertex and small RMAT
    graphs served together match the oracle."""
    sv = G.Server(budget_bytes=64 << 20, max_slots=4)
    gs = [G.Graph.from_edges(1, []), G.Graph.from_edges(2, [(0, 1)]), G.Graph.rmat(8, 8, seed=5)]
    jobs = [(sv.submit(g, s), g, s) for g in gs for s in (0, 3)]
    for jid, g, s in jobs:
        got, st = sv.wait(jid, g.n)
        assert st.state == 2 and st.steps == s
        off, nbr, _ = g.csr()
        want = O.classic(off, nbr, s)[-1] if s else np.arange(g.n, dtype=np.int32)
        assert np.array_equal(got, want), "graph n=%d steps %d" % (g.n, s)
    sv.close()


def test_calib_random_read_runs():
    """The random 4-byte read calibration (the label gather's access pattern)
    runs and reports a time consistent with a sane read rate."""
    ms = G.calib_random_read(1 << 20, 1 << 22, 2)
    assert ms > 0
    assert (1 << 22) / (ms / 1e3) > 1e8   # more than 0.1 G reads/s
