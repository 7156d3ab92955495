"""Pins the CPU oracle (oracle/liboracle.so) with checks that do not restate
its code: hand-derived closed forms, reductions between methods, published
known-answer vectors, and the reference's own loader compiled unchanged."""
import numpy as np
import pytest

import oracle_lib as O


def ident(n):
    return np.arange(n, dtype=np.int32)


# ---------------------------------------------------------------- classic ----

def test_clique_two_steps():
    k = 6
    edges = [(i, j) for i in range(k) for j in range(i + 1, k)]
    off, nbr, _ = O.csr_from_edges(k, edges)
    L = O.classic(off, nbr, 2)
    # step 1: every vertex sees each other label once -> highest other id
    assert L[0].tolist() == [k - 1] * (k - 1) + [k - 2]
    # step 2: k-2 neighbours show k-1, one shows k-2 -> k-1 everywhere
    assert L[1].tolist() == [k - 1] * k


def test_star_oscillates():
    k = 5
    off, nbr, _ = O.csr_from_edges(k + 1, [(0, i) for i in range(1, k + 1)])
    L = O.classic(off, nbr, 3)
    assert L[0].tolist() == [k] + [0] * k          # centre takes the max leaf id
    assert L[1].tolist() == [0] + [k] * k
    assert L[2].tolist() == L[0].tolist()


def test_path_is_synchronous():
    off, nbr, _ = O.csr_from_edges(6, [(0, 1), (1, 2), (2, 3), (3, 4)])  # vertex 5 isolated
    L = O.classic(off, nbr, 2)
    assert L[0].tolist() == [1, 2, 3, 4, 3, 5]
    assert L[1].tolist() == [2, 3, 4, 3, 4, 5]


def test_two_triangles_with_bridge():
    edges = [(0, 1), (1, 2), (0, 2), (3, 4), (4, 5), (3, 5), (2, 3)]
    off, nbr, _ = O.csr_from_edges(6, edges)
    L = O.classic(off, nbr, 2)
    assert L[0].tolist() == [2, 2, 3, 5, 5, 4]
    assert L[1].tolist() == [3, 3, 2, 5, 5, 5]


def test_multiplicity_counts():
    # duplicate edge 0-1: label 1 is seen twice and beats the higher label 2
    off, nbr, _ = O.csr_from_edges(3, [(0, 1), (0, 1), (0, 2)])
    assert O.classic_step(off, nbr, ident(3))[0] == 1


def test_weighted_sums():
    off, nbr, w = O.csr_from_edges(4, [(0, 1), (0, 2), (0, 3)], weights=[5, 1, 1])
    prev = np.array([0, 7, 9, 9], dtype=np.int32)
    assert O.classic_step(off, nbr, prev, w)[0] == 7      # 5 > 1+1
    assert O.classic_step(off, nbr, prev)[0] == 9         # unweighted: 2 > 1


def _random_graph(rng, n, m):
    e = rng.integers(0, n, size=(m, 2))
    return O.csr_from_edges(n, [tuple(x) for x in e])


def test_argmax_characterization_random():
    """Every winner has the maximum frequency, ties to the highest label, and
    isolated vertices keep their label."""
    rng = np.random.default_rng(5)
    for trial in range(20):
        n = int(rng.integers(2, 40))
        off, nbr, _ = _random_graph(rng, n, int(rng.integers(0, 3 * n)))
        prev = rng.integers(0, 6, size=n).astype(np.int32)
        nxt = O.classic_step(off, nbr, prev)
        for v in range(n):
            labs = prev[nbr[off[v]:off[v + 1]]]
            if len(labs) == 0:
                assert nxt[v] == prev[v]
                continue
            cnt = np.bincount(labs, minlength=8)
            best = cnt.max()
            assert cnt[nxt[v]] == best
            assert nxt[v] == np.flatnonzero(cnt == best).max()


def test_monotone_relabel_equivariance():
    rng = np.random.default_rng(9)
    off, nbr, _ = _random_graph(rng, 60, 200)
    prev = rng.integers(0, 10, size=60).astype(np.int32)
    f = np.cumsum(rng.integers(1, 5, size=10)).astype(np.int32)  # strictly increasing map
    assert np.array_equal(O.classic_step(off, nbr, f[prev]), f[O.classic_step(off, nbr, prev)])


def test_cpulp_sync_mode_matches_classic():
    rng = np.random.default_rng(3)
    off, nbr, _ = _random_graph(rng, 300, 1200)
    out, secs = O.cpulp_run(off, nbr, 5, 4, 1)
    assert np.array_equal(out, O.classic(off, nbr, 5)[-1])
    out0, _ = O.cpulp_run(off, nbr, 2, 1, 0)  # verbatim mode runs (timing only)
    assert out0.shape == (300,)


def test_classic_subset_equals_full_step():
    rng = np.random.default_rng(21)
    for trial in range(6):
        n = int(rng.integers(5, 300))
        e = rng.integers(0, n, size=(int(rng.integers(0, 6 * n)), 2))
        wts = rng.integers(1, 9, size=len(e)).tolist() if trial % 2 else None
        off, nbr, w = O.csr_from_edges(n, [tuple(x) for x in e], weights=wts)
        prev = rng.integers(0, 7, size=n).astype(np.int32)
        full = O.classic_step(off, nbr, prev, w)
        verts = rng.permutation(n)[: max(1, n // 2)].astype(np.int32)
        assert np.array_equal(O.classic_subset(off, nbr, prev, verts, w=w, threads=3), full[verts])


def test_cpulp_layered_sample_picks_neighbour_labels():
    rng = np.random.default_rng(22)
    off, nbr, _ = _random_graph(rng, 200, 800)
    lab = np.arange(200, dtype=np.int32)
    vr = np.ones(200, dtype=np.int32)
    vw = np.zeros(200, dtype=np.int32)
    before = lab.copy()
    O.cpulp_layered_sample(off, nbr, lab, vr, vw, np.arange(200, dtype=np.int32), 1, 1, 4)
    # one thread sweeps ids in order with live reads: a neighbour u < v
    # already shows its new label, u > v still its old one
    for v in range(200):
        nb = nbr[off[v]:off[v + 1]]
        if len(nb) == 0:
            assert lab[v] == before[v]
            continue
        shown = [lab[u] if u < v else before[u] for u in nb]
        assert lab[v] in shown
    assert vw.sum() == 200


# ---------------------------------------------------------------- layered ----

def test_layered_hand_scores():
    # vertex 0 sees labels 5,5,6; V(5)=6, V(6)=2 from the isolated vertices
    off, nbr, _ = O.csr_from_edges(10, [(0, 1), (0, 2), (0, 3)])
    prev = np.array([0, 5, 5, 6, 5, 5, 5, 5, 6, 9], dtype=np.int32)
    vprev = np.bincount(prev, minlength=10).astype(np.int32)
    got = {}
    for (p, q), want in {(0, 1): 5, (1, 4): 5, (1, 1): 6}.items():
        nl = np.zeros(10, dtype=np.int32)
        nv = np.zeros(10, dtype=np.int32)
        O._o.orc_layered_step(10, O._i(off), O._i(nbr), None, p, q, O._i(prev), O._i(vprev),
                              O._i(nl), O._i(nv))
        got[(p, q)] = int(nl[0])
        assert nl[0] == want  # (q+p)C - pV: 2C-V gives 5:-2, 6:0
        assert np.array_equal(nv, np.bincount(nl, minlength=10))
    assert got == {(0, 1): 5, (1, 4): 5, (1, 1): 6}


def test_layered_gamma_zero_is_classic():
    rng = np.random.default_rng(11)
    off, nbr, _ = _random_graph(rng, 200, 700)
    L, V = O.layered(off, nbr, 4, 0, 1)
    assert np.array_equal(L, O.classic(off, nbr, 4))
    assert np.array_equal(V[-1], np.bincount(L[-1], minlength=200))


# ------------------------------------------------------------------- SLPA ----

def test_stream_known_answers():
    # SplitMix64 published outputs: state 0 -> 0xe220a8397b1dcdaf, state
    # 1234567 -> 6457827717110365317 (Vigna's reference splitmix64.c)
    assert O.stream_u32(0, 0, 0) == 0x7B1DCDAF
    assert O.stream_u32(0, 0, 1234567) == 6457827717110365317 & 0xFFFFFFFF


def test_slpa_first_step_is_classic():
    rng = np.random.default_rng(2)
    off, nbr, _ = _random_graph(rng, 150, 500)
    mem = O.slpa(off, nbr, seed=7, T=1)
    assert np.array_equal(mem[0], ident(150))
    assert np.array_equal(mem[1], O.classic(off, nbr, 1)[0])  # R mod 1 = 0 -> speakers show M[0]


def test_slpa_memory_rows_are_neighbour_labels():
    rng = np.random.default_rng(4)
    off, nbr, _ = _random_graph(rng, 80, 300)
    mem = O.slpa(off, nbr, seed=3, T=6)
    for t in range(1, 7):
        for v in range(80):
            nb = nbr[off[v]:off[v + 1]]
            if len(nb) == 0:
                assert mem[t, v] == mem[t - 1, v]
            else:  # the listener picks a label some neighbour held earlier
                assert mem[t, v] in set(mem[:t, nb].ravel().tolist())


def _splitmix_first(state):
    """First output of Vigna's SplitMix64 from `state` (independent of the
    oracle; pinned by the published answers below)."""
    m = (1 << 64) - 1
    z = (state + 0x9e3779b97f4a7c15) & m
    z = ((z ^ (z >> 30)) * 0xbf58476d1ce4e5b9) & m
    z = ((z ^ (z >> 27)) * 0x94d049bb133111eb) & m
    return z ^ (z >> 31)


def test_slpa_hand_trace_three_steps():
    """Graph 0-1, 1-2, 2-3, 3-4, 1-3; seed 3.  Speaker u at step t shows
    M_u[R mod t], R = low 32 bits of SplitMix64's first output from state
    (seed<<40) ^ (t<<32) ^ u; the listener takes the most frequent shown
    label, ties to the highest.  The memory rows below are traced by hand."""
    assert _splitmix_first(0) == 0xe220a8397b1dcdaf
    assert _splitmix_first(1234567) == 6457827717110365317
    seed = 3
    rows = {t: [(_splitmix_first((seed << 40) ^ (t << 32) ^ u) & 0xFFFFFFFF) % t for u in range(5)]
            for t in (2, 3)}
    assert rows == {2: [1, 0, 1, 0, 1], 3: [2, 1, 2, 1, 0]}
    for t in (1, 2, 3):
        assert [O.stream_u32(seed, t, u) % t for u in range(5)] == ([0] * 5 if t == 1 else rows[t])
    off, nbr, _ = O.csr_from_edges(5, [(0, 1), (1, 2), (2, 3), (3, 4), (1, 3)])
    mem = O.slpa(off, nbr, seed=seed, T=3)
    # t=1: every speaker shows M[0] = id; v1 sees {0,2,3} -> 3, v3 {2,4,1} -> 4
    # t=2: shown = [M1[0], M0[1], M1[2], M0[3], M1[4]] = [1,1,3,3,3]
    # t=3: shown = [M2[0], M1[1], M2[2], M1[3], M0[4]] = [1,3,3,4,4]
    assert mem.tolist() == [[0, 1, 2, 3, 4],
                            [1, 3, 3, 4, 3],
                            [1, 3, 3, 3, 3],
                            [3, 4, 4, 3, 4]]


def test_slpa_post_hand():
    mem = np.array([[3], [3], [5], [5], [5], [7], [3]], dtype=np.int32)
    dom, ns, sets = O.slpa_post(mem, 1, 10, 4)
    assert dom.tolist() == [5] and ns.tolist() == [3] and sets[0].tolist() == [3, 5, 7, -1]
    dom, ns, sets = O.slpa_post(mem, 1, 3, 4)   # keep count*3 >= 7 -> count >= 3
    assert ns.tolist() == [2] and sets[0].tolist() == [3, 5, -1, -1]


# -------------------------------------------------------------- count-min ----

def _cm_hash_py(row, k):
    """Independent restatement of method/countmin.cuh:97-142."""
    M = 0xFFFFFFFF
    spec = {0: (16, 0x12ebcb3b, 13, 0xc123ae53, 16), 1: (16, 0x21eb21bb, 13, 0x2cb2ea35, 16),
            2: (16, 0x8febce61, 13, 0xa2b2be25, 16), 3: (12, 0x13ebdb3a, 12, 0xc123ae53, 14),
            4: (13, 0x312b521b, 5, 0x1cb3ea22, 17)}[row]
    k ^= k >> spec[0]
    k = (k * spec[1]) & M
    k ^= k >> spec[2]
    k = (k * spec[3]) & M
    k ^= k >> spec[4]
    return k


def test_cm_hash_known_answers():
    for row in range(5):
        for k in [0, 1, 2, 3, 255, 65536, 2 ** 31 - 1, 123456789, 0xFFFFFFFF]:
            assert O.cm_hash(row, k) == _cm_hash_py(row, k)
    assert O.cm_hash(0, 0) == 0


def test_countmin_without_collisions_is_classic():
    W = 4096
    prev = ident(120)

    def row0_collision_free(off, nbr):
        for v in range(120):
            keys = set((prev[nbr[off[v]:off[v + 1]]] + 1).tolist())
            if len({_cm_hash_py(0, k) % W for k in keys}) != len(keys):
                return False
        return True

    # first seeded instance whose row 0 has no collision among any vertex's
    # neighbour labels: its row-0 cells are exact, so every estimate is exact
    for seed in range(100):
        off, nbr, _ = _random_graph(np.random.default_rng(seed), 120, 400)
        if row0_collision_free(off, nbr):
            break
    else:
        pytest.fail("no collision-free instance found")
    for aliased in (0, 1):
        assert np.array_equal(O.countmin_step(off, nbr, prev, 3, W, aliased),
                              O.classic_step(off, nbr, prev))


# Hand-derived collision cases at width 4.  Vertex 0 is a star centre whose
# leaves show label a (count ca) and label b > a (count cb); every other
# vertex is a leaf.  The cells each key touches come from _cm_hash_py, and
# the expected winner is argued from them by hand below (C_r = row r's cell
# sums; est = minimum over the cells a key touches; ties go to the higher
# label).  rows are method/countmin.cuh:145-263: proper rows use h_r in row
# r; aliased d=3 adds h_3's increment to row 2 (:185-191), aliased d=4 adds
# h_3's cell of row 2 twice (:210-219).
CM_W = 4


def _cm_cells(label):
    return [_cm_hash_py(r, label + 1) % CM_W for r in range(5)]


def _cm_centre(counts, rows, aliased):
    """The oracle's label for the centre of a star whose leaves show
    `counts` = {label: multiplicity}."""
    prev = [0]
    for lab, c in counts.items():
        prev += [lab] * c
    n = len(prev)
    off, nbr, _ = O.csr_from_edges(n, [(0, i) for i in range(1, n)])
    return int(O.countmin_step(off, nbr, np.array(prev, dtype=np.int32), rows, CM_W, aliased)[0])


def test_countmin_rows_beyond_the_first_decide():
    """a=1 (x2), b=4 (x1): both keys share their row-0 cell (C_0 = 3 for
    both), rows 1 and 2 separate them.  Proper min over rows: est(a) = 2,
    est(b) = 1, so a wins; row 0 alone would tie 3:3 and give b."""
    a, b = 1, 4
    A, B = _cm_cells(a), _cm_cells(b)
    assert A[0] == B[0] and A[1] != B[1] and A[2] != B[2] and A[4] != B[4]
    for rows in (2, 3, 5):
        assert _cm_centre({a: 2, b: 1}, rows, 0) == a
    assert _cm_centre({a: 2, b: 1}, 1, 0) == b   # row 0 only: the 3:3 tie
    # d=4 proper: rows 3 collide (A[3] == B[3]) but rows 1, 2 still separate
    assert A[3] == B[3] and _cm_centre({a: 2, b: 1}, 4, 0) == a


def test_countmin_aliased_row_two_crosses_keys():
    """a=2 (x2), b=17 (x1), cells a=[1,2,0,..], b=[1,0,2,..].  Proper d=3:
    est(a) = min(3,2,2) = 2 > est(b) = min(3,1,1) = 1, a wins.  Aliased d=3
    puts every key's h_1 and h_2 increments in row 1, where a's h_1 cell (2)
    is b's h_2 cell and a's h_2 cell (0) is b's h_1 cell: both cells hold 3,
    so est(a) = est(b) = 3 and the tie goes to b.  Aliased d=4 adds h_2
    twice: C_1[0] = 4+1 = 5, C_1[2] = 2+2 = 4, still est 3:3, b."""
    a, b = 2, 17
    A, B = _cm_cells(a), _cm_cells(b)
    assert A[0] == B[0] and B[1] == A[2] and B[2] == A[1]
    assert A[1] != B[1] and A[2] != B[2] and A[3] != B[3]
    assert _cm_centre({a: 2, b: 1}, 3, 0) == a
    assert _cm_centre({a: 2, b: 1}, 4, 0) == a
    assert _cm_centre({a: 2, b: 1}, 3, 1) == b
    assert _cm_centre({a: 2, b: 1}, 4, 1) == b


def test_countmin_aliased_d3_and_d4_differ():
    """a=1, b=15, one edge each, cells a=[3,0,0,..], b=[3,0,2,..].  Proper:
    rows 0 and 1 collide (2:2), row 2 separates (1:1): tie, b.  Aliased d=3:
    row 1 gets a's h_1 and h_2 (both cell 0) and b's h_1 (cell 0) and h_2
    (cell 2): C_1[0] = 3, C_1[2] = 1, so est(a) = min(2,3,3) = 2 and
    est(b) = min(2,3,1) = 1: a wins although b is higher.  Aliased d=4 adds
    h_2 twice: C_1[0] = 1+2+1 = 4, C_1[2] = 2, est(a) = 2 = est(b): b."""
    a, b = 1, 15
    A, B = _cm_cells(a), _cm_cells(b)
    assert A[0] == B[0] and A[1] == A[2] == B[1] and B[2] != A[1] and A[3] != B[3]
    for rows in (3, 4, 5):
        assert _cm_centre({a: 1, b: 1}, rows, 0) == b
    assert _cm_centre({a: 1, b: 1}, 3, 1) == a
    assert _cm_centre({a: 1, b: 1}, 4, 1) == b


def test_countmin_width_one_is_highest_label():
    off, nbr, _ = O.csr_from_edges(4, [(0, 1), (0, 1), (0, 2), (0, 3)])
    # every label hits the single cell -> all estimates equal -> highest label
    assert O.countmin_step(off, nbr, ident(4), 3, 1, 0)[0] == 3
    assert O.classic_step(off, nbr, ident(4))[0] == 1


# ----------------------------------------------------------- bounded tables ----

def test_tables_with_room_are_classic():
    rng = np.random.default_rng(12)
    off, nbr, _ = _random_graph(rng, 100, 300)
    cap = int(np.diff(off).max())
    for buf in (0, 1):
        assert np.array_equal(O.table_step(off, nbr, ident(100), cap, buf),
                              O.classic_step(off, nbr, ident(100)))


def test_tables_hand_trace():
    off, nbr, w = O.csr_from_edges(13, [(0, 1), (0, 2), (0, 3)], weights=[1, 1, 3])
    prev = np.array([0, 10, 11, 12] + [4] * 9, dtype=np.int32)
    # Misra-Gries, 2 slots: {10:1, 11:1}; 12 (w=3) drops both by 1 and enters with 2
    assert O.table_step(off, nbr, prev, 2, 0, w)[0] == 12
    # buffer, 2 slots: 12 (w=3) empties both slots and is dropped -> keep own label
    assert O.table_step(off, nbr, prev, 2, 1, w)[0] == 0
    # unweighted MG, 2 slots: third label cancels both; table empty -> keep
    assert O.table_step(off, nbr, prev, 2, 0)[0] == 0
    assert O.table_step(off, nbr, prev, 3, 1, w)[0] == 12


# ------------------------------------------------------------------ seeded ----

def test_seeded_path_one_hop_per_step():
    off, nbr, _ = O.csr_from_edges(5, [(0, 1), (1, 2), (2, 3), (3, 4)])
    got = O.seeded(off, nbr, 4, [1, 0, 0, 0, 0])
    assert got.tolist() == [[0, 0, -1, -1, -1], [0, 0, 0, -1, -1], [0, 0, 0, 0, -1], [0, 0, 0, 0, 0]]


def test_seeded_competing_seeds_tie_to_higher_label():
    """Seeds 0 and 6 at the ends of a 7-path meet in the middle: vertex 3 sees
    one neighbour with each label and takes the higher (6)."""
    off, nbr, _ = O.csr_from_edges(7, [(i, i + 1) for i in range(6)])
    got = O.seeded(off, nbr, 4, [1, 0, 0, 0, 0, 0, 1])
    assert got[0].tolist() == [0, 0, -1, -1, -1, 6, 6]
    assert got[1].tolist() == [0, 0, 0, -1, 6, 6, 6]
    assert got[2].tolist() == [0, 0, 0, 6, 6, 6, 6]
    # vertex 2 then sees 0 (from 1) and 6 (from 3): tie, 6; vertex 4 sees 6 twice
    assert got[3].tolist() == [0, 0, 6, 6, 6, 6, 6]


def test_seeded_star_and_isolated():
    """Leaves 1..5 seeded: the centre takes the highest leaf; an isolated
    non-seed stays unlabeled, an isolated seed keeps its id; seeds never move."""
    off, nbr, _ = O.csr_from_edges(8, [(0, i) for i in range(1, 6)])
    seeds = [0, 1, 1, 1, 1, 1, 0, 1]
    got = O.seeded(off, nbr, 3, seeds)
    for t in range(3):
        assert got[t].tolist() == [5, 1, 2, 3, 4, 5, -1, 7]


def test_seeded_weights_decide():
    off, nbr, w = O.csr_from_edges(3, [(0, 1), (1, 2)], weights=[3, 1])
    got = O.seeded(off, nbr, 1, [1, 0, 1], w=w)
    assert got[0].tolist() == [0, 0, 2]     # 3 > 1 outweighs the higher label
    off, nbr, w = O.csr_from_edges(3, [(0, 1), (1, 2)], weights=[1, 1])
    assert O.seeded(off, nbr, 1, [1, 0, 1], w=w)[0].tolist() == [0, 2, 2]


def test_seeded_extremes():
    rng = np.random.default_rng(4)
    e = [tuple(x) for x in rng.integers(0, 200, size=(800, 2))]
    off, nbr, _ = O.csr_from_edges(200, e)
    n = 200
    assert (O.seeded(off, nbr, 3, np.ones(n)) == np.arange(n)).all()
    assert (O.seeded(off, nbr, 3, np.zeros(n)) == -1).all()


def test_seeded_invariants_random():
    """Seeds stay fixed, every label is -1 or a seed id, and the labeled set
    after t steps is exactly the vertices within t hops of a seed (BFS)."""
    rng = np.random.default_rng(5)
    for trial in range(5):
        n = 400
        e = [tuple(x) for x in rng.integers(0, n, size=(700, 2))]
        off, nbr, _ = O.csr_from_edges(n, e)
        seeds = (rng.random(n) < 0.02).astype(np.uint8)
        got = O.seeded(off, nbr, 6, seeds)
        dist = np.full(n, -1)
        frontier = list(np.flatnonzero(seeds))
        dist[frontier] = 0
        d = 0
        while frontier:
            d += 1
            nxt = []
            for u in frontier:
                for v in nbr[off[u]:off[u + 1]]:
                    if dist[v] < 0:
                        dist[v] = d
                        nxt.append(v)
            frontier = nxt
        seed_ids = set(np.flatnonzero(seeds).tolist()) | {-1}
        for t in range(6):
            assert (got[t][seeds != 0] == np.flatnonzero(seeds)).all()
            assert set(got[t].tolist()) <= seed_ids
            labeled = got[t] >= 0
            assert np.array_equal(labeled, (dist >= 0) & (dist <= t + 1))
