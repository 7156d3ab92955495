"""ctypes binding of oracle/liboracle.so (the CPU restatement used as the
checker).  Test infrastructure only."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
REF_DIR = os.path.join(ORACLE_DIR, "_ref")
_path = os.path.join(ORACLE_DIR, "liboracle.so")
if not os.path.exists(_path):
    raise ImportError("oracle/liboracle.so not built: run `make -C oracle`")
_o = ctypes.CDLL(_path)

_i32p = ctypes.POINTER(ctypes.c_int32)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_i64p = ctypes.POINTER(ctypes.c_int64)


def _s(name, res, *args):
    f = getattr(_o, name)
    f.restype = res
    f.argtypes = list(args)


_s("orc_load_bin_size", ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _i32p, _i64p)
_s("orc_load_bin", ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _i32p, _i32p, _u32p)
_s("orc_classic_step", None, ctypes.c_int32, _i32p, _i32p, _u32p, _i32p, _i32p)
_s("orc_seeded_step", None, ctypes.c_int32, _i32p, _i32p, _u32p, ctypes.c_void_p, _i32p, _i32p)
_s("orc_layered_step", None, ctypes.c_int32, _i32p, _i32p, _u32p, ctypes.c_int32, ctypes.c_int32,
   _i32p, _i32p, _i32p, _i32p)
_s("orc_stream_u32", ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32)
_s("orc_slpa", None, ctypes.c_int32, _i32p, _i32p, _u32p, ctypes.c_uint64, ctypes.c_int32, _i32p)
_s("orc_slpa_post", None, ctypes.c_int32, ctypes.c_int32, _i32p, ctypes.c_int32, ctypes.c_int32,
   ctypes.c_int32, _i32p, _i32p, _i32p)
_s("orc_cm_hash", ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint32)
_s("orc_countmin_step", None, ctypes.c_int32, _i32p, _i32p, _u32p, ctypes.c_int32, ctypes.c_int32,
   ctypes.c_int32, _i32p, _i32p)
_s("orc_table_step", None, ctypes.c_int32, _i32p, _i32p, _u32p, ctypes.c_int32, ctypes.c_int32,
   _i32p, _i32p)
_s("orc_cpulp_run", ctypes.c_double, ctypes.c_int32, _i32p, _i32p, ctypes.c_int32, ctypes.c_int32,
   ctypes.c_int32, _i32p)


def _i(a):
    return a.ctypes.data_as(_i32p)


def _u(a):
    return None if a is None else a.ctypes.data_as(_u32p)


def _csr(off, nbr, w):
    off = np.ascontiguousarray(off, dtype=np.int32)
    nbr = np.ascontiguousarray(nbr, dtype=np.int32)
    if len(nbr) == 0:
        nbr = np.zeros(1, dtype=np.int32)
    w = None if w is None else np.ascontiguousarray(w, dtype=np.uint32)
    return off, nbr, w


def load_bin(path, directed=False, weighted=False):
    n, m = ctypes.c_int32(), ctypes.c_int64()
    if _o.orc_load_bin_size(os.fsencode(path), int(directed), int(weighted), ctypes.byref(n),
                            ctypes.byref(m)) != 0:
        raise IOError("oracle could not load %s" % path)
    off = np.zeros(n.value + 1, dtype=np.int32)
    nbr = np.zeros(max(m.value, 1), dtype=np.int32)
    w = np.zeros(max(m.value, 1), dtype=np.uint32) if weighted else None
    _o.orc_load_bin(os.fsencode(path), int(directed), int(weighted), _i(off), _i(nbr), _u(w))
    return off, nbr[:m.value], (w[:m.value] if weighted else None)


def classic_step(off, nbr, prev, w=None):
    off, nbr, w = _csr(off, nbr, w)
    n = len(off) - 1
    prev = np.ascontiguousarray(prev, dtype=np.int32)
    nxt = np.zeros(max(n, 1), dtype=np.int32)
    _o.orc_classic_step(n, _i(off), _i(nbr), _u(w), _i(prev), _i(nxt))
    return nxt[:n]


def classic(off, nbr, steps, w=None, labels=None):
    """Per-step labels, shape (steps, n), from L0 = identity (or `labels`)."""
    n = len(off) - 1
    L = np.arange(n, dtype=np.int32) if labels is None else np.asarray(labels, dtype=np.int32)
    out = []
    for _ in range(steps):
        L = classic_step(off, nbr, L, w)
        out.append(L)
    return np.array(out).reshape(steps, n)


def seeded_labels0(seeds):
    """Seeded LP's L_0: v on seeds, -1 elsewhere."""
    seeds = np.asarray(seeds, dtype=np.uint8)
    return np.where(seeds != 0, np.arange(len(seeds), dtype=np.int32), -1).astype(np.int32)


def seeded(off, nbr, steps, seeds, w=None):
    """Per-step labels of seeded LP, shape (steps, n)."""
    off, nbr, w = _csr(off, nbr, w)
    n = len(off) - 1
    seeds = np.ascontiguousarray(seeds, dtype=np.uint8)
    L = seeded_labels0(seeds)
    out = []
    for _ in range(steps):
        nxt = np.zeros(max(n, 1), dtype=np.int32)
        _o.orc_seeded_step(n, _i(off), _i(nbr), _u(w), seeds.ctypes.data, _i(L), _i(nxt))
        L = nxt[:n]
        out.append(L)
    return np.array(out).reshape(steps, n)


def layered(off, nbr, steps, p, q, w=None):
    off, nbr, w = _csr(off, nbr, w)
    n = len(off) - 1
    L = np.arange(n, dtype=np.int32)
    V = np.ones(n, dtype=np.int32)
    out, vols = [], []
    for _ in range(steps):
        nl = np.zeros(max(n, 1), dtype=np.int32)
        nv = np.zeros(max(n, 1), dtype=np.int32)
        _o.orc_layered_step(n, _i(off), _i(nbr), _u(w), p, q, _i(L), _i(V), _i(nl), _i(nv))
        L, V = nl[:n], nv[:n]
        out.append(L)
        vols.append(V)
    return np.array(out).reshape(steps, n), np.array(vols).reshape(steps, n)


def stream_u32(seed, t, u):
    return _o.orc_stream_u32(seed, t, u)


def slpa(off, nbr, seed, T, w=None):
    off, nbr, w = _csr(off, nbr, w)
    n = len(off) - 1
    mem = np.zeros(max((T + 1) * n, 1), dtype=np.int32)
    _o.orc_slpa(n, _i(off), _i(nbr), _u(w), seed, T, _i(mem))
    return mem[:(T + 1) * n].reshape(T + 1, n)


def slpa_post(mem, r_num, r_den, kmax):
    rows, n = mem.shape
    mem = np.ascontiguousarray(mem, dtype=np.int32)
    dom = np.zeros(max(n, 1), dtype=np.int32)
    ns = np.zeros(max(n, 1), dtype=np.int32)
    sets = np.zeros(max(n * kmax, 1), dtype=np.int32)
    _o.orc_slpa_post(n, rows, _i(mem), r_num, r_den, kmax, _i(dom), _i(ns), _i(sets))
    return dom[:n], ns[:n], sets[:n * kmax].reshape(n, kmax)


def cm_hash(row, key):
    return _o.orc_cm_hash(row, key)


def countmin_step(off, nbr, prev, rows, width, aliased, w=None):
    off, nbr, w = _csr(off, nbr, w)
    n = len(off) - 1
    prev = np.ascontiguousarray(prev, dtype=np.int32)
    nxt = np.zeros(max(n, 1), dtype=np.int32)
    _o.orc_countmin_step(n, _i(off), _i(nbr), _u(w), rows, width, int(aliased), _i(prev), _i(nxt))
    return nxt[:n]


def table_step(off, nbr, prev, cap, buffer, w=None):
    off, nbr, w = _csr(off, nbr, w)
    n = len(off) - 1
    prev = np.ascontiguousarray(prev, dtype=np.int32)
    nxt = np.zeros(max(n, 1), dtype=np.int32)
    _o.orc_table_step(n, _i(off), _i(nbr), _u(w), cap, int(buffer), _i(prev), _i(nxt))
    return nxt[:n]


def cpulp_run(off, nbr, niter, threads, mode):
    off, nbr, _ = _csr(off, nbr, None)
    n = len(off) - 1
    out = np.zeros(max(n, 1), dtype=np.int32)
    secs = _o.orc_cpulp_run(n, _i(off), _i(nbr), niter, threads, mode, _i(out))
    return out[:n], secs


def csr_from_edges(n, edges, weights=None):
    """Undirected CSR with the .bin reader's neighbour order (pair order)."""
    adj = [[] for _ in range(n)]
    adjw = [[] for _ in range(n)]
    for k, (u, v) in enumerate(edges):
        if u == v:
            continue
        wt = 1 if weights is None else weights[k]
        adj[u].append(v)
        adjw[u].append(wt)
        adj[v].append(u)
        adjw[v].append(wt)
    off = np.zeros(n + 1, dtype=np.int32)
    for v in range(n):
        off[v + 1] = off[v] + len(adj[v])
    nbr = np.array([x for a in adj for x in a], dtype=np.int32)
    w = None if weights is None else np.array([x for a in adjw for x in a], dtype=np.uint32)
    return off, nbr, w


_s("orc_cpulp_sample", ctypes.c_double, ctypes.c_int32, _i32p, _i32p, _i32p, _i32p, ctypes.c_int64,
   ctypes.c_int32, ctypes.c_int32)


def cpulp_sample(off, nbr, labels, verts, threads, balanced=False):
    """Verbatim CPUlp step over `verts` only (the reference's static OpenMP
    schedule, or dynamic chunks when balanced); labels updated in place;
    seconds."""
    off, nbr, _ = _csr(off, nbr, None)
    verts = np.ascontiguousarray(verts, dtype=np.int32)
    assert labels.dtype == np.int32 and labels.flags.c_contiguous
    return _o.orc_cpulp_sample(len(off) - 1, _i(off), _i(nbr), _i(labels), _i(verts), len(verts),
                               threads, int(balanced))


_s("orc_classic_subset", None, ctypes.c_int32, _i32p, _i32p, _u32p, _i32p, _i32p, ctypes.c_int64,
   ctypes.c_int32, _i32p)
_s("orc_cpulp_layered_sample", ctypes.c_double, ctypes.c_int32, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p,
   ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32)


def classic_subset(off, nbr, prev, verts, w=None, threads=1):
    """One classic step for the vertices `verts` only (labels in that order)."""
    off, nbr, w = _csr(off, nbr, w)
    prev = np.ascontiguousarray(prev, dtype=np.int32)
    verts = np.ascontiguousarray(verts, dtype=np.int32)
    out = np.zeros(max(len(verts), 1), dtype=np.int32)
    _o.orc_classic_subset(len(off) - 1, _i(off), _i(nbr), _u(w), _i(prev), _i(verts), len(verts), threads,
                          _i(out))
    return out[:len(verts)]


def cpulp_layered_sample(off, nbr, labels, Vr, Vw, verts, threads, p, q):
    """One verbatim layered CPUlp step over `verts` (labels, Vw updated in
    place); seconds."""
    off, nbr, _ = _csr(off, nbr, None)
    verts = np.ascontiguousarray(verts, dtype=np.int32)
    for a in (labels, Vr, Vw):
        assert a.dtype == np.int32 and a.flags.c_contiguous
    return _o.orc_cpulp_layered_sample(len(off) - 1, _i(off), _i(nbr), _i(labels), _i(Vr), _i(Vw), _i(verts),
                                       len(verts), threads, p, q)
