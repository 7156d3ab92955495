"""Python binding of libglp.so, the MI355X label-propagation library.

This module is a thin ctypes mirror of the C ABI declared in include/glp.h
and include/glp_tools.h.  It never computes labels itself: every LP step
runs in the HIP kernels of libglp.so, and the first call into the library
fails loudly (ImportError) when it has not been built (python -m
glp_amd.build); importing the package alone does not need it.

Reference interface mirrored (Gibyeng/GLP):
  * CSRGraph(path, startvalue, direct)       -> Graph.load_bin   (header/graph.h:69)
  * XLprop(G).run(niter)                      -> run() / Session   (main.cu:63-101)
  * method ids 1 classic, 2 layered, 3 SLPA, 4 user hooks        (main.cu:65-98)
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libglp.so")

_SIGS = []  # (name, restype, argtypes) applied when the library loads


class _Library:
    """libglp.so, loaded on first use.  Importing the package never needs the
    library (so the build module runs from a clean tree); the first call into
    it raises ImportError when it has not been built."""

    handle = None

    def __getattr__(self, name):
        if _Library.handle is None:
            _load()
        return getattr(_Library.handle, name)


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError("libglp.so not built: run `python -m glp_amd.build` (looked in %s)" % _HERE)
    h = ctypes.CDLL(LIB_PATH)
    for name, restype, argtypes in _SIGS:
        f = getattr(h, name, None)
        if f is None:
            continue
        f.restype = restype
        f.argtypes = list(argtypes)
    _Library.handle = h


_lib = _Library()

CLASSIC, LAYERED, SLPA, USER, COUNTMIN, FREQUENT, BUFFER, SEEDED = 1, 2, 3, 4, 5, 6, 7, 8
METHOD_NAMES = {CLASSIC: "classic", LAYERED: "layered", SLPA: "slpa", USER: "user",
                COUNTMIN: "countmin", FREQUENT: "frequent", BUFFER: "buffer", SEEDED: "seeded"}
TIER_NAMES = ["packed", "wave", "block", "hub", "gather", "pick", "aux", "spare"]

_c_i32p = ctypes.POINTER(ctypes.c_int32)
_c_u32p = ctypes.POINTER(ctypes.c_uint32)
_c_i64p = ctypes.POINTER(ctypes.c_int64)


class GLPError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("glp error %d: %s" % (code, msg))
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int32), ("weighted", ctypes.c_int32),
                ("gamma_num", ctypes.c_int32), ("gamma_den", ctypes.c_int32),
                ("seed", ctypes.c_uint64), ("slpa_T", ctypes.c_int32),
                ("slpa_r_num", ctypes.c_int32), ("slpa_r_den", ctypes.c_int32),
                ("cm_rows", ctypes.c_int32), ("cm_width", ctypes.c_int32),
                ("cm_aliased", ctypes.c_int32), ("table_cap", ctypes.c_int32),
                ("gather", ctypes.c_int32), ("profile", ctypes.c_int32),
                ("frontier", ctypes.c_int32), ("reserved", ctypes.c_int32 * 5)]


class Timing(ctypes.Structure):
    _fields_ = [("upload_ms", ctypes.c_double), ("plan_ms", ctypes.c_double),
                ("steps_ms", ctypes.c_double), ("tier_ms", ctypes.c_double * 8),
                ("tier_launches", ctypes.c_int64 * 8), ("tier_vertices", ctypes.c_int64 * 8),
                ("tier_edges", ctypes.c_int64 * 8), ("steps", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]

    def as_dict(self):
        return {"upload_ms": self.upload_ms, "plan_ms": self.plan_ms, "steps_ms": self.steps_ms,
                "steps": self.steps,
                "tier_ms": {TIER_NAMES[i]: self.tier_ms[i] for i in range(8)},
                "tier_launches": {TIER_NAMES[i]: self.tier_launches[i] for i in range(8)},
                "tier_vertices": {TIER_NAMES[i]: self.tier_vertices[i] for i in range(8)},
                "tier_edges": {TIER_NAMES[i]: self.tier_edges[i] for i in range(8)}}


class JobStats(ctypes.Structure):
    _fields_ = [("submit_ms", ctypes.c_double), ("admit_ms", ctypes.c_double),
                ("done_ms", ctypes.c_double), ("bytes", ctypes.c_int64),
                ("steps", ctypes.c_int32), ("state", ctypes.c_int32),
                ("start_ms", ctypes.c_double)]


def _sig(name, restype, *argtypes):
    _SIGS.append((name, restype, argtypes))


_vp = ctypes.c_void_p
_pp = ctypes.POINTER(ctypes.c_void_p)
_sig("glp_last_error", ctypes.c_char_p)
_sig("glp_version", ctypes.c_char_p)
_sig("glp_params_default", None, ctypes.POINTER(Params))
_sig("glp_device_count", ctypes.c_int, _c_i32p)
_sig("glp_graph_load_bin", ctypes.c_int, ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32, _pp)
_sig("glp_graph_from_csr", ctypes.c_int, ctypes.c_int32, ctypes.c_int64, _c_i32p, _c_i32p, _c_u32p, _pp)
_sig("glp_graph_info", ctypes.c_int, _vp, _c_i32p, _c_i64p, _c_i32p, _c_i32p)
_sig("glp_graph_symmetric", ctypes.c_int, _vp, _c_i32p)
_sig("glp_graph_csr", ctypes.c_int, _vp, _c_i32p, _c_i32p, _c_u32p)
_sig("glp_graph_free", None, _vp)
_sig("glp_graph_save_bin", ctypes.c_int, _vp, ctypes.c_char_p)
_sig("glp_graph_degree_sort", ctypes.c_int, _vp, _pp, _c_i32p)
_sig("glp_tool_edgelist2bin", ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p)
_sig("glp_tool_normalize", ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p)
_sig("glp_tool_direct_edgelist2bin", ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int32)
_sig("glp_calib_random_read", ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32,
     ctypes.POINTER(ctypes.c_double))
_sig("glp_gen_rmat", ctypes.c_int, ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
     ctypes.c_double, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _pp)
_sig("glp_gen_planted", ctypes.c_int, ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
     ctypes.c_uint64, ctypes.c_int32, _c_i32p, _pp)
_sig("glp_gen_lfr", ctypes.c_int, ctypes.c_int32, ctypes.c_double, ctypes.c_int32, ctypes.c_int32,
     ctypes.c_double, ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_uint64, _c_i32p, _pp)
_sig("glp_gen_mix_params", ctypes.c_int, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
     ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, _c_i32p, _c_i32p,
     ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint64))
_sig("glp_stream_u32", ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32)
_sig("glp_stream_fill", ctypes.c_int, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _c_u32p)
_sig("glp_run", ctypes.c_int, _vp, ctypes.POINTER(Params), ctypes.c_int32, _c_i32p, _c_i32p,
     ctypes.POINTER(Timing))
_sig("glp_session_create", ctypes.c_int, _vp, ctypes.POINTER(Params), _pp)
_sig("glp_session_reset", ctypes.c_int, _vp)
_sig("glp_session_step", ctypes.c_int, _vp, ctypes.c_int32)
_sig("glp_session_step_graph", ctypes.c_int, _vp, ctypes.c_int32)
_sig("glp_session_sync", ctypes.c_int, _vp)
_sig("glp_session_labels", ctypes.c_int, _vp, _c_i32p)
_sig("glp_session_changes", ctypes.c_int, _vp, _c_i64p, ctypes.c_int32, _c_i32p)
_sig("glp_session_timing", ctypes.c_int, _vp, ctypes.POINTER(Timing))
_sig("glp_session_volume", ctypes.c_int, _vp, _c_i32p)
_sig("glp_session_save", ctypes.c_int, _vp, _vp, ctypes.POINTER(ctypes.c_size_t))
_sig("glp_session_load", ctypes.c_int, _vp, _vp, ctypes.c_size_t)
_sig("glp_session_slpa_memory", ctypes.c_int, _vp, _c_i32p)
_sig("glp_session_slpa_post", ctypes.c_int, _vp, ctypes.c_int32, _c_i32p, _c_i32p, _c_i32p)
_sig("glp_session_free", ctypes.c_int, _vp)
_sig("glp_batch_create", ctypes.c_int, ctypes.c_int32, _pp, ctypes.POINTER(Params), _pp)
_sig("glp_batch_run", ctypes.c_int, _vp, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(ctypes.c_double))
_sig("glp_batch_labels", ctypes.c_int, _vp, ctypes.c_int32, _c_i32p)
_sig("glp_batch_reset", ctypes.c_int, _vp)
_sig("glp_batch_free", None, _vp)
_sig("glp_batch_autotune", ctypes.c_int, _vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_double))
_sig("glp_server_create", ctypes.c_int, ctypes.c_uint64, ctypes.c_int32, _pp)
_sig("glp_server_submit", ctypes.c_int, _vp, _vp, ctypes.POINTER(Params), ctypes.c_int32, _c_i64p)
_sig("glp_server_poll", ctypes.c_int, _vp, ctypes.c_int64, ctypes.POINTER(JobStats))
_sig("glp_server_wait", ctypes.c_int, _vp, ctypes.c_int64, _c_i32p, ctypes.POINTER(JobStats))
_sig("glp_server_cancel", ctypes.c_int, _vp, ctypes.c_int64)
_sig("glp_server_release", ctypes.c_int, _vp, ctypes.c_int64)
_sig("glp_server_add_graph", ctypes.c_int, _vp, _vp, _c_i32p)
_sig("glp_server_add_rmat", ctypes.c_int, _vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
     ctypes.c_double, ctypes.c_uint64, ctypes.c_int32, _c_i32p)
_sig("glp_server_graph_info", ctypes.c_int, _vp, ctypes.c_int32, _c_i32p, _c_i64p, _c_i32p, _c_i64p)
_sig("glp_server_graph_csr", ctypes.c_int, _vp, ctypes.c_int32, _c_i32p, _c_i32p, _c_u32p)
_sig("glp_server_remove_graph", ctypes.c_int, _vp, ctypes.c_int32)
_sig("glp_server_submit_resident", ctypes.c_int, _vp, ctypes.c_int32, ctypes.POINTER(Params), ctypes.c_int32,
     _c_i64p)
_sig("glp_server_free", ctypes.c_int, _vp)


_sig("glp_tuning_set", ctypes.c_int, ctypes.c_char_p, ctypes.c_int32)
_sig("glp_tuning_get", ctypes.c_int, ctypes.c_char_p, _c_i32p)

TUNING_KEYS = ("pack_max_deg", "half_max_deg", "half_lanes", "wave_max_deg", "block_max_deg", "hub_chunk",
               "part_keys", "part_limit", "block_threads", "fused_small")


def lib():
    """The loaded libglp.so handle."""
    if _Library.handle is None:
        _load()
    return _Library.handle


_sig("glp_autotune", ctypes.c_int, _vp, ctypes.POINTER(Params), ctypes.c_int32, ctypes.c_char_p,
     ctypes.POINTER(ctypes.c_double))
_sig("glp_session_set_gamma", ctypes.c_int, _vp, ctypes.c_int32, ctypes.c_int32)
_sig("glp_run_noalloc", ctypes.c_int, _vp, ctypes.POINTER(Params), ctypes.c_int32, _c_i32p, _c_i32p)
_sig("glp_session_set_seeds", ctypes.c_int, _vp, ctypes.c_void_p, ctypes.c_int32)
_sig("glp_run_seeded", ctypes.c_int, _vp, ctypes.POINTER(Params), ctypes.c_void_p, ctypes.c_int32,
     _c_i32p, _c_i32p, ctypes.POINTER(Timing))


def autotune(graph, params, steps=5, out_path=None):
    """Times candidate tunings on `graph`, makes the fastest the process
    default (and writes it to out_path); returns its time for `steps` steps."""
    best = ctypes.c_double(0)
    _check(_lib.glp_autotune(graph.handle, ctypes.byref(params), steps,
                             os.fsencode(out_path) if out_path else None, ctypes.byref(best)))
    return best.value


def tuning_get(key):
    v = ctypes.c_int32()
    _check(_lib.glp_tuning_get(key.encode(), ctypes.byref(v)))
    return v.value


def tuning_set(key, value):
    _check(_lib.glp_tuning_set(key.encode(), int(value)))


class tuning:
    """Context manager: process-wide tuning for graphs uploaded inside it."""

    def __init__(self, **kw):
        self.kw = kw
        self.saved = {}

    def __enter__(self):
        self.saved = {k: tuning_get(k) for k in TUNING_KEYS}
        for k, v in self.kw.items():
            tuning_set(k, v)
        return self

    def __exit__(self, *a):
        for k, v in self.saved.items():
            tuning_set(k, v)


def _check(rc):
    if rc != 0:
        raise GLPError(rc, _lib.glp_last_error().decode(errors="replace"))


def _i32(a):
    return a.ctypes.data_as(_c_i32p)


def _u32(a):
    return a.ctypes.data_as(_c_u32p)


def version():
    return _lib.glp_version().decode()


def device_count():
    c = ctypes.c_int32(0)
    _check(_lib.glp_device_count(ctypes.byref(c)))
    return c.value


def make_params(method=CLASSIC, weighted=False, gamma=(0, 1), seed=1, slpa_T=20, slpa_r=(1, 10),
                cm_rows=3, cm_width=200, cm_aliased=False, table_cap=16, gather=False, profile=False,
                frontier=False):
    p = Params()
    _lib.glp_params_default(ctypes.byref(p))
    p.method = int(method)
    p.weighted = int(bool(weighted))
    p.gamma_num, p.gamma_den = int(gamma[0]), int(gamma[1])
    p.seed = int(seed)
    p.slpa_T = int(slpa_T)
    p.slpa_r_num, p.slpa_r_den = int(slpa_r[0]), int(slpa_r[1])
    p.cm_rows, p.cm_width, p.cm_aliased = int(cm_rows), int(cm_width), int(bool(cm_aliased))
    p.table_cap = int(table_cap)
    p.gather = int(bool(gather))
    p.profile = int(bool(profile))
    p.frontier = int(frontier)  # 2: with the gather's two-step lookahead
    return p


class Graph:
    """Host CSR graph owned by libglp.so (header/graph.h CSRGraph)."""

    def __init__(self, handle):
        self._h = ctypes.c_void_p(handle)

    @property
    def handle(self):
        return self._h

    def __del__(self):
        self.free()

    def free(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.glp_graph_free(self._h)
            self._h = ctypes.c_void_p(None)

    @classmethod
    def load_bin(cls, path, directed=False, weighted=False):
        h = ctypes.c_void_p()
        _check(_lib.glp_graph_load_bin(os.fsencode(path), int(directed), int(weighted), ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def from_csr(cls, offsets, neighbors, weights=None):
        off = np.ascontiguousarray(offsets, dtype=np.int32)
        nbr = np.ascontiguousarray(neighbors, dtype=np.int32)
        n = len(off) - 1
        w = None if weights is None else np.ascontiguousarray(weights, dtype=np.uint32)
        h = ctypes.c_void_p()
        _check(_lib.glp_graph_from_csr(n, len(nbr), _i32(off), _i32(nbr) if len(nbr) else None,
                                       _u32(w) if w is not None and len(w) else None, ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def from_edges(cls, n, edges, weights=None):
        """Undirected graph from (u,v) pairs: symmetrized, self-loops dropped,
        neighbours in pair order (the .bin reader's rule)."""
        edges = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
        keep = edges[:, 0] != edges[:, 1]
        e = edges[keep]
        w = None if weights is None else np.asarray(weights, dtype=np.uint32)[keep]
        src = np.concatenate([e[:, 0], e[:, 1]])
        dst = np.concatenate([e[:, 1], e[:, 0]])
        order = np.concatenate([np.arange(len(e)) * 2, np.arange(len(e)) * 2 + 1])
        idx = np.lexsort((order, src))
        deg = np.bincount(src, minlength=n) if len(src) else np.zeros(n, dtype=np.int64)
        off = np.zeros(n + 1, dtype=np.int64)
        np.cumsum(deg, out=off[1:])
        ww = None if w is None else np.concatenate([w, w])[idx]
        return cls.from_csr(off, dst[idx], ww)

    @classmethod
    def rmat(cls, scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1, degree_sort=True, wmax=0,
             threads=0):
        h = ctypes.c_void_p()
        _check(_lib.glp_gen_rmat(scale, edge_factor, a, b, c, seed, int(degree_sort), wmax, threads,
                                 ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def planted(cls, n, k, deg_in, deg_out, seed=1, threads=0, with_communities=False):
        comm = np.zeros(n, dtype=np.int32)
        h = ctypes.c_void_p()
        _check(_lib.glp_gen_planted(n, k, deg_in, deg_out, seed, threads, _i32(comm), ctypes.byref(h)))
        g = cls(h.value)
        return (g, comm) if with_communities else g

    @classmethod
    def lfr(cls, n, tau1=2.5, kmin=8, kmax=512, tau2=1.5, cmin=20, cmax=5000, mu=0.3, seed=1,
            with_communities=False):
        comm = np.zeros(n, dtype=np.int32)
        h = ctypes.c_void_p()
        _check(_lib.glp_gen_lfr(n, tau1, kmin, kmax, tau2, cmin, cmax, mu, seed, _i32(comm),
                                ctypes.byref(h)))
        g = cls(h.value)
        return (g, comm) if with_communities else g

    def degree_sorted(self):
        h = ctypes.c_void_p()
        perm = np.zeros(self.n, dtype=np.int32)
        _check(_lib.glp_graph_degree_sort(self._h, ctypes.byref(h), _i32(perm)))
        return Graph(h.value), perm

    def info(self):
        n, m, d, w = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32()
        _check(_lib.glp_graph_info(self._h, ctypes.byref(n), ctypes.byref(m), ctypes.byref(d),
                                   ctypes.byref(w)))
        return {"n": n.value, "m": m.value, "max_degree": d.value, "weighted": bool(w.value)}

    def symmetric(self):
        """Whether every edge u->v has an edge v->u (the frontier needs it)."""
        x = ctypes.c_int32()
        _check(_lib.glp_graph_symmetric(self._h, ctypes.byref(x)))
        return bool(x.value)

    @property
    def n(self):
        return self.info()["n"]

    @property
    def m(self):
        return self.info()["m"]

    def csr(self):
        inf = self.info()
        off = np.zeros(inf["n"] + 1, dtype=np.int32)
        nbr = np.zeros(max(inf["m"], 1), dtype=np.int32)
        w = np.zeros(max(inf["m"], 1), dtype=np.uint32) if inf["weighted"] else None
        _check(_lib.glp_graph_csr(self._h, _i32(off), _i32(nbr), _u32(w) if w is not None else None))
        nbr = nbr[:inf["m"]]
        if w is not None:
            w = w[:inf["m"]]
        return off, nbr, w

    def save_bin(self, path):
        _check(_lib.glp_graph_save_bin(self._h, os.fsencode(path)))


def mix_params(seed, i, nmin=1000, nmax=262144, dmin=4.0, dmax=64.0, fmin=0.1, fmax=0.4):
    n, k = ctypes.c_int32(), ctypes.c_int32()
    din, dout, gs = ctypes.c_double(), ctypes.c_double(), ctypes.c_uint64()
    _check(_lib.glp_gen_mix_params(seed, i, nmin, nmax, dmin, dmax, fmin, fmax, ctypes.byref(n),
                                   ctypes.byref(k), ctypes.byref(din), ctypes.byref(dout),
                                   ctypes.byref(gs)))
    return {"n": n.value, "k": k.value, "deg_in": din.value, "deg_out": dout.value, "seed": gs.value}


def stream_u32(seed, t, u0, count):
    out = np.zeros(count, dtype=np.uint32)
    _check(_lib.glp_stream_fill(seed, t, u0, count, _u32(out)))
    return out


def edgelist2bin(txt_path, out_path):
    _check(_lib.glp_tool_edgelist2bin(os.fsencode(txt_path), os.fsencode(out_path)))


def normalize(bin_path, out_path):
    _check(_lib.glp_tool_normalize(os.fsencode(bin_path), os.fsencode(out_path)))


# the id cut-off tools/direct_edgelist2bin.cpp:137 hard-codes
DIRECT_ID_CUTOFF = 5013066


def direct_edgelist2bin(bin_path, out_path, id_cutoff=DIRECT_ID_CUTOFF):
    """Directed .bin writer (tools/direct_edgelist2bin.cpp); id_cutoff <= 0
    keeps every edge."""
    _check(_lib.glp_tool_direct_edgelist2bin(os.fsencode(bin_path), os.fsencode(out_path),
                                             int(id_cutoff)))


def calib_random_read(table_entries=1 << 25, reads=1 << 28, repeats=5):
    """Fastest time (ms) of one launch of `reads` random 4-byte reads from a
    table of table_entries int32 (a power of two): the device's ceiling for
    the label gather's access pattern."""
    ms = ctypes.c_double(0.0)
    _check(_lib.glp_calib_random_read(int(table_entries), int(reads), int(repeats), ctypes.byref(ms)))
    return ms.value


class Session:
    """Device-resident LP state of one job (one HIP stream)."""

    def __init__(self, graph, params=None, **kw):
        self.graph = graph
        self.params = params if params is not None else make_params(**kw)
        h = ctypes.c_void_p()
        _check(_lib.glp_session_create(graph.handle, ctypes.byref(self.params), ctypes.byref(h)))
        self._h = h
        self.n = graph.n

    def __del__(self):
        self.free()

    def free(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.glp_session_free(self._h)
            self._h = ctypes.c_void_p(None)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.free()

    def reset(self):
        _check(_lib.glp_session_reset(self._h))

    def step(self, steps=1, graph=False):
        if graph:
            _check(_lib.glp_session_step_graph(self._h, steps))
        else:
            _check(_lib.glp_session_step(self._h, steps))

    def sync(self):
        _check(_lib.glp_session_sync(self._h))

    def set_gamma(self, num, den):
        _check(_lib.glp_session_set_gamma(self._h, num, den))

    def set_seeds(self, seeds):
        """Seeded LP: seeds is a length-n mask (nonzero = seed); resets the
        labels to v on seeds and -1 elsewhere."""
        m = np.ascontiguousarray(np.asarray(seeds) != 0, dtype=np.uint8)
        _check(_lib.glp_session_set_seeds(self._h, m.ctypes.data if len(m) else None, len(m)))

    def labels(self):
        out = np.zeros(max(self.n, 1), dtype=np.int32)
        _check(_lib.glp_session_labels(self._h, _i32(out)))
        return out[:self.n]

    def changes(self, cap=4096):
        out = np.zeros(cap, dtype=np.int64)
        k = ctypes.c_int32()
        _check(_lib.glp_session_changes(self._h, out.ctypes.data_as(_c_i64p), cap, ctypes.byref(k)))
        return out[:k.value]

    def timing(self):
        t = Timing()
        _check(_lib.glp_session_timing(self._h, ctypes.byref(t)))
        return t.as_dict()

    def volume(self):
        out = np.zeros(max(self.n, 1), dtype=np.int32)
        _check(_lib.glp_session_volume(self._h, _i32(out)))
        return out[:self.n]

    def save(self):
        size = ctypes.c_size_t(0)
        _check(_lib.glp_session_save(self._h, None, ctypes.byref(size)))
        buf = ctypes.create_string_buffer(size.value)
        _check(_lib.glp_session_save(self._h, buf, ctypes.byref(size)))
        return bytes(buf.raw[:size.value])

    def load(self, state):
        buf = ctypes.create_string_buffer(state, len(state))
        _check(_lib.glp_session_load(self._h, buf, len(state)))

    def slpa_memory(self, rows):
        out = np.zeros(rows * max(self.n, 1), dtype=np.int32)
        _check(_lib.glp_session_slpa_memory(self._h, _i32(out)))
        return out[:rows * self.n].reshape(rows, self.n)

    def slpa_post(self, kmax=8):
        dom = np.zeros(max(self.n, 1), dtype=np.int32)
        ns = np.zeros(max(self.n, 1), dtype=np.int32)
        sets = np.zeros(max(self.n * kmax, 1), dtype=np.int32)
        _check(_lib.glp_session_slpa_post(self._h, kmax, _i32(dom), _i32(ns), _i32(sets)))
        return dom[:self.n], ns[:self.n], sets[:self.n * kmax].reshape(self.n, kmax)


def run(graph, steps, params=None, dump=False, seeds=None, **kw):
    """Runs `steps` LP steps from L0[v]=v (seeded LP: from v on seeds, -1
    elsewhere; `seeds` is a length-n mask).  Returns labels (and the per-step
    label dump when dump=True) and the timing record."""
    p = params if params is not None else make_params(**kw)
    n = graph.n
    labels = np.zeros(max(n, 1), dtype=np.int32)
    dump_arr = np.zeros(max(steps * n, 1), dtype=np.int32) if dump else None
    t = Timing()
    if p.method == SEEDED:
        if seeds is None:
            raise GLPError(-1, "seeded LP needs seeds")
        m = np.ascontiguousarray(np.asarray(seeds) != 0, dtype=np.uint8)
        if len(m) != n:
            raise GLPError(-1, "seeds hold %d entries, the graph has %d vertices" % (len(m), n))
        _check(_lib.glp_run_seeded(graph.handle, ctypes.byref(p), m.ctypes.data if n else None, steps,
                                   _i32(labels), _i32(dump_arr) if dump else None, ctypes.byref(t)))
    else:
        _check(_lib.glp_run(graph.handle, ctypes.byref(p), steps, _i32(labels),
                            _i32(dump_arr) if dump else None, ctypes.byref(t)))
    labels = labels[:n]
    if dump:
        return labels, dump_arr[:steps * n].reshape(steps, n), t.as_dict()
    return labels, t.as_dict()


def run_noalloc(graph, steps, params=None, dump=False, **kw):
    """run() for classic and layered LP through the allocation-free entry
    point (all device state in the library's static array); labels, and the
    per-step dump when dump=True."""
    p = params if params is not None else make_params(**kw)
    n = graph.n
    labels = np.zeros(max(n, 1), dtype=np.int32)
    dump_arr = np.zeros(max(steps * n, 1), dtype=np.int32) if dump else None
    _check(_lib.glp_run_noalloc(graph.handle, ctypes.byref(p), steps, _i32(labels),
                                _i32(dump_arr) if dump else None))
    if dump:
        return labels[:n], dump_arr[:steps * n].reshape(steps, n)
    return labels[:n]


class Batch:
    """Many graphs stepped together as one disjoint union (one launch per tier per step)."""

    def __init__(self, graphs, params=None, **kw):
        self.graphs = list(graphs)
        self.params = params if params is not None else make_params(**kw)
        arr = (ctypes.c_void_p * len(self.graphs))(*[g.handle.value for g in self.graphs])
        h = ctypes.c_void_p()
        _check(_lib.glp_batch_create(len(self.graphs), ctypes.cast(arr, _pp), ctypes.byref(self.params),
                                     ctypes.byref(h)))
        self._h = h

    def __del__(self):
        self.free()

    def free(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.glp_batch_free(self._h)
            self._h = ctypes.c_void_p(None)

    def run(self, steps, mode=0):
        ms = ctypes.c_double(0)
        _check(_lib.glp_batch_run(self._h, steps, mode, ctypes.byref(ms)))
        return ms.value

    def reset(self):
        _check(_lib.glp_batch_reset(self._h))

    def autotune(self, steps=3):
        best = ctypes.c_double(0)
        _check(_lib.glp_batch_autotune(self._h, steps, ctypes.byref(best)))
        return best.value

    def labels(self, i):
        n = self.graphs[i].n
        out = np.zeros(max(n, 1), dtype=np.int32)
        _check(_lib.glp_batch_labels(self._h, i, _i32(out)))
        return out[:n]


class Server:
    """Continuous-batching LP request loop under an HBM budget."""

    def __init__(self, budget_bytes, max_slots=64):
        h = ctypes.c_void_p()
        _check(_lib.glp_server_create(int(budget_bytes), int(max_slots), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        self.close()

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _check(_lib.glp_server_free(self._h))
            self._h = ctypes.c_void_p(None)

    def submit(self, graph, steps, params=None, **kw):
        p = params if params is not None else make_params(**kw)
        jid = ctypes.c_int64()
        _check(_lib.glp_server_submit(self._h, graph.handle, ctypes.byref(p), steps, ctypes.byref(jid)))
        return jid.value

    def poll(self, jid):
        st = JobStats()
        _check(_lib.glp_server_poll(self._h, jid, ctypes.byref(st)))
        return st

    def wait(self, jid, n):
        out = np.zeros(max(n, 1), dtype=np.int32)
        st = JobStats()
        _check(_lib.glp_server_wait(self._h, jid, _i32(out), ctypes.byref(st)))
        return out[:n], st

    def cancel(self, jid):
        _check(_lib.glp_server_cancel(self._h, jid))

    def release(self, jid):
        """Forgets an ended job's labels and stats."""
        _check(_lib.glp_server_release(self._h, jid))

    def add_graph(self, graph):
        """Uploads `graph` once for any number of jobs; returns its id."""
        gid = ctypes.c_int32()
        _check(_lib.glp_server_add_graph(self._h, graph.handle, ctypes.byref(gid)))
        return gid.value

    def add_rmat(self, scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1, wmax=0):
        """Builds Graph.rmat(scale, ..., degree_sort=False, wmax)'s graph in
        HBM; returns its id."""
        gid = ctypes.c_int32()
        _check(_lib.glp_server_add_rmat(self._h, scale, edge_factor, a, b, c, seed, wmax, ctypes.byref(gid)))
        return gid.value

    def graph_info(self, gid):
        n, m, d, b = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int64()
        _check(_lib.glp_server_graph_info(self._h, gid, ctypes.byref(n), ctypes.byref(m), ctypes.byref(d),
                                          ctypes.byref(b)))
        return {"n": n.value, "m": m.value, "max_degree": d.value, "bytes": b.value}

    def graph_csr(self, gid, weights=False):
        inf = self.graph_info(gid)
        off = np.zeros(inf["n"] + 1, dtype=np.int32)
        nbr = np.zeros(max(inf["m"], 1), dtype=np.int32)
        w = np.zeros(max(inf["m"], 1), dtype=np.uint32) if weights else None
        _check(_lib.glp_server_graph_csr(self._h, gid, _i32(off), _i32(nbr), _u32(w) if weights else None))
        return off, nbr[:inf["m"]], (w[:inf["m"]] if weights else None)

    def remove_graph(self, gid):
        _check(_lib.glp_server_remove_graph(self._h, gid))

    def submit_resident(self, gid, steps, params=None, **kw):
        p = params if params is not None else make_params(**kw)
        jid = ctypes.c_int64()
        _check(_lib.glp_server_submit_resident(self._h, gid, ctypes.byref(p), steps, ctypes.byref(jid)))
        return jid.value
