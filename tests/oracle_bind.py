"""ctypes binding of oracle/liboracle.so (test infrastructure: the checker)."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_LIB = None

ALIAS_DTYPE = np.dtype([("prob", "<f8"), ("first", "<i4"), ("second", "<i4")])

APP = {"deepwalk": 0, "ppr": 1, "node2vec": 2, "metapath": 3, "genericwalk": 4}


class OrcSfmt(C.Structure):
    _fields_ = [("st", C.c_uint32 * 624), ("idx", C.c_int32)]


class OrcGraph(C.Structure):
    _fields_ = [("nv", C.c_int32), ("ne", C.c_int64), ("offset", C.c_void_p), ("adj", C.c_void_p),
                ("degree", C.c_void_p), ("edge_weight", C.c_void_p), ("edge_label", C.c_void_p),
                ("alias", C.c_void_p), ("prefix", C.c_void_p), ("rs_weight", C.c_void_p),
                ("rs_max", C.c_void_p)]


class OrcParams(C.Structure):
    _fields_ = [("app", C.c_int32), ("execution", C.c_int32), ("sampling", C.c_int32), ("length", C.c_int32),
                ("stop_prob", C.c_double), ("p", C.c_double), ("q", C.c_double), ("schema", C.c_void_p),
                ("schema_length", C.c_int32), ("max_trials", C.c_int32), ("n2v_weighted", C.c_int32)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(ROOT, "oracle", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        _LIB = C.CDLL(path)
        vp = C.c_void_p
        _LIB.orc_sfmt_genrand_uint32.restype = C.c_uint32
        _LIB.orc_sfmt_genrand_uint32.argtypes = [vp]
        _LIB.orc_sfmt_init_gen_rand.argtypes = [vp, C.c_uint32]
        _LIB.orc_sfmt_init_by_array.argtypes = [vp, vp, C.c_int32]
        _LIB.orc_sfmt_fill.argtypes = [vp, vp, C.c_int64]
        _LIB.orc_philox4x32_10.argtypes = [vp, vp, vp]
        _LIB.orc_alias_build.argtypes = [vp, vp, vp]
        _LIB.orc_its_build.argtypes = [vp, vp, vp]
        _LIB.orc_rs_build.argtypes = [vp, vp, vp]
        _LIB.orc_is_neighbor.argtypes = [vp, C.c_int32, C.c_int32]
        _LIB.orc_to_real1.restype = C.c_double
        _LIB.orc_to_real1.argtypes = [C.c_uint32]
        _LIB.orc_to_real2.restype = C.c_double
        _LIB.orc_to_real2.argtypes = [C.c_uint32]
        _LIB.orc_sfmt_stream.argtypes = [C.c_uint32, C.c_uint64, C.c_int64, C.c_void_p]
        _LIB.orc_alias_build.restype = C.c_int32
        _LIB.orc_upper_bound.restype = C.c_int32
        _LIB.orc_upper_bound.argtypes = [C.c_void_p, C.c_int32, C.c_double]
        _LIB.orc_is_neighbor.restype = C.c_int32
        _LIB.orc_walk.restype = C.c_int32
        _LIB.orc_walk.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_int64,
                                  C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p]
        _LIB.orc_ppr_power.restype = C.c_int32
        _LIB.orc_ppr_power.argtypes = [C.c_void_p, C.c_int32, C.c_double, C.c_void_p, C.c_int32, C.c_double,
                                       C.c_void_p]
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data


def sfmt_stream(seed, n, skip=0):
    out = np.empty(n, dtype=np.uint32)
    lib().orc_sfmt_stream(seed, skip, n, _p(out))
    return out


def sfmt_by_array(key, n):
    s = OrcSfmt()
    k = np.asarray(key, dtype=np.uint32)
    lib().orc_sfmt_init_by_array(C.byref(s), _p(k), C.c_int32(len(k)))
    out = np.empty(n, dtype=np.uint32)
    lib().orc_sfmt_fill(C.byref(s), _p(out), C.c_int64(n))
    return out


def real1(u):
    return lib().orc_to_real1(int(u))


def real2(u):
    return lib().orc_to_real2(int(u))


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.empty(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_p(c), _p(k), _p(out))
    return out


class Graph:
    """Keeps numpy arrays alive behind an OrcGraph."""

    def __init__(self, g, alias=None, prefix=None, rs_weight=None, rs_max=None):
        self.g = g
        self.alias, self.prefix, self.rs_weight, self.rs_max = alias, prefix, rs_weight, rs_max
        self.s = OrcGraph(g.nv, g.ne, _p(g.offset), _p(g.adj), _p(g.degree), _p(g.weight), _p(g.label),
                          _p(alias), _p(prefix), _p(rs_weight), _p(rs_max))


def alias_build(g, w):
    out = np.zeros(g.ne, dtype=ALIAS_DTYPE)
    zero = lib().orc_alias_build(C.byref(Graph(g).s), _p(np.ascontiguousarray(w, np.float64)), _p(out))
    return out, zero


def its_build(g, w):
    out = np.zeros(g.ne, dtype=np.float64)
    lib().orc_its_build(C.byref(Graph(g).s), _p(np.ascontiguousarray(w, np.float64)), _p(out))
    return out


def rs_build(g, w):
    out = np.zeros(g.nv, dtype=np.float64)
    lib().orc_rs_build(C.byref(Graph(g).s), _p(np.ascontiguousarray(w, np.float64)), _p(out))
    return out


def walk(graph, app, execution, sampling, length, sources, draws, K, term_draws=None, KT=0, stop_prob=0.2,
         p=2.0, q=0.5, schema=None, max_trials=0, weighted=False):
    """draws: walker-major (n_walkers, K) uint32; returns dict of outputs.
    weighted: Node2Vec over edge weights (factor times w_e, bound max w times max_)."""
    sources = np.ascontiguousarray(sources, dtype=np.int32)
    n = len(sources)
    L = max(length, 1)
    paths = np.full((n, L), -1, dtype=np.int32)
    lengths = np.zeros(n, np.int32)
    ends = np.zeros(n, np.int32)
    status = np.zeros(n, np.int32)
    used = np.zeros(n, np.int32)
    steps = np.zeros(1, np.uint64)
    sch = None if schema is None else np.ascontiguousarray(schema, dtype=np.int32)
    prm = OrcParams(APP[app] if isinstance(app, str) else app, execution, sampling, length, stop_prob, p, q,
                    _p(sch), 0 if sch is None else len(sch), max_trials, int(bool(weighted)))
    d = np.ascontiguousarray(draws, dtype=np.uint32)
    t = None if term_draws is None else np.ascontiguousarray(term_draws, dtype=np.uint32)
    rc = lib().orc_walk(C.addressof(graph.s), C.addressof(prm), _p(sources), n, _p(d), K, _p(t), KT, _p(paths),
                        _p(lengths), _p(ends), _p(status), _p(used), _p(steps))
    if rc != 0:
        raise ValueError("oracle rejected the walk parameters")
    return dict(paths=paths, lengths=lengths, ends=ends, status=status, draws_used=used, steps=int(steps[0]))


def ppr_power(g, alpha, sources, weighted=False, tol=1e-12):
    out = np.zeros(g.nv, dtype=np.float64)
    src = np.ascontiguousarray(sources, dtype=np.int32)
    it = lib().orc_ppr_power(C.byref(Graph(g).s), int(weighted), alpha, _p(src), len(src), tol, _p(out))
    return out, it
