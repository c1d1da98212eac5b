"""Python host side over the C ABI: contexts, device graphs, walks.

This mirrors the reference's app-facing surface for the step engine:
  * DeviceGraph       Graph + static_initialization (util/graph/graph.h, creeper.h:102-175)
  * walk()            compute<F>(graph, walkers, f) for a built-in app (creeper.h:689)
  * sources_*()       walker_generator.cpp:9-63
Device memory, copies and timing go through the library's own entry points, so a
process holds one HIP runtime (the one libthunderrw.so links); every walk runs in
the HIP kernels of libthunderrw.so.
"""
import ctypes as C

import numpy as np

from . import _abi as A
from ._abi import check, lib

_EXEC = {"uniform": A.EXEC_UNIFORM, "static": A.EXEC_STATIC, "dynamic": A.EXEC_DYNAMIC}
_SM = {"uniform": A.SM_UNIFORM, "its": A.SM_ITS, "alias": A.SM_ALIAS, "rejection": A.SM_REJECTION,
       "max_weight_rejection": A.SM_MAX_WEIGHT_REJECTION}
_RNG = {"replay": A.RNG_REPLAY, "philox": A.RNG_PHILOX}
_FORM = {"auto": A.FORM_AUTO, "fused": A.FORM_FUSED, "queue": A.FORM_QUEUE, "step": A.FORM_STEP,
         "graph": A.FORM_GRAPH}

ALIAS_DTYPE = np.dtype([("alias_value", "<f8"), ("first", "<i4"), ("second", "<i4")])

# Reference defaults per app (g_para of apps/*.cpp:7-8 and execute()).
APP_DEFAULTS = {
    "deepwalk": ("static", "alias", 80),
    "genericwalk": ("static", "its", 80),
    "ppr": ("uniform", "uniform", 1),
    "node2vec": ("dynamic", "max_weight_rejection", 80),
    "metapath": ("dynamic", "its", 80),
}


def _ptr(a):
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data


def _enum(table, v):
    return table[v] if isinstance(v, str) else int(v)


def device_count():
    n = C.c_int32(0)
    check(lib().trw_device_count(C.byref(n)))
    return n.value


class Context:
    """One device and its stream (replaces the reference's process globals)."""

    def __init__(self, device=0):
        self.h = C.c_void_p()
        check(lib().trw_context_create(device, C.byref(self.h)))
        self.device = device

    @property
    def stream(self):
        s = C.c_void_p()
        check(lib().trw_context_stream(self.h, C.byref(s)))
        return s.value

    def close(self):
        if self.h:
            lib().trw_context_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class HostGraph:
    """A CSR graph in host memory owned by the library (generator / reference files)."""

    def __init__(self, handle):
        self.h = handle

    @classmethod
    def chung_lu(cls, num_vertices, undirected_edges, gamma, dmin, dmax, seed, weights=True, n_labels=0):
        h = C.c_void_p()
        check(lib().trw_host_graph_chung_lu(num_vertices, undirected_edges, gamma, dmin, dmax, seed, int(weights),
                                            n_labels, C.byref(h)))
        return cls(h)

    @classmethod
    def load(cls, directory, weights=False, labels=False):
        h = C.c_void_p()
        check(lib().trw_host_graph_load(directory.encode(), int(weights), int(labels), C.byref(h)))
        return cls(h)

    @classmethod
    def load_edge_list(cls, directory, skip="#"):
        """<directory>/b_edge_list.bin as text edge list (Graph::load_edge_list semantics)."""
        h = C.c_void_p()
        check(lib().trw_host_graph_load_edge_list(directory.encode(), ord(skip), C.byref(h)))
        return cls(h)

    def store(self, directory):
        check(lib().trw_host_graph_store(self.h, directory.encode()))

    def arrays(self):
        """Zero-copy numpy views (degree, adj, weight or None, label or None); valid while self lives."""
        nv, ne = C.c_int32(), C.c_int64()
        d, a, w, l_ = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
        check(lib().trw_host_graph_arrays(self.h, C.byref(nv), C.byref(ne), C.byref(d), C.byref(a), C.byref(w),
                                          C.byref(l_)))

        def view(p, n, ct, dt, optional=True):
            if not p.value or n == 0:
                return None if (optional and not p.value) else np.zeros(0, dt)
            return np.ctypeslib.as_array(C.cast(p.value, C.POINTER(ct)), shape=(n,))

        return (view(d, nv.value, C.c_int32, np.int32, False), view(a, ne.value, C.c_int32, np.int32, False),
                view(w, ne.value, C.c_double, np.float64), view(l_, ne.value, C.c_int32, np.int32))

    def close(self):
        if self.h:
            lib().trw_host_graph_free(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceGraph:
    """The CSR graph resident in HBM (Graph::load_csr, util/graph/graph.cpp:45-101)."""

    def __init__(self, ctx, degree, adj, weight=None, label=None):
        degree = np.ascontiguousarray(degree, dtype=np.int32)
        adj = np.ascontiguousarray(adj, dtype=np.int32)
        weight = None if weight is None else np.ascontiguousarray(weight, dtype=np.float64)
        label = None if label is None else np.ascontiguousarray(label, dtype=np.int32)
        self.ctx = ctx
        self.h = C.c_void_p()
        check(lib().trw_graph_upload(ctx.h, len(degree), len(adj), _ptr(degree), _ptr(adj), _ptr(weight),
                                     _ptr(label), C.byref(self.h)))
        self.num_vertices, self.num_edges = len(degree), len(adj)

    def info(self):
        nv, ne, md = C.c_int32(), C.c_int64(), C.c_int32()
        check(lib().trw_graph_info(self.h, C.byref(nv), C.byref(ne), C.byref(md)))
        return nv.value, ne.value, md.value

    def build_tables(self, sampling, weights="edge"):
        w = A.WEIGHTS_EDGE if weights == "edge" else A.WEIGHTS_UNIT
        check(lib().trw_build_tables(self.h, _enum(_SM, sampling), w))

    def build_membership(self):
        """trw_build_membership: the hashed neighbour sets is_neighbor is answered from on the device."""
        check(lib().trw_build_membership(self.h))

    def store_tables(self, directory):
        """Writes the built tables as the reference's b_edge_weight_*.bin files."""
        check(lib().trw_graph_store_tables(self.h, directory.encode()))

    def load_tables(self, directory, sampling, weights="edge"):
        """Reads one of the reference's table files in place of a build."""
        w = A.WEIGHTS_EDGE if weights == "edge" else A.WEIGHTS_UNIT
        check(lib().trw_graph_load_tables(self.h, directory.encode(), _enum(_SM, sampling), w))

    def last_build_ms(self):
        ms = C.c_double()
        check(lib().trw_last_build_ms(self.h, C.byref(ms)))
        return ms.value

    TUNING = {"walkers_per_thread": A.TUNE_WALKERS_PER_THREAD, "its_tree_min_degree": A.TUNE_ITS_TREE_MIN_DEGREE,
              "its_guide_off": A.TUNE_ITS_GUIDE_OFF,
              "queue_chunk": A.TUNE_QUEUE_CHUNK}
    SHAPES = {"fused_block": A.TUNE_FUSED_BLOCK, "fused_tile": A.TUNE_FUSED_TILE,
              "queue_blocks_per_cu": A.TUNE_QUEUE_BLOCKS_PER_CU, "n2v_block": A.TUNE_N2V_BLOCK,
              "graph_burst": A.TUNE_GRAPH_BURST}

    def set_tuning(self, move=None, **knobs):
        """Launch choices that change speed, never results (trw_graph_set_tuning); the shape
        knobs (SHAPES) for every Move, or for one Move ("alias", ...) when move is given."""
        for k, v in knobs.items():
            if k in self.SHAPES and move is not None:
                check(lib().trw_graph_set_move_tuning(self.h, A.MOVES.index(move), self.SHAPES[k], int(v)))
            else:
                check(lib().trw_graph_set_tuning(self.h, {**self.TUNING, **self.SHAPES}[k], int(v)))

    def autotune(self, params, sample_walkers=1 << 20, seed=0, with_paths=True):
        """trw_autotune: times the applicable knobs' candidates on a sample batch and keeps the
        fastest on the graph; returns {"chosen": {knob: value}, "default_steps_per_s", "tuned_steps_per_s"}."""
        rep = A.TuneReport()
        check(lib().trw_autotune(self.h, C.byref(params), int(sample_walkers), int(seed), int(bool(with_paths)),
                                 C.byref(rep)))
        return self._report(rep)

    def autotune_mix(self, params_a, params_b, split, sample_walkers=1 << 20, seed=0, with_paths=True):
        """trw_autotune_mix: the same for a mixed QUEUE-form batch, walkers [0, split) of the
        sample as params_a and the rest as params_b."""
        rep = A.TuneReport()
        check(lib().trw_autotune_mix(self.h, C.byref(params_a), C.byref(params_b), int(split), int(sample_walkers),
                                     int(seed), int(bool(with_paths)), C.byref(rep)))
        return self._report(rep)

    def _report(self, rep):
        names = {v: k for k, v in {**self.TUNING, **self.SHAPES}.items()}
        return {"move": A.MOVES[rep.move], "chosen": {names[rep.knob[i]]: rep.chosen[i] for i in range(rep.knobs)},
                "default_steps_per_s": rep.default_steps_per_s, "tuned_steps_per_s": rep.tuned_steps_per_s}

    def tuning(self, move=None):
        """The per-graph knobs, and the shape knobs of one Move when move is given."""
        out = {}
        for k, knob in self.TUNING.items():
            v = C.c_int64()
            check(lib().trw_graph_get_tuning(self.h, knob, C.byref(v)))
            out[k] = v.value
        if move is not None:
            for k, knob in self.SHAPES.items():
                v = C.c_int64()
                check(lib().trw_graph_get_move_tuning(self.h, A.MOVES.index(move), knob, C.byref(v)))
                out[k] = v.value
        return out

    def alias_table(self):
        out = np.zeros(self.num_edges, dtype=ALIAS_DTYPE)
        check(lib().trw_copy_alias_table(self.h, _ptr(out)))
        return out

    def prefix_sum(self):
        out = np.zeros(self.num_edges, dtype=np.float64)
        check(lib().trw_copy_prefix_sum(self.h, _ptr(out)))
        return out

    def rejection_max(self):
        out = np.zeros(self.num_vertices, dtype=np.float64)
        check(lib().trw_copy_rejection_max(self.h, _ptr(out)))
        return out

    def close(self):
        if self.h:
            lib().trw_graph_free(self.h)
            self.h = C.c_void_p()


def make_params(app, execution=None, sampling=None, length=None, stop_probability=0.2, p=2.0, q=0.5, schema=None,
                rng="philox", seed=0, form="auto", max_trials=0, walkers_per_thread=0, weights="app"):
    """g_para + app functor arguments.  Unspecified fields take the app's reference defaults.

    p and q are Node2Vec's command-line values; the functor's weights are 1/p and 1/q
    (node2vec.cpp:64-76), which is what the C ABI takes.
    """
    name = app if isinstance(app, str) else {v: k for k, v in A.APPS.items()}[app]
    de, ds, dl = APP_DEFAULTS[name]
    prm = A.WalkParams()
    prm.app = A.APPS[name]
    prm.execution = _enum(_EXEC, execution if execution is not None else de)
    prm.sampling = _enum(_SM, sampling if sampling is not None else ds)
    prm.length = dl if length is None else int(length)
    prm.stop_probability = stop_probability
    prm.p, prm.q = 1.0 / p, 1.0 / q
    sch = list(schema or [])
    if len(sch) > 16:
        raise ValueError("schema longer than 16")
    for i, s in enumerate(sch):
        prm.schema[i] = int(s)
    prm.schema_length = len(sch)
    prm.rng = _enum(_RNG, rng)
    prm.seed = int(seed)
    prm.form = _enum(_FORM, form)
    prm.max_trials = int(max_trials)
    prm.walkers_per_thread = int(walkers_per_thread)
    prm.weights = {"app": A.WALK_WEIGHTS_APP, "edge": A.WALK_WEIGHTS_EDGE, "unit": A.WALK_WEIGHTS_UNIT}[weights]
    return prm


def make_io(sources, n_walkers, path_stride=0, walker_id_base=0, draws=None, draws_per_walker=0, term_draws=None,
            term_draws_per_walker=0, paths=None, lengths=None, ends=None, status=None, draws_used=None,
            counters=None, walker_ids=None, unpadded=False, path_stride_b=0):
    io = A.WalkIO()
    io.sources = _ptr(sources)
    io.n_walkers = int(n_walkers)
    io.path_stride = int(path_stride)
    io.walker_id_base = int(walker_id_base)
    io.draws = _ptr(draws)
    io.draws_per_walker = int(draws_per_walker)
    io.term_draws = _ptr(term_draws)
    io.term_draws_per_walker = int(term_draws_per_walker)
    io.paths, io.lengths, io.ends = _ptr(paths), _ptr(lengths), _ptr(ends)
    io.status, io.draws_used, io.counters = _ptr(status), _ptr(draws_used), _ptr(counters)
    io.walker_ids = _ptr(walker_ids)
    io.path_flags = A.PATHS_UNPADDED if unpadded else A.PATHS_PADDED
    io.path_stride_b = int(path_stride_b)
    return io


def walk_async(graph, params, io, stream=None):
    check(lib().trw_walk_async(graph.h, C.byref(params), C.byref(io), stream))


def walk_noalloc(graph, params, io, stream=None):
    """trw_walk_noalloc: trw_walk_async that allocates nothing (tables must be built; STEP form
    at most 2^18 walkers; calls on one device must not overlap)."""
    check(lib().trw_walk_noalloc(graph.h, C.byref(params), C.byref(io), stream))


def walk_mix_async(graph, params_a, params_b, split, io, stream=None):
    check(lib().trw_walk_mix_async(graph.h, C.byref(params_a), C.byref(params_b), int(split), C.byref(io), stream))


def replay_fill(seed, n_walkers, draws_per_walker, first_walker=0):
    """Replay draws in the device layout [k][walker]."""
    out = np.empty((int(draws_per_walker), int(n_walkers)), dtype=np.uint32)
    check(lib().trw_replay_fill(seed, first_walker, n_walkers, draws_per_walker, _ptr(out)))
    return out


def sources_all(degree):
    degree = np.ascontiguousarray(degree, dtype=np.int32)
    out = np.empty(len(degree), dtype=np.int32)
    cnt = C.c_int32()
    check(lib().trw_sources_all(_ptr(degree), len(degree), _ptr(out), C.byref(cnt)))
    return out[:cnt.value].copy()


def sources_single(degree, n):
    degree = np.ascontiguousarray(degree, dtype=np.int32)
    out = np.empty(n, dtype=np.int32)
    check(lib().trw_sources_single(_ptr(degree), len(degree), n, _ptr(out)))
    return out


def sources_random(degree, n, seed=0):
    degree = np.ascontiguousarray(degree, dtype=np.int32)
    out = np.empty(n, dtype=np.int32)
    check(lib().trw_sources_random(_ptr(degree), len(degree), n, seed, _ptr(out)))
    return out


class DeviceBuffer:
    """A buffer in HBM owned through the C ABI (trw_device_alloc)."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        check(lib().trw_device_alloc(ctx.h, self.nbytes, C.byref(self.ptr)))

    @classmethod
    def from_array(cls, ctx, arr):
        arr = np.ascontiguousarray(arr)
        buf = cls(ctx, arr.nbytes)
        check(lib().trw_copy_to_device(ctx.h, buf.ptr, arr.ctypes.data, arr.nbytes))
        return buf

    @classmethod
    def filled(cls, ctx, shape, dtype, value):
        return cls.from_array(ctx, np.full(shape, value, dtype=dtype))

    def data_ptr(self):
        return self.ptr.value

    def to_numpy(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        check(lib().trw_copy_to_host(self.ctx.h, out.ctypes.data, self.ptr, out.nbytes))
        return out

    def zero(self, stream=None):
        check(lib().trw_device_zero(self.ctx.h, self.ptr, self.nbytes, stream))

    def free(self):
        if self.ptr.value:
            lib().trw_device_free(self.ctx.h, self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize(ctx, stream=None):
    check(lib().trw_synchronize(ctx.h, stream))


class DeviceTimer:
    """Device time between begin() and end() from HIP events on the given stream."""

    def __init__(self, ctx, stream=None):
        self.ctx, self.stream = ctx, stream

    def begin(self):
        check(lib().trw_timer_begin(self.ctx.h, self.stream))

    def end(self):
        ms = C.c_double()
        check(lib().trw_timer_end(self.ctx.h, self.stream, C.byref(ms)))
        return ms.value


def walk(graph, params, sources, draws=None, term_draws=None, want_paths=True, walker_id_base=0, stream=None,
         mix=None, walker_ids=None, noalloc=False, unpadded=False, kind_rows=False):
    """Runs one batch in HBM buffers and returns numpy results.

    draws / term_draws: replay streams in the [k][walker] layout, or None (Philox).
    mix: optional (params_b, split) for a two-kind batch.
    walker_ids: optional per-walker global ids keying the Philox stream (uint64).
    noalloc: run the batch through trw_walk_noalloc instead of trw_walk_async.
    unpadded: TRW_PATHS_UNPADDED -- entries past a walker's length are left as they were (the
    buffer starts filled with -2 here).
    kind_rows: with mix, each kind's rows as long as its own walks (io.path_stride_b); the
    result then holds paths_a (split x length_a) and paths_b ((n - split) x length_b).
    """
    ctx = graph.ctx
    src_h = np.ascontiguousarray(sources, dtype=np.int32)
    n = len(src_h)
    L = max(params.length, mix[0].length if mix else 0, 1)
    bufs = []

    def up(a):
        b = DeviceBuffer.from_array(ctx, a)
        bufs.append(b)
        return b

    try:
        src = up(src_h if n else np.zeros(1, np.int32))
        d_draws = None if draws is None else up(np.ascontiguousarray(draws, dtype=np.uint32))
        d_tdraws = None if term_draws is None else up(np.ascontiguousarray(term_draws, dtype=np.uint32))
        K = 0 if draws is None else draws.shape[0]
        KT = 0 if term_draws is None else term_draws.shape[0]
        m = max(n, 1)
        per_kind = bool(mix) and kind_rows
        if per_kind:
            split = int(mix[1])
            La, Lb = max(params.length, 1), max(mix[0].length, 1)
            paths = up(np.full(max(split * La + (n - split) * Lb, 1), -2, np.int32)) if want_paths else None
        else:
            paths = up(np.full((m, L), -2, np.int32)) if want_paths else None
        lengths = up(np.zeros(m, np.int32))
        ends = up(np.zeros(m, np.int32))
        status = up(np.full(m, -9, np.int32))
        used = up(np.zeros(m, np.int32))
        counters = up(np.zeros(4, np.uint64))
        ids = None if walker_ids is None else up(np.ascontiguousarray(walker_ids, dtype=np.uint64))
        io = make_io(src, n, La if per_kind else L, walker_id_base, d_draws, K, d_tdraws, KT, paths, lengths, ends,
                     status, used, counters, walker_ids=ids, unpadded=unpadded, path_stride_b=Lb if per_kind else 0)
        if mix is None:
            (walk_noalloc if noalloc else walk_async)(graph, params, io, stream)
        else:
            if noalloc:
                raise ValueError("trw_walk_noalloc runs one walker kind per batch")
            walk_mix_async(graph, params, mix[0], mix[1], io, stream)
        synchronize(ctx, stream)
        c = counters.to_numpy(np.uint64, 4)
        if per_kind:
            flat = None if paths is None else paths.to_numpy(np.int32, split * La + (n - split) * Lb)
            extra = dict(paths_a=None if flat is None else flat[:split * La].reshape(split, La),
                         paths_b=None if flat is None else flat[split * La:].reshape(n - split, Lb))
        else:
            extra = {}
        return dict(paths=None if paths is None or per_kind else paths.to_numpy(np.int32, (n, L)), **extra,
                    lengths=lengths.to_numpy(np.int32, n), ends=ends.to_numpy(np.int32, n),
                    status=status.to_numpy(np.int32, n), draws_used=used.to_numpy(np.int32, n),
                    steps=int(c[0]), gather_terminated=int(c[1]), dead_ends=int(c[2]), errors=int(c[3]))
    finally:
        for b in bufs:
            b.free()


def _struct_dict(st):
    return {name: getattr(st, name) for name, _ in st._fields_}


class Server:
    """Continuous-batching query server over one DeviceGraph (trw_server_*).

    Build the tables the queries sample from before creating it.  Queries draw
    from the Philox stream: give each one a disjoint walker-id range (the default
    is query_number << 32) and its result is the same alone or in a mix."""

    def __init__(self, graph, streams=4, max_batch_walkers=1 << 20, path_bytes_per_batch=64 << 20,
                 max_pending_queries=1 << 16, completion_threads=4, max_queries_per_batch=1024):
        self.graph = graph
        cfg = A.ServerConfig(streams, max_batch_walkers, path_bytes_per_batch, max_pending_queries, completion_threads,
                             max_queries_per_batch, 0)
        self.h = C.c_void_p()
        check(lib().trw_server_create(graph.h, C.byref(cfg), C.byref(self.h)))
        self._count = 0

    def close(self):
        if self.h:
            check(lib().trw_server_destroy(self.h))
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def query(self, params, sources, n_walkers, walker_id_base=None, result="ppr"):
        """A trw_query struct (keeps a reference to its source array)."""
        src = np.ascontiguousarray(sources, dtype=np.int32)
        if walker_id_base is None:
            self._count += 1
            walker_id_base = self._count << 32
        q = A.Query(_ptr(src), len(src), int(n_walkers), int(walker_id_base), params,
                    A.RESULT_PPR if result == "ppr" else A.RESULT_PATHS, 0)
        q._src = src
        return q

    def submit(self, params, sources, n_walkers, walker_id_base=None, result="ppr"):
        q = self.query(params, sources, n_walkers, walker_id_base, result)
        qid = C.c_uint64()
        check(lib().trw_server_submit(self.h, C.byref(q), C.byref(qid)))
        return qid.value

    def wait(self, qid, timeout_ms=-1.0):
        info = A.QueryInfo()
        check(lib().trw_server_wait(self.h, qid, float(timeout_ms), C.byref(info)))
        return _struct_dict(info)

    def info(self, qid):
        info = A.QueryInfo()
        check(lib().trw_server_query_info(self.h, qid, C.byref(info)))
        return _struct_dict(info)

    def ppr(self, qid):
        """(vertices ascending, walker counts) of a done PPR query."""
        n = self.info(qid)["ppr_entries"]
        v = np.empty(max(n, 1), np.int32)
        c = np.empty(max(n, 1), np.uint32)
        check(lib().trw_server_ppr_result(self.h, qid, _ptr(v), _ptr(c), n))
        return v[:n], c[:n]

    def walk_result(self, qid, length=None):
        info = self.info(qid)
        n = info["n_walkers"]
        m = max(n, 1)
        lengths, ends, status = np.empty(m, np.int32), np.empty(m, np.int32), np.empty(m, np.int32)
        paths = None if length is None else np.empty((m, max(length, 1)), np.int32)
        check(lib().trw_server_walk_result(self.h, qid, _ptr(paths), _ptr(lengths), _ptr(ends), _ptr(status)))
        return dict(paths=None if paths is None else paths[:n], lengths=lengths[:n], ends=ends[:n],
                    status=status[:n])

    def release(self, qid):
        check(lib().trw_server_release(self.h, qid))

    def stats(self):
        st = A.ServerStats()
        check(lib().trw_server_stats_get(self.h, C.byref(st)))
        return _struct_dict(st)

    def run_open_loop(self, queries, qps, seed=0):
        """Submits trw_query structs at Poisson arrivals of rate qps; returns
        (report dict, query ids, latencies in ms)."""
        n = len(queries)
        arr = (A.Query * max(n, 1))(*queries)
        ids = np.zeros(max(n, 1), np.uint64)
        lat = np.zeros(max(n, 1), np.float64)
        rep = A.OpenLoopReport()
        check(lib().trw_server_run_open_loop(self.h, arr, n, float(qps), int(seed), _ptr(ids), _ptr(lat),
                                             C.byref(rep)))
        return _struct_dict(rep), ids[:n], lat[:n]
