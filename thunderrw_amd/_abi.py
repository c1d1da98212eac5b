"""ctypes view of the C ABI in include/thunderrw.h (libthunderrw.so)."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libthunderrw.so")

OK = 0
E_INVALID, E_HIP, E_NOMEM, E_UNSUPPORTED, E_STATE, E_IO, E_BUSY, E_TIMEOUT = -1, -2, -3, -4, -5, -6, -7, -8

EXEC_UNIFORM, EXEC_STATIC, EXEC_DYNAMIC = 0, 1, 2
SM_UNIFORM, SM_ITS, SM_ALIAS, SM_REJECTION, SM_MAX_WEIGHT_REJECTION = 0, 1, 2, 3, 4
APP_DEEPWALK, APP_PPR, APP_NODE2VEC, APP_METAPATH, APP_GENERIC = 0, 1, 2, 3, 4
RNG_REPLAY, RNG_PHILOX = 0, 1
FORM_AUTO, FORM_FUSED, FORM_QUEUE, FORM_STEP, FORM_GRAPH = 0, 1, 2, 3, 4
WEIGHTS_EDGE, WEIGHTS_UNIT, WEIGHTS_CUSTOM = 0, 1, 2
WALK_WEIGHTS_APP, WALK_WEIGHTS_EDGE, WALK_WEIGHTS_UNIT = 0, 1, 2
TUNE_WALKERS_PER_THREAD, TUNE_ITS_TREE_MIN_DEGREE, TUNE_QUEUE_CHUNK, TUNE_ITS_GUIDE_OFF = 0, 1, 2, 3
TUNE_FUSED_BLOCK, TUNE_FUSED_TILE, TUNE_QUEUE_BLOCKS_PER_CU, TUNE_N2V_BLOCK, TUNE_GRAPH_BURST = 4, 5, 6, 7, 8
RESULT_PPR, RESULT_PATHS = 0, 1
PATHS_PADDED, PATHS_UNPADDED = 0, 1
QUERY_PENDING, QUERY_RUNNING, QUERY_DONE, QUERY_FAILED = 0, 1, 2, 3
WALKER_DONE, WALKER_GATHER_TERMINATED, WALKER_DEAD_END, WALKER_REPLAY_EXHAUSTED, WALKER_TRIAL_CAP, \
    WALKER_BAD_SOURCE, WALKER_STEP_CAP = 0, 1, 2, 3, 4, 5, 6

APPS = {"deepwalk": APP_DEEPWALK, "ppr": APP_PPR, "node2vec": APP_NODE2VEC, "metapath": APP_METAPATH,
        "genericwalk": APP_GENERIC}


class TrwError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("thunderrw error %d: %s" % (code, msg))
        self.code = code


class AliasSlot(C.Structure):
    _fields_ = [("alias_value", C.c_double), ("first", C.c_int32), ("second", C.c_int32)]


class DeviceView(C.Structure):
    _fields_ = [("num_vertices", C.c_int32), ("max_degree", C.c_int32), ("num_edges", C.c_int64),
                ("offset", C.c_void_p), ("degree", C.c_void_p), ("adj", C.c_void_p), ("edge_weight", C.c_void_p),
                ("edge_label", C.c_void_p), ("alias_table", C.c_void_p), ("prefix_sum", C.c_void_p),
                ("rejection_max", C.c_void_p), ("rejection_weight", C.c_void_p), ("alias_weights", C.c_int32),
                ("its_weights", C.c_int32), ("rejection_weights", C.c_int32), ("reserved", C.c_int32),
                ("membership_lines", C.c_void_p), ("membership_root", C.c_void_p)]


class WalkParams(C.Structure):
    _fields_ = [("app", C.c_int32), ("execution", C.c_int32), ("sampling", C.c_int32), ("length", C.c_int32),
                ("stop_probability", C.c_double), ("p", C.c_double), ("q", C.c_double),
                ("schema", C.c_int32 * 16), ("schema_length", C.c_int32), ("rng", C.c_int32),
                ("seed", C.c_uint64), ("form", C.c_int32), ("max_trials", C.c_int32),
                ("walkers_per_thread", C.c_int32), ("weights", C.c_int32)]


class WalkIO(C.Structure):
    _fields_ = [("sources", C.c_void_p), ("n_walkers", C.c_int32), ("path_stride", C.c_int32),
                ("walker_id_base", C.c_int64), ("draws", C.c_void_p), ("draws_per_walker", C.c_int64),
                ("term_draws", C.c_void_p), ("term_draws_per_walker", C.c_int64), ("paths", C.c_void_p),
                ("lengths", C.c_void_p), ("ends", C.c_void_p), ("status", C.c_void_p), ("draws_used", C.c_void_p),
                ("counters", C.c_void_p), ("walker_ids", C.c_void_p), ("path_flags", C.c_int32),
                ("path_stride_b", C.c_int32)]


TUNE_KNOBS = 9
MOVES = ("uniform", "alias", "its", "rejection", "node2vec", "metapath")  # TRW_MOVE_* order


class TuneReport(C.Structure):
    _fields_ = [("knobs", C.c_int32), ("move", C.c_int32), ("knob", C.c_int32 * TUNE_KNOBS),
                ("candidates", C.c_int32 * TUNE_KNOBS), ("chosen", C.c_int64 * TUNE_KNOBS),
                ("default_steps_per_s", C.c_double), ("tuned_steps_per_s", C.c_double)]


class ServerConfig(C.Structure):
    _fields_ = [("streams", C.c_int32), ("max_batch_walkers", C.c_int32), ("path_bytes_per_batch", C.c_int64),
                ("max_pending_queries", C.c_int32), ("completion_threads", C.c_int32),
                ("max_queries_per_batch", C.c_int32), ("reserved", C.c_int32)]


class Query(C.Structure):
    _fields_ = [("sources", C.c_void_p), ("n_sources", C.c_int32), ("n_walkers", C.c_int32),
                ("walker_id_base", C.c_uint64), ("params", WalkParams), ("result", C.c_int32),
                ("reserved", C.c_int32)]


class QueryInfo(C.Structure):
    _fields_ = [("state", C.c_int32), ("n_walkers", C.c_int32), ("steps", C.c_uint64), ("errors", C.c_uint64),
                ("ppr_entries", C.c_int64), ("queue_ms", C.c_double), ("latency_ms", C.c_double),
                ("device_ms", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("queries_done", C.c_uint64), ("walkers", C.c_uint64), ("steps", C.c_uint64),
                ("batches", C.c_uint64), ("rejected", C.c_uint64), ("mean_batch_walkers", C.c_double),
                ("mean_batch_queries", C.c_double)]


class OpenLoopReport(C.Structure):
    _fields_ = [("wall_ms", C.c_double), ("offered_qps", C.c_double), ("achieved_qps", C.c_double),
                ("steps", C.c_uint64), ("steps_per_s", C.c_double), ("p50_ms", C.c_double), ("p90_ms", C.c_double),
                ("p99_ms", C.c_double), ("p999_ms", C.c_double), ("mean_ms", C.c_double), ("max_ms", C.c_double),
                ("batches", C.c_uint64), ("mean_batch_queries", C.c_double), ("rejected", C.c_uint64)]


_SIGS = {
    "trw_last_error": (C.c_char_p, []),
    "trw_version": (C.c_int, [C.c_void_p, C.c_void_p]),
    "trw_device_count": (C.c_int, [C.c_void_p]),
    "trw_context_create": (C.c_int, [C.c_int32, C.c_void_p]),
    "trw_context_destroy": (C.c_int, [C.c_void_p]),
    "trw_context_stream": (C.c_int, [C.c_void_p, C.c_void_p]),
    "trw_device_alloc": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p]),
    "trw_device_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "trw_copy_to_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    "trw_copy_to_host": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]),
    "trw_device_zero": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "trw_device_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "trw_probe_random_gather": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int32, C.c_int32, C.c_int32, C.c_void_p]),
    "trw_synchronize": (C.c_int, [C.c_void_p, C.c_void_p]),
    "trw_timer_begin": (C.c_int, [C.c_void_p, C.c_void_p]),
    "trw_timer_end": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "trw_graph_upload": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_void_p]),
    "trw_graph_free": (C.c_int, [C.c_void_p]),
    "trw_graph_info": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "trw_build_tables": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "trw_copy_alias_table": (C.c_int, [C.c_void_p, C.c_void_p]),
    "trw_copy_prefix_sum": (C.c_int, [C.c_void_p, C.c_void_p]),
    "trw_copy_rejection_max": (C.c_int, [C.c_void_p, C.c_void_p]),
    "trw_last_build_ms": (C.c_int, [C.c_void_p, C.c_void_p]),
    "trw_build_tables_from": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    "trw_graph_device_view": (C.c_int, [C.c_void_p, C.c_void_p]),
    "trw_build_membership": (C.c_int, [C.c_void_p]),
    "trw_graph_store_tables": (C.c_int, [C.c_void_p, C.c_char_p]),
    "trw_graph_load_tables": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int32, C.c_int32]),
    "trw_graph_set_tuning": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64]),
    "trw_graph_get_tuning": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    "trw_graph_set_move_tuning": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int64]),
    "trw_graph_get_move_tuning": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]),
    "trw_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_uint64, C.c_int32, C.c_void_p]),
    "trw_autotune_mix": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_uint64, C.c_int32,
                                   C.c_void_p]),
    "trw_walk_async": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "trw_walk_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "trw_walk_supported": (C.c_int, [C.c_void_p]),
    "trw_walk_mix_async": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "trw_walk_host": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.c_void_p,
                                C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "trw_replay_fill": (C.c_int, [C.c_uint32, C.c_int64, C.c_int32, C.c_int64, C.c_void_p]),
    "trw_sources_all": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "trw_sources_single": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]),
    "trw_sources_random": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_uint32, C.c_void_p]),
    "trw_host_graph_chung_lu": (C.c_int, [C.c_int32, C.c_int64, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                          C.c_int32, C.c_int32, C.c_void_p]),
    "trw_host_graph_load": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.c_void_p]),
    "trw_host_graph_load_edge_list": (C.c_int, [C.c_char_p, C.c_int32, C.c_void_p]),
    "trw_host_graph_store": (C.c_int, [C.c_void_p, C.c_char_p]),
    "trw_host_graph_arrays": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.c_void_p]),
    "trw_host_graph_free": (C.c_int, [C.c_void_p]),
    "trw_server_create": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "trw_server_destroy": (C.c_int, [C.c_void_p]),
    "trw_server_submit": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "trw_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.c_double, C.c_void_p]),
    "trw_server_query_info": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p]),
    "trw_server_ppr_result": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_int64]),
    "trw_server_walk_result": (C.c_int, [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "trw_server_release": (C.c_int, [C.c_void_p, C.c_uint64]),
    "trw_server_stats_get": (C.c_int, [C.c_void_p, C.c_void_p]),
    "trw_server_run_open_loop": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_double, C.c_uint64, C.c_void_p,
                                           C.c_void_p, C.c_void_p]),
}

_LIB = None


def lib():
    """The loaded library; raises if it has not been built (there is no fallback)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("thunderrw_amd/libthunderrw.so is missing: build it with "
                               "`python -m thunderrw_amd.build` (there is no CPU fallback)")
        lib_ = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib_, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib_
    return _LIB


def check(rc):
    if rc != OK:
        raise TrwError(rc, lib().trw_last_error().decode(errors="replace"))
    return rc


def exported_names():
    return list(_SIGS)
