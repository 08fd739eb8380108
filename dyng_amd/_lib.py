"""ctypes view of libdyng_hip.so (include/dyng_hip.h). Fails loudly when the library is missing."""

from __future__ import annotations

import ctypes as C
import os

from . import errors

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdyng_hip.so")

OK = 0
INVALID_ARGUMENT = 1
STALE_RESULT = 2
NOT_SUPPORTED = 3
OUT_OF_MEMORY = 4
DEVICE_ERROR = 5
RUNTIME_ERROR = 6
CAPACITY = 7

ENGINES = {"automatic": 0, "fused": 1, "operators": 2}
ENGINE_NAMES = {v: k for k, v in ENGINES.items()}

_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)


class GraphProps(C.Structure):
    _fields_ = [("directed", C.c_int32), ("sorted_rows", C.c_int32), ("allow_parallel", C.c_int32),
                ("on_existing_insert", C.c_int32), ("on_missing_delete", C.c_int32),
                ("on_self_loop", C.c_int32), ("deletions_first", C.c_int32),
                ("allow_vertex_growth", C.c_int32), ("headroom", C.c_double)]


SUMMARY_FIELDS = ("inserted_edges", "updated_edges", "deleted_edges", "ignored_deletions",
                  "dropped_self_loops", "inserted_vertices", "num_vertices_after",
                  "ignored_insertions", "relocated_rows", "compactions", "host_plans")


APPLY_STAGES = ("scan", "make_ops", "sort", "pairs_out", "unused_4", "unused_5", "pairs_in", "scans",
                "plan", "merge_tails", "merge_full")


class ApplySummary(C.Structure):
    _fields_ = [(f, C.c_int64) for f in SUMMARY_FIELDS] + [("stage_ms", C.c_double * len(APPLY_STAGES))]


class Batch(C.Structure):
    _fields_ = [("ins_src", C.c_void_p), ("ins_dst", C.c_void_p), ("ins_w", C.c_void_p),
                ("num_ins", C.c_int64), ("del_src", C.c_void_p), ("del_dst", C.c_void_p),
                ("num_del", C.c_int64), ("num_weights", C.c_int32), ("on_device", C.c_int32)]


class Tuning(C.Structure):
    _fields_ = [("engine", C.c_int32), ("grid_blocks", C.c_int32), ("blocks_per_cu", C.c_int32),
                ("light_degree", C.c_int32), ("chunk_edges", C.c_int32), ("launch", C.c_int32),
                ("recompute", C.c_int32), ("barrier_timeout_us", C.c_int32)]


COUNTER_FIELDS = ("invalidated", "iterations", "epochs", "pushes", "affected")


class Counters(C.Structure):
    _fields_ = [(f, C.c_int64) for f in COUNTER_FIELDS] + [
        ("packed_parents", C.c_int32), ("engine_used", C.c_int32), ("device_errors", C.c_uint32),
        ("relax_vertices", C.c_int64), ("relax_edges", C.c_int64), ("relax_improved", C.c_int64),
        ("pull_edges", C.c_int64), ("jump_rounds", C.c_int64), ("jump_items", C.c_int64),
        ("pull_vertices", C.c_int64), ("step1_ms", C.c_double), ("pull_ms", C.c_double),
        ("relax_ms", C.c_double), ("blocks", C.c_int64), ("kernel_ms", C.c_double), ("pack_ms", C.c_double),
        ("jump_ms", C.c_double), ("unpack_ms", C.c_double), ("tree_token", C.c_uint64),
        ("words_kept", C.c_int32), ("pad_", C.c_int32)]


class Options(C.Structure):
    _fields_ = [("delta", C.c_int64), ("objective", C.c_int32), ("cuda_engine", C.c_int32),
                ("validate_inputs", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("affected", C.c_int64), ("iterations", C.c_int64), ("frontier_visits", C.c_int64),
                ("fallback_used", C.c_int32), ("converged", C.c_int32), ("engine_used", C.c_int32),
                ("packed_parents", C.c_int32), ("batch", ApplySummary), ("invalidated", C.c_int64),
                ("epochs", C.c_int64), ("pushes", C.c_int64), ("device", Counters),
                ("apply_ms", C.c_double), ("engine_ms", C.c_double)]


class SsspRun(C.Structure):
    _fields_ = [("graph", C.c_void_p), ("objective", C.c_int32), ("source", C.c_int32),
                ("delta", C.c_int64), ("max_weight", C.c_int64), ("distances", C.c_void_p),
                ("parents", C.c_void_p), ("check_cycles", C.c_int32), ("workspace", C.c_void_p),
                ("tuning", Tuning), ("sources", C.c_void_p), ("num_sources", C.c_int64),
                ("tree_token", C.c_uint64)]


class Changes(C.Structure):
    _fields_ = [("del_src", C.c_void_p), ("del_dst", C.c_void_p), ("num_del", C.c_int64),
                ("ins_src", C.c_void_p), ("ins_dst", C.c_void_p), ("ins_increased", C.c_void_p),
                ("num_ins", C.c_int64), ("num_weights", C.c_int32)]


# every exported symbol of include/dyng_hip.h
EXPORTS = (
    "dyng_hip_last_error", "dyng_hip_graph_create", "dyng_hip_graph_create_from_edges",
    "dyng_hip_graph_destroy", "dyng_hip_graph_num_vertices", "dyng_hip_graph_num_edges",
    "dyng_hip_graph_column", "dyng_hip_graph_download", "dyng_hip_graph_download_in",
    "dyng_hip_graph_apply", "dyng_hip_graph_changes", "dyng_hip_workspace_create",
    "dyng_hip_workspace_destroy", "dyng_hip_sssp_update", "dyng_hip_sssp_compute",
    "dyng_hip_graph_version", "dyng_hip_api_compute", "dyng_hip_api_update",
    "dyng_hip_api_from_arrays", "dyng_hip_result_destroy", "dyng_hip_result_clone",
    "dyng_hip_result_arrays", "dyng_hip_result_source", "dyng_hip_result_version",
    "dyng_hip_result_options", "dyng_hip_result_set_options", "dyng_hip_result_download",
    "dyng_hip_result_set_tuning", "dyng_hip_device_count", "dyng_hip_malloc", "dyng_hip_free",
    "dyng_hip_memcpy_htod", "dyng_hip_memcpy_dtoh", "dyng_hip_synchronize", "dyng_hip_api_compute_multi",
    "dyng_hip_api_repair", "dyng_hip_batch_features", "dyng_hip_mlp_create", "dyng_hip_mlp_destroy",
    "dyng_hip_mlp_get_params", "dyng_hip_mlp_set_params", "dyng_hip_mlp_gradients", "dyng_hip_mlp_train",
    "dyng_hip_mlp_predict", "dyng_hip_result_equal", "dyng_hip_api_update_scheduled",
    "dyng_hip_mosp_combined", "dyng_hip_api_update_each", "dyng_hip_mosp_create", "dyng_hip_mosp_update",
    "dyng_hip_mosp_arrays", "dyng_hip_mosp_path_costs", "dyng_hip_mosp_destroy",
    "dyng_hip_sssp_update_noalloc", "dyng_hip_result_counters",
    "dyng_hip_hypergraph_create", "dyng_hip_hypergraph_destroy", "dyng_hip_hypergraph_clone",
    "dyng_hip_hypergraph_info", "dyng_hip_hypergraph_rows", "dyng_hip_hypergraph_apply",
    "dyng_hip_hyper_line_graph", "dyng_hip_hyper_line_graph_delta",
)

NOALLOC_SCRATCH_BYTES = 64 << 20  # DYNG_HIP_NOALLOC_SCRATCH_BYTES

MLP_INPUTS, MLP_HIDDEN, MLP_OUTPUTS, MLP_TRAINED, MLP_PARAMS = 12, 64, 5, 5317, 5341


class BatchStats(C.Structure):
    _fields_ = [("num_ins", C.c_int64), ("num_del", C.c_int64), ("tree_deletions", C.c_int64),
                ("touched", C.c_int64), ("mean_log_degree", C.c_double), ("max_degree", C.c_int64),
                ("num_vertices", C.c_int64), ("num_edges", C.c_int64)]

class MospStats(C.Structure):
    _fields_ = [("combined_edges", C.c_int64), ("weight_sum", C.c_int64), ("preference_scale", C.c_int64),
                ("base", C.c_int64), ("affected", C.c_int64), ("deletions", C.c_int64), ("insertions", C.c_int64),
                ("solve", Counters)]


class HyperProps(C.Structure):
    _fields_ = [("id_reuse", C.c_int32), ("initial_headroom", C.c_double), ("initial_slack", C.c_int64),
                ("min_initial_payload", C.c_int64), ("grow_factor", C.c_double), ("compact_before_grow", C.c_int32),
                ("max_payload", C.c_int64), ("row_alignment", C.c_int32), ("compaction_headroom", C.c_int32)]


class HyperBatch(C.Structure):
    _fields_ = [("delete_hyperedges", C.c_void_p), ("num_deletions", C.c_int64),
                ("insert_offsets", C.c_void_p), ("insert_members", C.c_void_p), ("insert_weights", C.c_void_p),
                ("num_insertions", C.c_int64),
                ("delete_incidence_hyperedge", C.c_void_p), ("delete_incidence_vertex", C.c_void_p),
                ("num_removals", C.c_int64),
                ("insert_incidence_hyperedge", C.c_void_p), ("insert_incidence_vertex", C.c_void_p),
                ("num_additions", C.c_int64)]


HYPER_SUMMARY_FIELDS = ("inserted_hyperedges", "deleted_hyperedges", "inserted_incidences", "deleted_incidences",
                        "skipped_ops", "ignored_deletions", "dropped_members", "num_hyperedges_after",
                        "id_bound_after")


class HyperSummary(C.Structure):
    _fields_ = [(f, C.c_int64) for f in HYPER_SUMMARY_FIELDS]


class HyperDelta(C.Structure):
    _fields_ = [("num_insertions", C.c_int64), ("insert_src", _i32p), ("insert_dst", _i32p),
                ("insert_weights", _i32p), ("num_deletions", C.c_int64), ("delete_src", _i32p),
                ("delete_dst", _i32p), ("line_graph_insertions", C.c_int64), ("line_graph_deletions", C.c_int64)]


_lib = None


def lib() -> C.CDLL:
    """The loaded library (raises NotSupportedError when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise errors.NotSupportedError(
            f"dyng_amd: {LIB_PATH} is not built; run `python -m dyng_amd.build` (the HIP path has "
            "no fallback)")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.dyng_hip_last_error.restype = C.c_char_p
    L.dyng_hip_graph_create.argtypes = [C.c_int, C.POINTER(GraphProps), C.c_int64, C.c_int64, C.c_int32,
                                        vp, vp, vp, C.c_int32, vp, C.POINTER(vp)]
    L.dyng_hip_graph_create_from_edges.argtypes = [C.c_int, C.POINTER(GraphProps), C.c_int64, C.c_int64,
                                                   C.c_int32, vp, vp, vp, C.c_int32, vp, C.POINTER(vp)]
    L.dyng_hip_graph_destroy.argtypes = [vp]
    L.dyng_hip_graph_num_vertices.argtypes = [vp, _i64p]
    L.dyng_hip_graph_num_edges.argtypes = [vp, _i64p]
    L.dyng_hip_graph_version.argtypes = [vp, C.POINTER(C.c_uint64)]
    L.dyng_hip_graph_column.argtypes = [vp, C.c_int32, _i64p, _i64p]
    L.dyng_hip_graph_download.argtypes = [vp, vp, vp, vp]
    L.dyng_hip_graph_download_in.argtypes = [vp, vp, vp, vp]
    L.dyng_hip_graph_apply.argtypes = [vp, C.POINTER(Batch), vp, C.POINTER(ApplySummary)]
    L.dyng_hip_graph_changes.argtypes = [vp, C.POINTER(Changes)]
    L.dyng_hip_workspace_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.dyng_hip_workspace_destroy.argtypes = [vp]
    L.dyng_hip_sssp_update.argtypes = [C.POINTER(SsspRun), vp, C.POINTER(Counters)]
    L.dyng_hip_sssp_compute.argtypes = [C.POINTER(SsspRun), vp, C.POINTER(Counters)]
    L.dyng_hip_sssp_update_noalloc.argtypes = [C.POINTER(SsspRun), vp, C.POINTER(Counters)]
    L.dyng_hip_api_compute.argtypes = [vp, C.c_int32, C.POINTER(Options), vp, C.POINTER(vp)]
    L.dyng_hip_api_update.argtypes = [vp, C.POINTER(Batch), vp, vp, C.POINTER(Stats)]
    L.dyng_hip_api_compute_multi.argtypes = [vp, vp, C.c_int64, C.POINTER(Options), vp, C.POINTER(vp)]
    L.dyng_hip_api_repair.argtypes = [vp, vp, vp, C.POINTER(Stats)]
    L.dyng_hip_api_update_each.argtypes = [vp, C.POINTER(Batch), vp, C.c_int64, vp, vp]
    L.dyng_hip_mosp_create.argtypes = [vp, C.c_int32, C.c_int32, vp, vp, vp, C.POINTER(vp), C.POINTER(MospStats)]
    L.dyng_hip_mosp_update.argtypes = [vp, vp, vp, vp, C.POINTER(MospStats)]
    L.dyng_hip_mosp_arrays.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(C.c_int64)]
    L.dyng_hip_mosp_path_costs.argtypes = [vp, vp, C.c_int32, vp, vp]
    L.dyng_hip_mosp_destroy.argtypes = [vp]
    L.dyng_hip_mosp_combined.argtypes = [vp, C.c_int32, C.c_int32, vp, vp, C.c_int32, vp, vp, C.c_int32, vp, vp,
                                         C.POINTER(MospStats)]
    L.dyng_hip_batch_features.argtypes = [vp, C.POINTER(Batch), vp, C.c_int64, vp, C.POINTER(BatchStats)]
    L.dyng_hip_mlp_create.argtypes = [C.c_int, C.c_uint64, C.POINTER(vp)]
    L.dyng_hip_mlp_destroy.argtypes = [vp]
    L.dyng_hip_mlp_get_params.argtypes = [vp, vp, C.c_int64]
    L.dyng_hip_mlp_set_params.argtypes = [vp, vp, C.c_int64]
    L.dyng_hip_mlp_gradients.argtypes = [vp, vp, vp, C.c_int64, vp, vp]
    L.dyng_hip_mlp_train.argtypes = [vp, vp, vp, C.c_int64, C.c_int32, C.c_int32, C.c_float, C.c_uint64, vp]
    L.dyng_hip_mlp_predict.argtypes = [vp, vp, C.c_int64, vp]
    L.dyng_hip_result_equal.argtypes = [vp, vp, vp, C.POINTER(C.c_int64)]
    L.dyng_hip_api_update_scheduled.argtypes = [vp, C.POINTER(Batch), vp, vp, C.c_int64, C.c_int64, vp,
                                                C.POINTER(Stats), C.POINTER(C.c_int32)]
    L.dyng_hip_api_from_arrays.argtypes = [vp, C.c_int32, vp, vp, C.c_int32, C.POINTER(Options), vp,
                                           C.POINTER(vp)]
    L.dyng_hip_result_destroy.argtypes = [vp]
    L.dyng_hip_result_clone.argtypes = [vp, vp, C.POINTER(vp)]
    L.dyng_hip_result_arrays.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), _i64p]
    L.dyng_hip_result_source.argtypes = [vp, _i32p]
    L.dyng_hip_result_counters.argtypes = [vp, C.POINTER(Counters)]
    L.dyng_hip_result_version.argtypes = [vp, C.POINTER(C.c_uint64)]
    L.dyng_hip_result_options.argtypes = [vp, C.POINTER(Options)]
    L.dyng_hip_result_set_options.argtypes = [vp, C.POINTER(Options)]
    L.dyng_hip_result_download.argtypes = [vp, vp, vp]
    L.dyng_hip_result_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
    L.dyng_hip_device_count.argtypes = [C.POINTER(C.c_int)]
    L.dyng_hip_malloc.argtypes = [C.c_int, C.c_int64, C.POINTER(vp)]
    L.dyng_hip_free.argtypes = [vp]
    L.dyng_hip_memcpy_htod.argtypes = [vp, vp, C.c_int64]
    L.dyng_hip_memcpy_dtoh.argtypes = [vp, vp, C.c_int64]
    L.dyng_hip_synchronize.argtypes = [C.c_int]
    L.dyng_hip_hypergraph_create.argtypes = [vp, vp, C.c_int64, vp, C.c_int32, C.POINTER(HyperProps), C.POINTER(vp)]
    L.dyng_hip_hypergraph_destroy.argtypes = [vp]
    L.dyng_hip_hypergraph_clone.argtypes = [vp, C.POINTER(vp)]
    L.dyng_hip_hypergraph_info.argtypes = [vp, _i32p, _i64p, _i64p, _i64p, _i32p]
    L.dyng_hip_hypergraph_rows.argtypes = [vp, C.POINTER(_i64p), C.POINTER(_i32p), C.POINTER(_i64p)]
    L.dyng_hip_hypergraph_apply.argtypes = [vp, C.POINTER(HyperBatch), C.c_int64, C.POINTER(HyperSummary), vp]
    L.dyng_hip_hyper_line_graph.argtypes = [vp, C.c_int32, _i64p, C.POINTER(_i32p), C.POINTER(_i32p),
                                            C.POINTER(_i32p)]
    L.dyng_hip_hyper_line_graph_delta.argtypes = [vp, C.c_int32, C.POINTER(HyperDelta)]
    _lib = L
    return L


def check(status: int) -> None:
    """Raise the dynG exception class of a status (cpp/include/dyng/core/error.hpp:28-164)."""
    if status == OK:
        return
    msg = (lib().dyng_hip_last_error() or b"").decode(errors="replace")
    cls = {
        INVALID_ARGUMENT: errors.InvalidArgumentError,
        STALE_RESULT: errors.StaleResultError,
        NOT_SUPPORTED: errors.NotSupportedError,
        OUT_OF_MEMORY: errors.OutOfMemoryError,
        RUNTIME_ERROR: errors.DeviceError,
        CAPACITY: errors.CapacityError,
    }.get(status)
    if status == DEVICE_ERROR:
        # invalid input trees (parent cycle, distances out of range) are invalid_argument_error in
        # the reference; the other device errors are internal
        if "parent cycle" in msg or "initial tree" in msg:
            raise errors.InvalidArgumentError(msg)
        raise errors.DeviceError(msg)
    raise (cls or errors.DyngError)(msg)
