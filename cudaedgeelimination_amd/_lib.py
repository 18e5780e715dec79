"""ctypes binding of ``libee_hip.so`` (the C ABI in ``include/ee_hip.h`` and
``include/ee_gen.h``).

The library is loaded from this package directory only.  There is no CPU
fallback: when the library is missing every entry point raises
``BackendMissing``.
"""
from __future__ import annotations

import ctypes as C
import pathlib
import threading

import numpy as np

LIBRARY_PATH = pathlib.Path(__file__).resolve().with_name("libee_hip.so")

EE_OK = 0
# ee_kopt_batch_run flags (EE_KOPT_RUN_* in include/ee_hip.h)
KOPT_RUN_GRAPH = 1
KOPT_RUN_PIPELINED = 2
STATUS_NAMES = {
    1: "EE_ERR_INVALID_ARGUMENT",
    2: "EE_ERR_OVERFLOW",
    3: "EE_ERR_RUNTIME",
    4: "EE_ERR_LOGIC",
    5: "EE_ERR_UNAVAILABLE",
}


class BackendMissing(RuntimeError):
    """The HIP backend library has not been built."""


class EeError(RuntimeError):
    """A C entry point returned a non-OK status."""

    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status
        self.message = message


class InvalidArgument(EeError, ValueError):
    pass


class OverflowFailure(EeError, OverflowError):
    pass


class Unavailable(EeError):
    pass


# ---------------------------------------------------------------- structs
KOPT_TASK_DTYPE = np.dtype([("port_nodes", np.int32, (10,)), ("deleted_cost", np.int64)],
                           align=True)
assert KOPT_TASK_DTYPE.itemsize == 48
MATCHING_DTYPE = np.dtype([("endpoint_count", np.uint8), ("mate", np.uint8, (14,))])
assert MATCHING_DTYPE.itemsize == 15


class GraphView(C.Structure):
    _fields_ = [
        ("dimension", C.c_int32),
        ("distance_type", C.c_uint8),
        ("integer_coordinates", C.c_uint8),
        ("reserved", C.c_uint8 * 2),
        ("x", C.c_void_p),
        ("y", C.c_void_p),
        ("edge_count", C.c_int32),
        ("edge_u", C.c_void_p),
        ("edge_v", C.c_void_p),
        ("edge_weight", C.c_void_p),
        ("edge_active", C.c_void_p),
        ("row_offsets", C.c_void_p),
        ("neighbors", C.c_void_p),
        ("csr_edge_ids", C.c_void_p),
    ]


class JvUsage(C.Structure):
    _fields_ = [
        ("static_hit", C.c_int32),
        ("workspace_hit", C.c_int32),
        ("resident_bytes", C.c_uint64),
        ("h2d_ms", C.c_double),
        ("kernel_ms", C.c_double),
        ("d2h_ms", C.c_double),
    ]


class KoptUsage(C.Structure):
    _fields_ = [
        ("snapshot_hit", C.c_int32),
        ("template_hit", C.c_int32),
        ("workspace_hit", C.c_int32),
        ("reserved", C.c_int32),
        ("resident_bytes", C.c_uint64),
        ("h2d_ms", C.c_double),
        ("kernel_ms", C.c_double),
        ("d2h_ms", C.c_double),
    ]


class HtUsage(C.Structure):
    _fields_ = [
        ("graph_hit", C.c_int32),
        ("workspace_hit", C.c_int32),
        ("resident_bytes", C.c_uint64),
        ("h2d_ms", C.c_double),
        ("kernel_ms", C.c_double),
        ("d2h_ms", C.c_double),
    ]


class TourView(C.Structure):
    _fields_ = [
        ("tour", C.c_void_p),
        ("edge_costs", C.c_void_p),
        ("selectable", C.c_void_p),
        ("tour_length", C.c_uint32),
        ("selectable_count", C.c_uint32),
    ]


class LeafBlock(C.Structure):
    _fields_ = [("tag", C.c_uint64), ("first", C.c_uint64), ("k", C.c_uint32), ("rows", C.c_uint32),
                ("template_count", C.c_uint32), ("mask_words", C.c_uint32), ("tasks", C.c_void_p),
                ("positions", C.c_void_p), ("costs", C.c_void_p), ("improving", C.c_void_p),
                ("improving_cells", C.c_uint64)]


class LeafStats(C.Structure):
    _fields_ = [("blocks", C.c_uint64), ("rows", C.c_uint64), ("cells", C.c_uint64),
                ("improving_cells", C.c_uint64), ("kernel_ms", C.c_double), ("copy_ms", C.c_double),
                ("wait_ms", C.c_double), ("certify_ms", C.c_double), ("certify_threads", C.c_uint32),
                ("reserved", C.c_uint32)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class SegmentDesc(C.Structure):
    _fields_ = [("instance", C.c_uint32), ("k", C.c_uint32), ("task_count", C.c_uint64)]


class KoptTuning(C.Structure):
    _fields_ = [("rows_per_tile", C.c_uint32 * 3), ("grid", C.c_uint32), ("block", C.c_uint32),
                ("non_temporal", C.c_uint32)]

    def as_dict(self) -> dict:
        return {"rows_per_tile": list(self.rows_per_tile), "grid": self.grid,
                "block": self.block, "non_temporal": self.non_temporal}


class JvRunStats(C.Structure):
    _fields_ = [
        ("epochs", C.c_uint32), ("reserved", C.c_uint32), ("edges_scanned", C.c_uint64),
        ("proposed", C.c_uint64), ("committed", C.c_uint64), ("initial_hash", C.c_uint64),
        ("final_hash", C.c_uint64), ("device_ms", C.c_double), ("h2d_ms", C.c_double),
        ("d2h_ms", C.c_double), ("propose_ms", C.c_double), ("verify_ms", C.c_double),
        ("commit_ms", C.c_double), ("total_ms", C.c_double), ("compact_ms", C.c_double),
        ("hash_wait_ms", C.c_double),
    ]


class JvFleetStats(C.Structure):
    _fields_ = [
        ("instances", C.c_uint32), ("epochs", C.c_uint32), ("edge_slots", C.c_uint64),
        ("edges_scanned", C.c_uint64), ("proposed", C.c_uint64), ("committed", C.c_uint64),
        ("device_ms", C.c_double), ("assemble_ms", C.c_double), ("propose_ms", C.c_double),
        ("host_ms", C.c_double), ("total_ms", C.c_double), ("graph_launches", C.c_uint32),
        ("reserved", C.c_uint32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_batch_cells", C.c_uint64),
                ("max_batch_requests", C.c_uint32), ("pipeline_depth", C.c_uint32)]


class RequestResult(C.Structure):
    _fields_ = [("request_id", C.c_uint64), ("status", C.c_int32), ("batch_size", C.c_uint32),
                ("batch_id", C.c_uint64), ("queue_ms", C.c_double), ("service_ms", C.c_double),
                ("latency_ms", C.c_double)]


class ServerRequest(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("graph", C.c_uint32), ("k", C.c_uint32),
                ("reserved", C.c_uint32), ("input", C.c_void_p), ("count", C.c_uint64),
                ("output", C.c_void_p)]


class ServerStats(C.Structure):
    _fields_ = [("requests", C.c_uint64), ("batches", C.c_uint64), ("cells", C.c_uint64),
                ("jv_edges", C.c_uint64), ("peak_inflight_bytes", C.c_uint64),
                ("device_ms", C.c_double)]


class QueueConfig(C.Structure):
    _fields_ = [("task_rows", C.c_uint64), ("cost_cells", C.c_uint64), ("max_nodes", C.c_uint64),
                ("ring_entries", C.c_uint32), ("workers", C.c_uint32), ("idle_exit_ms", C.c_double),
                ("max_jv_nodes", C.c_uint64), ("max_jv_edges", C.c_uint64),
                ("max_jv_adjacency", C.c_uint64)]


class QueueStats(C.Structure):
    _fields_ = [("requests", C.c_uint64), ("cells", C.c_uint64), ("launches", C.c_uint64),
                ("peak_inflight_cells", C.c_uint64), ("jv_edges", C.c_uint64)]


P = C.c_void_p
I32, U32, I64, U64, U8, SZ, DBL = (C.c_int32, C.c_uint32, C.c_int64, C.c_uint64, C.c_uint8,
                                   C.c_size_t, C.c_double)

# name -> (restype, argtypes)
PROTOTYPES = {
    "ee_last_error": (C.c_char_p, []),
    "ee_version": (C.c_char_p, []),
    "ee_device_count": (C.c_int, []),
    "ee_backend_available": (C.c_int, [C.c_int, C.c_char_p, SZ]),
    "ee_ctx_create": (C.c_int, [C.c_int, C.POINTER(P)]),
    "ee_ctx_destroy": (None, [P]),
    "ee_ctx_device": (C.c_int, [P]),
    "ee_ctx_synchronize": (C.c_int, [P]),
    "ee_ctx_set_jv_edges_per_wave": (C.c_int, [P, U32]),
    "ee_leaf_inject_cell": (C.c_int, [P, U64, I64]),
    "ee_jv_session_inject_witness": (C.c_int, [P, I32, I32]),
    "ee_graph_content_hash_hip": (C.c_int, [P, I32, U8, P, P, I32, P, P, P, P, C.POINTER(C.c_uint64)]),
    "ee_kopt_batch_create": (C.c_int, [P, C.POINTER(P)]),
    "ee_kopt_batch_destroy": (None, [P]),
    "ee_kopt_batch_add_instance": (C.c_int, [P, P, P, I32, U8, C.POINTER(U32)]),
    "ee_kopt_batch_set_work": (C.c_int, [P, C.POINTER(SegmentDesc), U32, P, P]),
    "ee_kopt_batch_set_tuning": (C.c_int, [P, C.POINTER(KoptTuning)]),
    "ee_kopt_batch_get_tuning": (C.c_int, [P, C.POINTER(KoptTuning)]),
    "ee_kopt_batch_autotune": (C.c_int, [P, U32, C.c_char_p, C.POINTER(KoptTuning)]),
    "ee_kopt_batch_autotune_runs": (C.c_int, [P, U32, C.c_int, C.c_char_p, C.POINTER(KoptTuning)]),
    "ee_kopt_batch_run": (C.c_int, [P, U32, C.c_int, C.POINTER(DBL)]),
    "ee_kopt_batch_set_pipeline_depth": (C.c_int, [P, U32]),
    "ee_kopt_batch_run_stamped": (C.c_int, [P, U32, C.c_int, P, P]),
    "ee_kopt_batch_download": (C.c_int, [P, U64, U64, P]),
    "ee_kopt_batch_download_pool": (C.c_int, [P, U32, U64, U64, P]),
    "ee_kopt_batch_set_output": (C.c_int, [P, U32]),
    "ee_kopt_batch_mask_words": (C.c_int, [P, C.POINTER(U32)]),
    "ee_kopt_batch_download_masks": (C.c_int, [P, U64, U64, P]),
    "ee_kopt_batch_resident_bytes": (U64, [P]),
    "ee_jv_session_create": (C.c_int, [P, C.POINTER(GraphView), C.POINTER(P)]),
    "ee_jv_session_destroy": (None, [P]),
    "ee_jv_session_run": (C.c_int, [P, U32, C.POINTER(JvRunStats)]),
    "ee_jv_session_proof": (C.c_int, [P, C.c_char_p, SZ, C.POINTER(SZ)]),
    "ee_jv_session_device_snapshot": (C.c_int, [P, P, P, P, P, SZ]),
    "ee_jv_session_active": (C.c_int, [P, P]),
    "ee_jv_fleet_create": (C.c_int, [P, C.POINTER(GraphView), U32, C.POINTER(P)]),
    "ee_jv_fleet_destroy": (None, [P]),
    "ee_jv_fleet_run": (C.c_int, [P, U32, C.c_int, C.POINTER(JvFleetStats)]),
    "ee_jv_fleet_instance_stats": (C.c_int, [P, U32, C.POINTER(JvRunStats)]),
    "ee_jv_fleet_proof": (C.c_int, [P, U32, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "ee_jv_fleet_active": (C.c_int, [P, U32, P]),
    "ee_server_create": (C.c_int, [P, C.POINTER(ServerConfig), C.POINTER(P)]),
    "ee_server_destroy": (None, [P]),
    "ee_server_add_graph": (C.c_int, [P, P, P, I32, U8, C.POINTER(U32)]),
    "ee_server_submit": (C.c_int, [P, U32, U32, P, U64, P, C.POINTER(U64)]),
    "ee_server_add_jv_graph": (C.c_int, [P, C.POINTER(GraphView), C.POINTER(U32)]),
    "ee_server_submit_jv": (C.c_int, [P, U32, P, U64, P, C.POINTER(U64)]),
    "ee_server_wait": (C.c_int, [P, U64, DBL, C.POINTER(RequestResult)]),
    "ee_server_submit_many": (C.c_int, [P, U32, P, P]),
    "ee_server_wait_many": (C.c_int, [P, U32, P, DBL, P]),
    "ee_server_stats_get": (C.c_int, [P, C.POINTER(ServerStats)]),
    "ee_queue_create": (C.c_int, [P, C.POINTER(QueueConfig), C.POINTER(P)]),
    "ee_queue_destroy": (None, [P]),
    "ee_queue_add_graph": (C.c_int, [P, P, P, I32, U8, C.POINTER(U32)]),
    "ee_queue_submit": (C.c_int, [P, U32, U32, P, U64, P, C.POINTER(U64)]),
    "ee_queue_add_jv_graph": (C.c_int, [P, C.POINTER(GraphView), C.POINTER(U32)]),
    "ee_queue_submit_jv": (C.c_int, [P, U32, P, U64, P, C.POINTER(U64)]),
    "ee_queue_submit_many": (C.c_int, [P, U32, P, P]),
    "ee_queue_wait": (C.c_int, [P, U64, DBL, C.POINTER(RequestResult)]),
    "ee_queue_wait_many": (C.c_int, [P, U32, P, DBL, P]),
    "ee_queue_stats_get": (C.c_int, [P, C.POINTER(QueueStats)]),
    "ee_queue_inject_refusal": (C.c_int, [P, U32, P, P, C.POINTER(U64)]),
    "ee_jv_find_candidates": (C.c_int, [P, C.POINTER(GraphView), P, C.POINTER(JvUsage)]),
    "ee_jv_clear_cache": (C.c_int, [P]),
    "ee_kopt_build_templates": (C.c_int, [U32, P, U32, C.POINTER(U32), C.POINTER(U64)]),
    "ee_ht_hamilton_replies": (C.c_int, [P, P, I32, I32, P, SZ, P, P, U64, C.POINTER(U64),
                                         C.POINTER(HtUsage)]),
    "ee_ht_end_replies": (C.c_int, [P, P, P, SZ, P, P, U64, C.POINTER(U64), C.POINTER(HtUsage)]),
    "ee_ht_copy_replies": (C.c_int, [P, U64, U64, P]),
    "ee_ht_clear_cache": (C.c_int, [P]),
    "ee_kopt_cursor_count": (C.c_int, [C.POINTER(TourView), U32, C.POINTER(U64)]),
    "ee_kopt_cursor_eval": (C.c_int, [P, P, P, I32, U8, C.POINTER(TourView), U32, P, U32, U64, U64,
                                      U32, P, P, P, P, C.POINTER(U32), C.POINTER(KoptUsage)]),
    "ee_kopt_eval_costs": (C.c_int, [P, P, P, I32, U8, U32, P, U32, U64, P, SZ, P,
                                     C.POINTER(KoptUsage)]),
    "ee_kopt_eval_costs_noalloc": (C.c_int, [P, P, P, I32, U8, U32, P, U32, U64, P, SZ, P,
                                             C.POINTER(KoptUsage)]),
    "ee_kopt_clear_cache": (C.c_int, [P]),
    "ee_kopt_certify": (C.c_int, [P, P, I32, U8, U32, P, U32, P, SZ, P, P, C.POINTER(U64),
                                  C.POINTER(U64)]),
    "ee_leaf_create": (C.c_int, [P, P, P, I32, U8, U32, U32, C.POINTER(P)]),
    "ee_leaf_destroy": (None, [P]),
    "ee_leaf_set_tour": (C.c_int, [P, C.POINTER(TourView)]),
    "ee_leaf_submit_tasks": (C.c_int, [P, U32, P, U32, U64]),
    "ee_leaf_submit_cursor": (C.c_int, [P, U32, U64, U32, U64, C.POINTER(U32)]),
    "ee_leaf_next": (C.c_int, [P, C.POINTER(LeafBlock)]),
    "ee_leaf_stats_get": (C.c_int, [P, C.POINTER(LeafStats)]),
    "ee_gen_points_uniform": (C.c_int, [U64, I32, I64, I64, P, P]),
    "ee_gen_points_lattice": (C.c_int, [U64, I32, I32, I64, DBL, I64, C.POINTER(I32), P, P]),
    "ee_gen_points_clustered": (C.c_int, [U64, I32, I32, DBL, I64, I64, P, P]),
    "ee_gen_knn_graph": (C.c_int, [P, P, I32, U8, I32, P, P, P, I64, C.POINTER(I64)]),
    "ee_graph_build_csr": (C.c_int, [I32, I32, P, P, P, P, P, P]),
    "ee_graph_content_hash": (U64, [I32, U8, P, P, I32, P, P, P, P]),
    "ee_gen_kopt_tasks": (C.c_int, [U64, P, P, I32, U8, P, P, U32, P, U64, P]),
    "ee_gen_shuffled_ks": (C.c_int, [U64, P, P]),
    "ee_isqrt_host": (None, [P, P, SZ]),
    "ee_isqrt_device": (C.c_int, [P, P, P, SZ]),
    "ee_distances_device": (C.c_int, [P, P, P, I32, U8, P, P, SZ, P]),
}

_lock = threading.Lock()
_lib: C.CDLL | None = None


def lib() -> C.CDLL:
    """The loaded backend library; raises ``BackendMissing`` when absent."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIBRARY_PATH.exists():
                raise BackendMissing(
                    f"{LIBRARY_PATH} is missing: build the HIP backend first "
                    "(python -c 'import __graft_entry__ as g; g.build()')")
            handle = C.CDLL(str(LIBRARY_PATH))
            for name, (restype, argtypes) in PROTOTYPES.items():
                fn = getattr(handle, name)
                fn.restype = restype
                fn.argtypes = argtypes
            _lib = handle
        return _lib


def last_error() -> str:
    return lib().ee_last_error().decode(errors="replace")


def check(status: int) -> None:
    if status == EE_OK:
        return
    message = last_error()
    if status == 1:
        raise InvalidArgument(status, message)
    if status == 2:
        raise OverflowFailure(status, message)
    if status == 5:
        raise Unavailable(status, message)
    raise EeError(status, message)


def ptr(array: np.ndarray | None) -> int | None:
    """Address of a C-contiguous numpy array (None for None)."""
    if array is None:
        return None
    if not array.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return array.ctypes.data
