"""Host mirror of the reference's device-backend interface for the two hot
paths, over the C ABI of ``libee_hip.so``.

Names and meaning follow the reference:

* ``cuda_backend_available``   - CudaBackendAvailable (elimination.hpp:333)
* ``find_jv_candidates_cuda``  - FindJvCandidatesCuda (elimination.hpp:334-336)
* ``clear_jv_cuda_cache``      - ClearJvCudaCache (elimination.hpp:338)
* ``build_kopt_reconnect_table`` - BuildKOptReconnectTable (local_search.hpp:50)
* ``evaluate_kopt_template_costs_cuda`` - detail::EvaluateKOptTemplateCostsCuda
  (local_search.hpp:236-239)
* ``kopt_cost_cuda_available`` / ``clear_kopt_cost_cuda_cache``
  (local_search.hpp:235, 242)

The reference keeps its device caches in ``thread_local`` state; here each
host thread owns a default ``Context`` (one per thread, created on first use)
and every function also accepts an explicit one.  Failures raise the
exception classes of ``_lib`` (``InvalidArgument`` is a ``ValueError``,
``OverflowFailure`` an ``OverflowError``), matching the reference's
invalid_argument / overflow_error / runtime_error split.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import threading
import weakref

import numpy as np

from . import _lib
from .graph import GraphSnapshot

JV_METHOD = 1  # EliminationMethod::kJv (types.hpp:28-31)


class Context:
    """One host thread's device-resident caches on one device."""

    def __init__(self, device: int = -1):
        handle = C.c_void_p()
        _lib.check(_lib.lib().ee_ctx_create(device, C.byref(handle)))
        self._handle = handle
        # batches, sessions and servers built on this context; they are
        # closed before the context itself
        self._children: weakref.WeakSet = weakref.WeakSet()

    def adopt(self, child) -> None:
        self._children.add(child)

    @property
    def handle(self) -> C.c_void_p:
        if self._handle is None:
            raise RuntimeError("context is closed")
        return self._handle

    @property
    def device(self) -> int:
        return int(_lib.lib().ee_ctx_device(self.handle))

    def close(self) -> None:
        if self._handle is not None:
            for child in list(self._children):
                child.close()
            _lib.lib().ee_ctx_destroy(self._handle)
            self._handle = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass


_tls = threading.local()


def default_context() -> Context:
    ctx = getattr(_tls, "ctx", None)
    if ctx is None:
        ctx = Context(-1)
        _tls.ctx = ctx
    return ctx


def _ctx(ctx: Context | None) -> Context:
    return ctx if ctx is not None else default_context()


def cuda_backend_available() -> tuple[bool, str]:
    reason = C.create_string_buffer(256)
    status = _lib.lib().ee_backend_available(-1, reason, len(reason))
    return status == _lib.EE_OK, reason.value.decode()


kopt_cost_cuda_available = cuda_backend_available


# ---------------------------------------------------------------------- JV
@dataclasses.dataclass
class JvCudaCacheUsage:
    static_hit: bool = False
    workspace_hit: bool = False
    resident_bytes: int = 0
    h2d_ms: float = 0.0
    kernel_ms: float = 0.0
    d2h_ms: float = 0.0


def find_jv_witnesses(graph: GraphSnapshot, ctx: Context | None = None):
    """Per edge slot: first JV witness or -1, plus the cache usage."""
    ctx = _ctx(ctx)
    witnesses = np.full(graph.edge_count, -1, dtype=np.int32)
    usage = _lib.JvUsage()
    view = graph.view()
    _lib.check(_lib.lib().ee_jv_find_candidates(ctx.handle, C.byref(view),
                                                _lib.ptr(witnesses) if graph.edge_count else None,
                                                C.byref(usage)))
    return witnesses, JvCudaCacheUsage(bool(usage.static_hit), bool(usage.workspace_hit),
                                       int(usage.resident_bytes), usage.h2d_ms, usage.kernel_ms,
                                       usage.d2h_ms)


def find_jv_candidates_cuda(graph: GraphSnapshot, ctx: Context | None = None):
    """FindJvCandidatesCuda: ([(edge_id, witness, JV)], selected_device, usage),
    candidates in ascending edge id (src/cuda/jv_cuda.cu:558-565)."""
    ctx = _ctx(ctx)
    witnesses, usage = find_jv_witnesses(graph, ctx)
    ids = np.flatnonzero(witnesses >= 0)
    return [(int(e), int(witnesses[e]), JV_METHOD) for e in ids], ctx.device, usage


def clear_jv_cuda_cache(ctx: Context | None = None) -> None:
    _lib.check(_lib.lib().ee_jv_clear_cache(_ctx(ctx).handle))


# ------------------------------------------------------------------- k-opt
@dataclasses.dataclass
class KOptReconnectTable:
    k: int
    templates: np.ndarray  # MATCHING_DTYPE
    generator_hash: int


@dataclasses.dataclass
class KOptCudaCacheUsage:
    snapshot_hit: bool = False
    template_hit: bool = False
    workspace_hit: bool = False
    resident_bytes: int = 0
    h2d_ms: float = 0.0
    kernel_ms: float = 0.0
    d2h_ms: float = 0.0


def build_kopt_reconnect_table(k: int) -> KOptReconnectTable:
    templates = np.zeros(945, dtype=_lib.MATCHING_DTYPE)
    count, hashed = C.c_uint32(), C.c_uint64()
    _lib.check(_lib.lib().ee_kopt_build_templates(k, _lib.ptr(templates), len(templates),
                                                  C.byref(count), C.byref(hashed)))
    return KOptReconnectTable(k, templates[:count.value].copy(), int(hashed.value))


def make_tasks(port_nodes, deleted_cost=None) -> np.ndarray:
    """KOptCostTask rows from an (R, <=10) port array (unused ports 0)."""
    port_nodes = np.asarray(port_nodes, dtype=np.int32)
    tasks = np.zeros(len(port_nodes), dtype=_lib.KOPT_TASK_DTYPE)
    tasks["port_nodes"][:, :port_nodes.shape[1]] = port_nodes
    if deleted_cost is not None:
        tasks["deleted_cost"] = deleted_cost
    return tasks


def evaluate_kopt_template_costs_cuda(graph: GraphSnapshot, table: KOptReconnectTable,
                                      tasks: np.ndarray, ctx: Context | None = None):
    """detail::EvaluateKOptTemplateCostsCuda: (costs[R*T] int64 row-major,
    selected_device, usage)."""
    ctx = _ctx(ctx)
    tasks = np.ascontiguousarray(tasks, dtype=_lib.KOPT_TASK_DTYPE)
    templates = np.ascontiguousarray(table.templates, dtype=_lib.MATCHING_DTYPE)
    costs = np.empty(len(tasks) * len(templates), dtype=np.int64)
    usage = _lib.KoptUsage()
    _lib.check(_lib.lib().ee_kopt_eval_costs(
        ctx.handle, _lib.ptr(graph.x), _lib.ptr(graph.y), graph.dimension, graph.distance_type,
        table.k, _lib.ptr(templates), len(templates), table.generator_hash,
        _lib.ptr(tasks) if len(tasks) else None, len(tasks),
        _lib.ptr(costs) if len(costs) else None, C.byref(usage)))
    return costs, ctx.device, KOptCudaCacheUsage(
        bool(usage.snapshot_hit), bool(usage.template_hit), bool(usage.workspace_hit),
        int(usage.resident_bytes), usage.h2d_ms, usage.kernel_ms, usage.d2h_ms)


def clear_kopt_cost_cuda_cache(ctx: Context | None = None) -> None:
    _lib.check(_lib.lib().ee_kopt_clear_cache(_ctx(ctx).handle))
