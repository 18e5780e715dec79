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

    def set_jv_edges_per_wave(self, edges: int) -> None:
        """JV kernel shape for later launches: 2 (default) pairs edge slots
        on 32-lane halves where both rows fit, 1 gives each a wavefront."""
        _lib.check(_lib.lib().ee_ctx_set_jv_edges_per_wave(self.handle, edges))

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
                                      tasks: np.ndarray, ctx: Context | None = None,
                                      no_alloc: bool = False):
    """detail::EvaluateKOptTemplateCostsCuda: (costs[R*T] int64 row-major,
    selected_device, usage).  no_alloc: the same evaluation through
    ee_kopt_eval_costs_noalloc, staged in the device's fixed scratch array
    (no allocation, no cache, at most EE_NOALLOC_SCRATCH_BYTES of staging)."""
    ctx = _ctx(ctx)
    tasks = np.ascontiguousarray(tasks, dtype=_lib.KOPT_TASK_DTYPE)
    templates = np.ascontiguousarray(table.templates, dtype=_lib.MATCHING_DTYPE)
    costs = np.empty(len(tasks) * len(templates), dtype=np.int64)
    usage = _lib.KoptUsage()
    entry = _lib.lib().ee_kopt_eval_costs_noalloc if no_alloc else _lib.lib().ee_kopt_eval_costs
    _lib.check(entry(
        ctx.handle, _lib.ptr(graph.x), _lib.ptr(graph.y), graph.dimension, graph.distance_type,
        table.k, _lib.ptr(templates), len(templates), table.generator_hash,
        _lib.ptr(tasks) if len(tasks) else None, len(tasks),
        _lib.ptr(costs) if len(costs) else None, C.byref(usage)))
    return costs, ctx.device, KOptCudaCacheUsage(
        bool(usage.snapshot_hit), bool(usage.template_hit), bool(usage.workspace_hit),
        int(usage.resident_bytes), usage.h2d_ms, usage.kernel_ms, usage.d2h_ms)


def certify_kopt_costs(graph: GraphSnapshot, table: KOptReconnectTable, tasks: np.ndarray,
                       cells: np.ndarray):
    """ee_kopt_certify: the host's exact CPU matrix of `tasks` compared cell
    by cell with `cells` (the certification of EvaluateKOptTemplateCosts,
    src/cpu/kopt_search.cpp:1455-1475).  Returns (improving masks uint32
    [rows, ceil(T / 32)], improving cell count); raises EeError
    (EE_ERR_RUNTIME) with .mismatch = the first differing cell otherwise."""
    tasks = np.ascontiguousarray(tasks, dtype=_lib.KOPT_TASK_DTYPE)
    templates = np.ascontiguousarray(table.templates, dtype=_lib.MATCHING_DTYPE)
    T = len(templates)
    cells = np.ascontiguousarray(cells, dtype=np.int64).reshape(-1)
    if cells.size != len(tasks) * T:
        raise ValueError("cells must hold rows x templates values")
    words = (T + 31) // 32
    improving = np.zeros((len(tasks), words), np.uint32)
    mismatch, count = C.c_uint64(), C.c_uint64()
    status = _lib.lib().ee_kopt_certify(
        _lib.ptr(graph.x), _lib.ptr(graph.y), graph.dimension, graph.distance_type, table.k,
        _lib.ptr(templates), T, _lib.ptr(tasks) if len(tasks) else None, len(tasks),
        _lib.ptr(cells) if cells.size else None, _lib.ptr(improving) if improving.size else None,
        C.byref(mismatch), C.byref(count))
    try:
        _lib.check(status)
    except _lib.EeError as err:
        err.mismatch = int(mismatch.value)
        raise
    return improving, int(count.value)


@dataclasses.dataclass
class HtReplyCudaCacheUsage:
    graph_hit: bool = False
    workspace_hit: bool = False
    resident_bytes: int = 0
    h2d_ms: float = 0.0
    kernel_ms: float = 0.0
    d2h_ms: float = 0.0


def _ht_usage(u) -> HtReplyCudaCacheUsage:
    return HtReplyCudaCacheUsage(bool(u.graph_hit), bool(u.workspace_hit), int(u.resident_bytes),
                                 u.h2d_ms, u.kernel_ms, u.d2h_ms)


def evaluate_ht_hamilton_replies_cuda(graph: GraphSnapshot, target, centers,
                                      ctx: Context | None = None):
    """EvaluateHtHamiltonRepliesCuda (src/cuda/ht_replies.cu:585-663):
    (offsets uint64[len(centers) + 1], replies int32[total, 3] as
    (center, first, second), usage)."""
    ctx = _ctx(ctx)
    centers = np.ascontiguousarray(centers, dtype=np.int32)
    offsets = np.zeros(len(centers) + 1, dtype=np.uint64)
    total, usage = C.c_uint64(), _lib.HtUsage()
    view = graph.view()
    # the reply count is known after the device's count pass: size the
    # buffer from it and copy the replies left on the device
    _lib.check(_lib.lib().ee_ht_hamilton_replies(
        ctx.handle, C.byref(view), int(target[0]), int(target[1]),
        _lib.ptr(centers) if len(centers) else None, len(centers), _lib.ptr(offsets),
        None, 0, C.byref(total), C.byref(usage)))
    replies = np.empty((total.value, 3), dtype=np.int32)
    if total.value:
        _lib.check(_lib.lib().ee_ht_copy_replies(ctx.handle, 0, total.value, _lib.ptr(replies)))
    return offsets, replies, _ht_usage(usage)


def evaluate_ht_end_replies_cuda(graph: GraphSnapshot, tasks, ctx: Context | None = None):
    """EvaluateHtEndRepliesCuda (src/cuda/ht_replies.cu:667-751): tasks as
    (endpoint, internal_neighbor) rows; (offsets uint64[len + 1], replies
    int32[total, 2] canonical edges, usage)."""
    ctx = _ctx(ctx)
    tasks = np.ascontiguousarray(np.asarray(tasks, dtype=np.int32).reshape(-1, 2))
    offsets = np.zeros(len(tasks) + 1, dtype=np.uint64)
    total, usage = C.c_uint64(), _lib.HtUsage()
    view = graph.view()
    _lib.check(_lib.lib().ee_ht_end_replies(
        ctx.handle, C.byref(view), _lib.ptr(tasks) if len(tasks) else None, len(tasks),
        _lib.ptr(offsets), None, 0, C.byref(total), C.byref(usage)))
    replies = np.empty((total.value, 2), dtype=np.int32)
    if total.value:
        _lib.check(_lib.lib().ee_ht_copy_replies(ctx.handle, 0, total.value, _lib.ptr(replies)))
    return offsets, replies, _ht_usage(usage)


def clear_ht_reply_cuda_cache(ctx: Context | None = None) -> None:
    _lib.check(_lib.lib().ee_ht_clear_cache(_ctx(ctx).handle))


@dataclasses.dataclass
class TourContext:
    """The TourContext fields the reference's k-opt cost cursor reads
    (src/cpu/kopt_search.cpp:125-134): the tour, the exact cost of the edge
    leaving each position (-1 off the paths) and the selectable positions
    (path-edge positions, the required edge's first)."""
    tour: np.ndarray        # int32 [L]
    edge_costs: np.ndarray  # int64 [L]
    selectable: np.ndarray  # uint64 [S]

    def __post_init__(self):
        self.tour = np.ascontiguousarray(self.tour, dtype=np.int32)
        self.edge_costs = np.ascontiguousarray(self.edge_costs, dtype=np.int64)
        self.selectable = np.ascontiguousarray(self.selectable, dtype=np.uint64)

    def view(self) -> _lib.TourView:
        return _lib.TourView(_lib.ptr(self.tour), _lib.ptr(self.edge_costs),
                             _lib.ptr(self.selectable) if len(self.selectable) else None,
                             len(self.tour), len(self.selectable))


def kopt_cursor_count(tour: TourContext, k: int) -> int:
    """Deletion sets of size k the reference cursor enumerates:
    C(|selectable| - 1, k - 1) for 3 <= k <= 5 and k < |selectable|."""
    out = C.c_uint64()
    view = tour.view()
    _lib.check(_lib.lib().ee_kopt_cursor_count(C.byref(view), k, C.byref(out)))
    return int(out.value)


@dataclasses.dataclass
class KOptCursorBlock:
    """One block of KOptSearchCursor::PrepareNextBlock
    (src/cpu/kopt_search.cpp:2123-2156) with its evaluated costs."""
    k: int
    first: int               # rank of the block's first deletion set within k
    tasks: np.ndarray        # KOPT_TASK_DTYPE [rows]
    positions: np.ndarray    # uint32 [rows, k], each row's sorted deleted positions
    costs: np.ndarray | None  # int64 [rows * T] (output "costs")
    masks: np.ndarray | None  # uint32 [rows, ceil(T / 32)] (output "masks")
    usage: KOptCudaCacheUsage


def evaluate_kopt_cursor_cuda(graph: GraphSnapshot | None, table: KOptReconnectTable | None,
                              tour: TourContext, k: int, first: int, count: int,
                              output: str | None = "costs",
                              ctx: Context | None = None) -> KOptCursorBlock:
    """Deletion sets [first, first + count) of size k built and evaluated on
    the device (ee_kopt_cursor_eval).  output: "costs", "masks" or None (the
    rows only; graph may then be a node count holder and table None)."""
    ctx = _ctx(ctx)
    view = tour.view()
    total = kopt_cursor_count(tour, k)
    rows = max(0, min(count, total - first))
    T = len(table.templates) if table is not None else 0
    tasks = np.zeros(max(rows, 1), dtype=_lib.KOPT_TASK_DTYPE)
    positions = np.zeros((max(rows, 1), k), dtype=np.uint32)
    costs = masks = None
    if output == "costs":
        costs = np.empty(max(rows * T, 1), dtype=np.int64)
    elif output == "masks":
        masks = np.zeros((max(rows, 1), (T + 31) // 32), dtype=np.uint32)
    elif output is not None:
        raise ValueError("output must be 'costs', 'masks' or None")
    templates = (np.ascontiguousarray(table.templates, dtype=_lib.MATCHING_DTYPE)
                 if table is not None else None)
    produced = C.c_uint32()
    usage = _lib.KoptUsage()
    _lib.check(_lib.lib().ee_kopt_cursor_eval(
        ctx.handle, _lib.ptr(graph.x) if output else None, _lib.ptr(graph.y) if output else None,
        graph.dimension, graph.distance_type if output else 0, C.byref(view), k,
        _lib.ptr(templates) if templates is not None else None, T,
        table.generator_hash if table is not None else 0, first, count,
        _lib.ptr(tasks), _lib.ptr(positions),
        _lib.ptr(costs) if costs is not None else None,
        _lib.ptr(masks) if masks is not None else None, C.byref(produced), C.byref(usage)))
    n = produced.value
    return KOptCursorBlock(
        k, first, tasks[:n], positions[:n], costs[:n * T] if costs is not None else None,
        masks[:n] if masks is not None else None,
        KOptCudaCacheUsage(bool(usage.snapshot_hit), bool(usage.template_hit),
                           bool(usage.workspace_hit), int(usage.resident_bytes), usage.h2d_ms,
                           usage.kernel_ms, usage.d2h_ms))


class KOptCostCursor:
    """The block sequence of the reference's k-opt cost cursor
    (KOptSearchCursor, src/cpu/kopt_search.cpp:2098-2156, 2258-2271) with each
    block's task rows built and evaluated on the device: for k = 3 .. max_k
    (k < |selectable|), blocks of at most cost_batch_size deletion sets, a
    block never spanning two k; max_deletion_sets > 0 caps the sets over all
    blocks, as the cursor's budget does."""

    def __init__(self, graph: GraphSnapshot, tour: TourContext, max_k: int = 5,
                 cost_batch_size: int = 4096, max_deletion_sets: int = 0,
                 output: str = "costs", ctx: Context | None = None):
        if not 3 <= max_k <= 5 or cost_batch_size <= 0:
            raise ValueError("max_k must lie in [3, 5] and cost_batch_size must be positive")
        self.graph, self.tour, self.max_k = graph, tour, max_k
        self.batch, self.budget, self.output, self.ctx = (cost_batch_size, max_deletion_sets,
                                                          output, ctx)
        self.tables = {k: build_kopt_reconnect_table(k) for k in range(3, max_k + 1)}
        self.deletion_sets = 0
        self.budget_exhausted = False

    def __iter__(self):
        for k in range(3, self.max_k + 1):
            total = kopt_cursor_count(self.tour, k)
            first = 0
            while first < total:
                rows = min(self.batch, total - first)
                if self.budget:
                    left = self.budget - self.deletion_sets
                    if left <= 0:
                        self.budget_exhausted = True
                        return
                    rows = min(rows, left)
                block = evaluate_kopt_cursor_cuda(self.graph, self.tables[k], self.tour, k, first,
                                                  rows, self.output, self.ctx)
                self.deletion_sets += len(block.tasks)
                first += len(block.tasks)
                yield block


def clear_kopt_cost_cuda_cache(ctx: Context | None = None) -> None:
    _lib.check(_lib.lib().ee_kopt_clear_cache(_ctx(ctx).handle))
