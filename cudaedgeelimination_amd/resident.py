"""Device-resident forms of the two hot paths: many k-opt segments per launch
(``KoptBatch``), multi-epoch JV with the snapshot kept on the device
(``JvSession``, ``JvFleet``), served queries (``KoptServer``, and
``DeviceQueue`` through a persistent kernel) and the
certified leaf pipeline (``LeafPipeline``).  Thin owners over the C ABI in
include/ee_hip.h."""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import _lib
from .backend import Context, TourContext
from .graph import GraphSnapshot

TEMPLATES_PER_K = {3: 4, 4: 25, 5: 208}


class KoptBatch:
    """Instances (coordinates) plus (instance, k, rows) segments evaluated in
    one launch; the cost output stays on the device until downloaded."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        handle = C.c_void_p()
        _lib.check(_lib.lib().ee_kopt_batch_create(ctx.handle, C.byref(handle)))
        self._h = handle
        ctx.adopt(self)
        self.cell_offsets = np.zeros(1, np.uint64)
        self.segments: list[tuple[int, int, int]] = []

    def close(self):
        if self._h is not None:
            _lib.lib().ee_kopt_batch_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def add_instance(self, graph: GraphSnapshot) -> int:
        inst = C.c_uint32()
        _lib.check(_lib.lib().ee_kopt_batch_add_instance(
            self._h, _lib.ptr(graph.x), _lib.ptr(graph.y), graph.dimension, graph.distance_type,
            C.byref(inst)))
        return inst.value

    def set_work(self, segments, tasks: np.ndarray) -> np.ndarray:
        """segments: [(instance, k, task_count)], tasks concatenated in
        segment order.  Returns the first cell of every segment in the cost
        pool (segments start on 16-cell boundaries) and the pool size."""
        descs = (_lib.SegmentDesc * max(len(segments), 1))()
        for i, (inst, k, count) in enumerate(segments):
            descs[i].instance, descs[i].k, descs[i].task_count = inst, k, count
        tasks = np.ascontiguousarray(tasks, dtype=_lib.KOPT_TASK_DTYPE)
        offsets = np.zeros(len(segments) + 1, np.uint64)
        _lib.check(_lib.lib().ee_kopt_batch_set_work(
            self._h, descs, len(segments), _lib.ptr(tasks) if len(tasks) else None,
            _lib.ptr(offsets)))
        self.cell_offsets = offsets
        self.segments = list(segments)
        return offsets

    @property
    def cells(self) -> int:
        """Cells of all segments (the pool may hold padding between them)."""
        return sum(count * TEMPLATES_PER_K[k] for _, k, count in getattr(self, "segments", []))

    @property
    def pool_cells(self) -> int:
        return int(self.cell_offsets[-1])

    def tuning(self) -> dict:
        t = _lib.KoptTuning()
        _lib.check(_lib.lib().ee_kopt_batch_get_tuning(self._h, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, rows_per_tile, grid=0, block=256, non_temporal=0):
        t = _lib.KoptTuning()
        for i, r in enumerate(rows_per_tile):
            t.rows_per_tile[i] = r
        t.grid, t.block, t.non_temporal = grid, block, non_temporal
        _lib.check(_lib.lib().ee_kopt_batch_set_tuning(self._h, C.byref(t)))

    def autotune(self, launches_per_candidate: int = 10, cache_path: str | None = None,
                 pipelined: bool = False) -> dict:
        """Times tuning candidates on the batch and keeps the fastest;
        pipelined: in pipelined runs at the batch's pipeline depth."""
        t = _lib.KoptTuning()
        _lib.check(_lib.lib().ee_kopt_batch_autotune_runs(
            self._h, launches_per_candidate, _lib.KOPT_RUN_PIPELINED if pipelined else 0,
            cache_path.encode() if cache_path else None, C.byref(t)))
        return t.as_dict()

    def run(self, launches: int = 1, use_graph: bool = False, pipelined: bool = False) -> float:
        """Mean device milliseconds of one launch (HIP events).  pipelined:
        the launches are dealt over the pipeline depth's streams and cost
        pools, so a launch starts while the previous one drains (cost output
        only)."""
        ms = C.c_double()
        flags = (_lib.KOPT_RUN_GRAPH if use_graph else 0) | (_lib.KOPT_RUN_PIPELINED if pipelined else 0)
        _lib.check(_lib.lib().ee_kopt_batch_run(self._h, launches, flags, C.byref(ms)))
        return ms.value

    def run_stamped(self, launches: int, pipelined: bool = True):
        """(start_us, end_us) of each launch: its first wavefront's start and
        last wavefront's end on the device clock, from the run's first start."""
        start = np.zeros(launches, np.float64)
        end = np.zeros(launches, np.float64)
        _lib.check(_lib.lib().ee_kopt_batch_run_stamped(
            self._h, launches, _lib.KOPT_RUN_PIPELINED if pipelined else 0, _lib.ptr(start), _lib.ptr(end)))
        return start, end

    def set_pipeline_depth(self, depth: int) -> None:
        """Streams and cost pools of pipelined runs (2 to 4)."""
        _lib.check(_lib.lib().ee_kopt_batch_set_pipeline_depth(self._h, depth))

    def download(self, first: int = 0, count: int | None = None) -> np.ndarray:
        count = self.pool_cells - first if count is None else count
        out = np.empty(count, np.int64)
        _lib.check(_lib.lib().ee_kopt_batch_download(self._h, first, count,
                                                     _lib.ptr(out) if count else None))
        return out

    def set_output(self, mode: str) -> None:
        """"costs" (int64 cells) or "masks" (per task row, the bit mask of the
        templates whose cell is below the row's deleted_cost)."""
        _lib.check(_lib.lib().ee_kopt_batch_set_output(self._h, {"costs": 0, "masks": 1}[mode]))

    def mask_words(self) -> int:
        w = C.c_uint32()
        _lib.check(_lib.lib().ee_kopt_batch_mask_words(self._h, C.byref(w)))
        return w.value

    def download_masks(self, first_row: int = 0, count: int | None = None) -> np.ndarray:
        """[count, mask_words] uint32; rows in segment order."""
        rows = sum(c for _, _, c in getattr(self, "segments", []))
        count = rows - first_row if count is None else count
        w = self.mask_words()
        out = np.empty((count, w), np.uint32)
        _lib.check(_lib.lib().ee_kopt_batch_download_masks(self._h, first_row, count,
                                                           _lib.ptr(out) if count else None))
        return out

    def segment_costs(self, index: int, pool: int = 0) -> np.ndarray:
        """Cells of segment `index`, from cost pool `pool` of pipelined runs
        (0: the pool ``download`` reads)."""
        _, k, count = self.segments[index]
        first, cells = int(self.cell_offsets[index]), count * TEMPLATES_PER_K[k]
        if pool == 0:
            return self.download(first, cells)
        out = np.empty(cells, np.int64)
        _lib.check(_lib.lib().ee_kopt_batch_download_pool(self._h, pool, first, cells,
                                                          _lib.ptr(out) if cells else None))
        return out

    def resident_bytes(self) -> int:
        return int(_lib.lib().ee_kopt_batch_resident_bytes(self._h))


@dataclasses.dataclass
class JvRun:
    epochs: int
    edges_scanned: int
    proposed: int
    committed: int
    initial_hash: int
    final_hash: int
    device_ms: float
    h2d_ms: float
    d2h_ms: float
    propose_ms: float
    verify_ms: float
    commit_ms: float
    total_ms: float
    proof: str
    compact_ms: float = 0.0    # device commit (deactivate + CSR compaction)
    hash_wait_ms: float = 0.0  # end-of-run wait for the snapshot hashes


class JvSession:
    """RunJvElimination with device proposal over a resident snapshot."""

    def __init__(self, ctx: Context, graph: GraphSnapshot):
        self.ctx = ctx
        self.graph = graph
        handle = C.c_void_p()
        view = graph.view()
        _lib.check(_lib.lib().ee_jv_session_create(ctx.handle, C.byref(view), C.byref(handle)))
        self._h = handle
        ctx.adopt(self)

    def close(self):
        if self._h is not None:
            _lib.lib().ee_jv_session_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def run(self, max_rounds: int = 100) -> JvRun:
        st = _lib.JvRunStats()
        _lib.check(_lib.lib().ee_jv_session_run(self._h, max_rounds, C.byref(st)))
        length = C.c_size_t()
        _lib.check(_lib.lib().ee_jv_session_proof(self._h, None, 0, C.byref(length)))
        buf = C.create_string_buffer(length.value + 1)
        _lib.check(_lib.lib().ee_jv_session_proof(self._h, buf, len(buf), C.byref(length)))
        return JvRun(st.epochs, st.edges_scanned, st.proposed, st.committed, st.initial_hash,
                     st.final_hash, st.device_ms, st.h2d_ms, st.d2h_ms, st.propose_ms,
                     st.verify_ms, st.commit_ms, st.total_ms, buf.value.decode(), st.compact_ms,
                     st.hash_wait_ms)

    def inject_witness(self, edge: int, witness: int) -> None:
        """Test hook: the next run's first epoch sees `witness` for `edge` in
        place of the device's, before host re-verification."""
        _lib.check(_lib.lib().ee_jv_session_inject_witness(self._h, edge, witness))

    def active(self) -> np.ndarray:
        out = np.empty(self.graph.edge_count, np.uint8)
        _lib.check(_lib.lib().ee_jv_session_active(self._h, _lib.ptr(out)))
        return out

    def device_snapshot(self) -> dict:
        """The device's active flags and CSR after the last run."""
        m, n = self.graph.edge_count, self.graph.dimension
        active = np.empty(max(m, 1), np.uint8)
        ro = np.empty(n + 1, np.int32)
        nb = np.empty(max(2 * m, 1), np.int32)
        eid = np.empty(max(2 * m, 1), np.int32)
        _lib.check(_lib.lib().ee_jv_session_device_snapshot(
            self._h, _lib.ptr(active), _lib.ptr(ro), _lib.ptr(nb), _lib.ptr(eid), 2 * m))
        a = int(ro[n])
        return {"edge_active": active[:m], "row_offsets": ro, "neighbors": nb[:a],
                "csr_edge_ids": eid[:a]}


class JvFleet:
    """Many independent JV runs in lockstep: one upload, launch and download
    per epoch for all graphs (optionally a replayed HIP graph); each graph's
    proof equals its standalone JvSession proof."""

    def __init__(self, ctx: Context, graphs: list[GraphSnapshot]):
        self.ctx = ctx
        self.graphs = list(graphs)
        self._views = (_lib.GraphView * len(self.graphs))(*[g.view() for g in self.graphs])
        handle = C.c_void_p()
        _lib.check(_lib.lib().ee_jv_fleet_create(ctx.handle, self._views, len(self.graphs),
                                                 C.byref(handle)))
        self._h = handle
        ctx.adopt(self)

    def close(self):
        if self._h is not None:
            _lib.lib().ee_jv_fleet_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def run(self, max_rounds: int = 100, use_graph: bool = False) -> dict:
        st = _lib.JvFleetStats()
        _lib.check(_lib.lib().ee_jv_fleet_run(self._h, max_rounds, 1 if use_graph else 0,
                                              C.byref(st)))
        return st.as_dict()

    def proof(self, index: int) -> str:
        length = C.c_size_t()
        _lib.check(_lib.lib().ee_jv_fleet_proof(self._h, index, None, 0, C.byref(length)))
        buf = C.create_string_buffer(length.value + 1)
        _lib.check(_lib.lib().ee_jv_fleet_proof(self._h, index, buf, len(buf), C.byref(length)))
        return buf.value.decode()

    def instance_stats(self, index: int) -> dict:
        st = _lib.JvRunStats()
        _lib.check(_lib.lib().ee_jv_fleet_instance_stats(self._h, index, C.byref(st)))
        return {name: getattr(st, name) for name, _ in st._fields_ if name != "reserved"}

    def active(self, index: int) -> np.ndarray:
        out = np.empty(self.graphs[index].edge_count, np.uint8)
        _lib.check(_lib.lib().ee_jv_fleet_active(self._h, index, _lib.ptr(out)))
        return out


class KoptServer:
    """Serving loop over device-resident graphs (ee_server_* in ee_hip.h).

    batching=True: continuous batching with two pipeline slots;
    batching=False: the plain per-query path (one request per launch, one
    synchronous upload / launch / download at a time)."""

    def __init__(self, ctx: Context, hbm_budget_bytes: int = 1 << 30,
                 max_batch_cells: int = 1 << 21, max_batch_requests: int = 256,
                 pipeline_depth: int = 2, batching: bool = True):
        cfg = _lib.ServerConfig()
        cfg.hbm_budget_bytes = hbm_budget_bytes
        cfg.max_batch_cells = max_batch_cells
        cfg.max_batch_requests = max_batch_requests if batching else 1
        cfg.pipeline_depth = pipeline_depth if batching else 1
        handle = C.c_void_p()
        _lib.check(_lib.lib().ee_server_create(ctx.handle, C.byref(cfg), C.byref(handle)))
        self._h = handle
        self.ctx = ctx
        ctx.adopt(self)
        self._buffers: dict[int, tuple[np.ndarray, np.ndarray]] = {}

    def close(self):
        if self._h is not None:
            _lib.lib().ee_server_destroy(self._h)
            self._h = None
            self._buffers.clear()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def add_graph(self, graph: GraphSnapshot) -> int:
        gid = C.c_uint32()
        _lib.check(_lib.lib().ee_server_add_graph(self._h, _lib.ptr(graph.x), _lib.ptr(graph.y),
                                                  graph.dimension, graph.distance_type,
                                                  C.byref(gid)))
        return gid.value

    def submit(self, graph: int, k: int, tasks: np.ndarray, out: np.ndarray | None = None) -> int:
        tasks = np.ascontiguousarray(tasks, dtype=_lib.KOPT_TASK_DTYPE)
        if out is None:
            out = np.empty(len(tasks) * TEMPLATES_PER_K[k], np.int64)
        rid = C.c_uint64()
        _lib.check(_lib.lib().ee_server_submit(self._h, graph, k, _lib.ptr(tasks), len(tasks),
                                               _lib.ptr(out), C.byref(rid)))
        self._buffers[rid.value] = (tasks, out)
        return rid.value

    def add_jv_graph(self, graph: GraphSnapshot) -> int:
        """Makes a JV snapshot resident for recheck requests."""
        gid = C.c_uint32()
        view = graph.view()
        _lib.check(_lib.lib().ee_server_add_jv_graph(self._h, C.byref(view), C.byref(gid)))
        return gid.value

    def submit_jv(self, jv_graph: int, edges: np.ndarray, out: np.ndarray | None = None) -> int:
        """Queues a JV recheck of the given edge slots; wait() returns the
        witness per slot (-1: not eliminable)."""
        edges = np.ascontiguousarray(edges, dtype=np.int32)
        if out is None:
            out = np.empty(len(edges), np.int32)
        rid = C.c_uint64()
        _lib.check(_lib.lib().ee_server_submit_jv(self._h, jv_graph, _lib.ptr(edges), len(edges),
                                                  _lib.ptr(out), C.byref(rid)))
        self._buffers[rid.value] = (edges, out)
        return rid.value

    def wait(self, request_id: int, timeout_ms: float = -1.0):
        """(results, result dict) of a completed request: costs of a leaf-cost
        query, witnesses of a JV recheck."""
        res = _lib.RequestResult()
        status = _lib.lib().ee_server_wait(self._h, request_id, timeout_ms, C.byref(res))
        _, out = self._buffers.pop(request_id, (None, None))
        _lib.check(status)
        return out, {"batch_size": res.batch_size, "batch_id": res.batch_id,
                     "queue_ms": res.queue_ms, "service_ms": res.service_ms,
                     "latency_ms": res.latency_ms}

    def submit_many(self, requests) -> list[int]:
        """Queues many requests at once: each is (graph, k, task rows) for a
        leaf-cost query or (jv_graph, 0, int32 edge slots) for a JV recheck.
        The inputs are not copied; this object keeps them (and the outputs)
        alive until wait_many() collects the requests."""
        descs = (_lib.ServerRequest * max(len(requests), 1))()
        keep = []
        for i, (graph, k, data) in enumerate(requests):
            if k:
                data = np.ascontiguousarray(data, dtype=_lib.KOPT_TASK_DTYPE)
                out = np.empty(len(data) * TEMPLATES_PER_K[k], np.int64)
            else:
                data = np.ascontiguousarray(data, dtype=np.int32)
                out = np.empty(len(data), np.int32)
            descs[i].kind, descs[i].graph, descs[i].k = (0 if k else 1), graph, k
            descs[i].input, descs[i].count, descs[i].output = _lib.ptr(data), len(data), _lib.ptr(out)
            keep.append((data, out))
        ids = np.zeros(max(len(requests), 1), np.uint64)
        _lib.check(_lib.lib().ee_server_submit_many(self._h, len(requests), descs, _lib.ptr(ids)))
        ids = [int(v) for v in ids[:len(requests)]]
        for rid, bufs in zip(ids, keep):
            self._buffers[rid] = bufs
        return ids

    def wait_many(self, request_ids, timeout_ms: float = -1.0):
        """[(results, latency_ms)] of the requests, in the given order."""
        ids = np.ascontiguousarray(request_ids, dtype=np.uint64)
        res = (_lib.RequestResult * max(len(ids), 1))()
        status = _lib.lib().ee_server_wait_many(self._h, len(ids), _lib.ptr(ids), timeout_ms, res)
        outs = [self._buffers.pop(int(rid), (None, None))[1] for rid in ids]
        _lib.check(status)
        return [(outs[i], res[i].latency_ms) for i in range(len(ids))]

    def stats(self) -> dict:
        st = _lib.ServerStats()
        _lib.check(_lib.lib().ee_server_stats_get(self._h, C.byref(st)))
        return {"requests": st.requests, "batches": st.batches, "cells": st.cells,
                "jv_edges": st.jv_edges, "peak_inflight_bytes": st.peak_inflight_bytes,
                "device_ms": st.device_ms}


class DeviceQueue:
    """Leaf-cost queries and JV rechecks through the device work queue
    (ee_queue_* in ee_hip.h): a persistent kernel polls a request ring in
    coherent host memory and raises a completion flag per request; no launch
    or copy per request.  Same submit / wait surface as KoptServer."""

    def __init__(self, ctx: Context, task_rows: int = 1 << 18, cost_cells: int = 1 << 23,
                 max_nodes: int = 1 << 20, ring_entries: int = 4096, workers: int = 256,
                 idle_exit_ms: float = 20.0, max_jv_nodes: int = 1 << 17, max_jv_edges: int = 1 << 20,
                 max_jv_adjacency: int = 1 << 21):
        cfg = _lib.QueueConfig()
        cfg.task_rows, cfg.cost_cells, cfg.max_nodes = task_rows, cost_cells, max_nodes
        cfg.ring_entries, cfg.workers, cfg.idle_exit_ms = ring_entries, workers, idle_exit_ms
        cfg.max_jv_nodes, cfg.max_jv_edges, cfg.max_jv_adjacency = max_jv_nodes, max_jv_edges, max_jv_adjacency
        handle = C.c_void_p()
        _lib.check(_lib.lib().ee_queue_create(ctx.handle, C.byref(cfg), C.byref(handle)))
        self._h = handle
        self.ctx = ctx
        ctx.adopt(self)
        self._buffers: dict[int, tuple[np.ndarray, np.ndarray]] = {}

    def close(self):
        if self._h is not None:
            _lib.lib().ee_queue_destroy(self._h)
            self._h = None
            self._buffers.clear()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def add_graph(self, graph: GraphSnapshot) -> int:
        gid = C.c_uint32()
        _lib.check(_lib.lib().ee_queue_add_graph(self._h, _lib.ptr(graph.x), _lib.ptr(graph.y),
                                                 graph.dimension, graph.distance_type, C.byref(gid)))
        return gid.value

    def submit(self, graph: int, k: int, tasks: np.ndarray, out: np.ndarray | None = None) -> int:
        tasks = np.ascontiguousarray(tasks, dtype=_lib.KOPT_TASK_DTYPE)
        if out is None:
            out = np.empty(len(tasks) * TEMPLATES_PER_K[k], np.int64)
        rid = C.c_uint64()
        _lib.check(_lib.lib().ee_queue_submit(self._h, graph, k, _lib.ptr(tasks), len(tasks),
                                              _lib.ptr(out), C.byref(rid)))
        self._buffers[rid.value] = (tasks, out)
        return rid.value

    def add_jv_graph(self, graph: GraphSnapshot) -> int:
        """Makes a JV snapshot resident for recheck requests."""
        gid = C.c_uint32()
        view = graph.view()
        _lib.check(_lib.lib().ee_queue_add_jv_graph(self._h, C.byref(view), C.byref(gid)))
        return gid.value

    def submit_jv(self, jv_graph: int, edges: np.ndarray, out: np.ndarray | None = None) -> int:
        """Queues a JV recheck of the given edge slots; wait() returns the
        witness per slot (-1: not eliminable)."""
        edges = np.ascontiguousarray(edges, dtype=np.int32)
        if out is None:
            out = np.empty(len(edges), np.int32)
        rid = C.c_uint64()
        _lib.check(_lib.lib().ee_queue_submit_jv(self._h, jv_graph, _lib.ptr(edges), len(edges),
                                                 _lib.ptr(out), C.byref(rid)))
        self._buffers[rid.value] = (edges, out)
        return rid.value

    def wait(self, request_id: int, timeout_ms: float = -1.0):
        """(costs, result dict) of a completed request."""
        res = _lib.RequestResult()
        status = _lib.lib().ee_queue_wait(self._h, request_id, timeout_ms, C.byref(res))
        _, out = self._buffers.pop(request_id, (None, None))
        _lib.check(status)
        return out, {"batch_size": res.batch_size, "batch_id": res.batch_id,
                     "queue_ms": res.queue_ms, "service_ms": res.service_ms,
                     "latency_ms": res.latency_ms}

    def submit_many(self, requests) -> list[int]:
        """Queues requests in order: (graph, k, task rows) leaf-cost queries
        or (jv_graph, 0, int32 edge slots) JV rechecks; the inputs are copied
        into the task arena, the outputs kept until wait_many()."""
        descs = (_lib.ServerRequest * max(len(requests), 1))()
        keep = []
        for i, (graph, k, data) in enumerate(requests):
            if k:
                data = np.ascontiguousarray(data, dtype=_lib.KOPT_TASK_DTYPE)
                out = np.empty(len(data) * TEMPLATES_PER_K[k], np.int64)
            else:
                data = np.ascontiguousarray(data, dtype=np.int32)
                out = np.empty(len(data), np.int32)
            descs[i].kind, descs[i].graph, descs[i].k = (0 if k else 1), graph, k
            descs[i].input, descs[i].count, descs[i].output = _lib.ptr(data), len(data), _lib.ptr(out)
            keep.append((data, out))
        ids = np.zeros(max(len(requests), 1), np.uint64)
        _lib.check(_lib.lib().ee_queue_submit_many(self._h, len(requests), descs, _lib.ptr(ids)))
        ids = [int(v) for v in ids[:len(requests)]]
        for rid, bufs in zip(ids, keep):
            self._buffers[rid] = bufs
        return ids

    def wait_many(self, request_ids, timeout_ms: float = -1.0):
        """[(costs, latency_ms)] of the requests, in the given order."""
        ids = np.ascontiguousarray(request_ids, dtype=np.uint64)
        res = (_lib.RequestResult * max(len(ids), 1))()
        status = _lib.lib().ee_queue_wait_many(self._h, len(ids), _lib.ptr(ids), timeout_ms, res)
        outs = [self._buffers.pop(int(rid), (None, None))[1] for rid in ids]
        _lib.check(status)
        return [(outs[i], res[i].latency_ms) for i in range(len(ids))]

    def inject_refusal(self, graph: int, task: np.ndarray) -> int:
        """Test hook: queues a one-row k = 3 request whose ring entry points
        past the task arena; the kernel refuses it, wait() raises and the
        queue refuses later submits."""
        task = np.ascontiguousarray(task[:1], dtype=_lib.KOPT_TASK_DTYPE)
        out = np.empty(TEMPLATES_PER_K[3], np.int64)
        rid = C.c_uint64()
        _lib.check(_lib.lib().ee_queue_inject_refusal(self._h, graph, _lib.ptr(task), _lib.ptr(out),
                                                      C.byref(rid)))
        self._buffers[rid.value] = (task, out)
        return rid.value

    def stats(self) -> dict:
        st = _lib.QueueStats()
        _lib.check(_lib.lib().ee_queue_stats_get(self._h, C.byref(st)))
        return {"requests": st.requests, "cells": st.cells, "launches": st.launches,
                "peak_inflight_cells": st.peak_inflight_cells, "jv_edges": st.jv_edges}


@dataclasses.dataclass
class LeafBlock:
    """One certified block of ee_leaf_next (copies of the pipeline's views)."""
    tag: int
    first: int
    k: int
    tasks: np.ndarray             # KOPT_TASK_DTYPE [rows]
    positions: np.ndarray | None  # uint32 [rows, k] (cursor blocks)
    costs: np.ndarray             # int64 [rows, T], certified
    improving: np.ndarray         # uint32 [rows, ceil(T / 32)]
    improving_cells: int


class LeafPipeline:
    """The certified leaf loop with device work ahead of host certification:
    submit up to `depth` blocks, then take them back in order with next()."""

    def __init__(self, ctx: Context, graph: GraphSnapshot, depth: int = 2, max_rows: int = 4096):
        self.ctx = ctx
        self.graph = graph
        handle = C.c_void_p()
        _lib.check(_lib.lib().ee_leaf_create(ctx.handle, _lib.ptr(graph.x), _lib.ptr(graph.y),
                                             graph.dimension, graph.distance_type, depth, max_rows,
                                             C.byref(handle)))
        self._h = handle
        ctx.adopt(self)

    def close(self):
        if self._h is not None:
            _lib.lib().ee_leaf_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def set_tour(self, tour: TourContext) -> None:
        self._tour = tour
        view = tour.view()
        _lib.check(_lib.lib().ee_leaf_set_tour(self._h, C.byref(view)))

    def submit_tasks(self, k: int, tasks: np.ndarray, tag: int = 0) -> None:
        tasks = np.ascontiguousarray(tasks, dtype=_lib.KOPT_TASK_DTYPE)
        _lib.check(_lib.lib().ee_leaf_submit_tasks(self._h, k, _lib.ptr(tasks) if len(tasks) else None,
                                                   len(tasks), tag))

    def submit_cursor(self, k: int, first: int, count: int, tag: int = 0) -> int:
        rows = C.c_uint32()
        _lib.check(_lib.lib().ee_leaf_submit_cursor(self._h, k, first, count, tag, C.byref(rows)))
        return int(rows.value)

    def next(self, copy: bool = True) -> LeafBlock:
        b = _lib.LeafBlock()
        _lib.check(_lib.lib().ee_leaf_next(self._h, C.byref(b)))
        rows, T, W = b.rows, b.template_count, b.mask_words

        def view(ptr, dtype, shape):
            count = int(np.prod(shape))
            if count == 0:
                return np.zeros(shape, dtype)
            buf = (C.c_char * (count * np.dtype(dtype).itemsize)).from_address(ptr)
            a = np.frombuffer(buf, dtype=dtype, count=count).reshape(shape)
            return a.copy() if copy else a

        return LeafBlock(
            int(b.tag), int(b.first), int(b.k),
            view(b.tasks, _lib.KOPT_TASK_DTYPE, (rows,)),
            view(b.positions, np.uint32, (rows, b.k)) if b.positions else None,
            view(b.costs, np.int64, (rows, T)),
            view(b.improving, np.uint32, (rows, W)),
            int(b.improving_cells))

    def inject_cell(self, cell: int, value: int) -> None:
        """Test hook: the next block handed out has `cell` of its host copy
        set to `value` before certification."""
        _lib.check(_lib.lib().ee_leaf_inject_cell(self._h, cell, value))

    def stats(self) -> dict:
        st = _lib.LeafStats()
        _lib.check(_lib.lib().ee_leaf_stats_get(self._h, C.byref(st)))
        return st.as_dict()
