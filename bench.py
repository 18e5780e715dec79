"""Benchmark of the MI355X backend on the k-opt leaf cost matrix workload
(n = 15,000 uniform points in [0, 10^6]^2, 8-NN graph, 65,536 task rows with
k = 3 / 4 / 5 quotas 32,768 / 19,661 / 13,107: 3,348,853 int64 cells) and,
beside it, the multi-epoch JV workload (60 x 60 drilling-grid lattice, 10-NN).

One step = one evaluation of the whole cost matrix by the resident batched
path (one kernel launch over all three k segments, inputs already in HBM).
Prints one JSON line; see the keys below.

  python bench.py [--gpus 1] [--steps 20] [--warmup 5] [--no-cpu] [--no-jv]
"""
from __future__ import annotations

import argparse
import ctypes as C
import hashlib
import json
import os
import pathlib
import statistics
import sys
import tempfile
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import numpy as np  # noqa: E402

from cudaedgeelimination_amd import backend, workloads as W  # noqa: E402
from cudaedgeelimination_amd.resident import DeviceQueue, JvSession, KoptBatch, KoptServer, LeafPipeline  # noqa: E402

METRIC = "exact k-opt cost cells/s + JV edges/s; p50/p99 query latency vs plain GPU path"
# streams (and cost pools) the timed k-opt launches are dealt over
PIPELINE_DEPTH = 4
HBM_PEAK_GBS = 8000.0
TEMPLATES = {3: 4, 4: 25, 5: 208}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def kopt_algorithmic_bytes(wl) -> int:
    """Tasks read once (48 B/row), cells written once (8 B), coordinates
    read once (16 B/node), templates read once (15 B each)."""
    rows = len(wl.tasks)
    cells = wl.cells()
    return 48 * rows + 8 * cells + 16 * wl.graph.dimension + 15 * sum(TEMPLATES.values())


# distinct port pairs the canonical templates use per row (one exact distance
# each), against k distances per cell in a per-cell evaluation
PAIRS = {3: 9, 4: 20, 5: 35}


def kopt_distance_evals(wl) -> tuple[int, int]:
    """(distances the kernel evaluates per launch, distances a per-cell
    evaluation would: k per cell)."""
    rows = {k: int((wl.ks == k).sum()) for k in (3, 4, 5)}
    return (sum(rows[k] * PAIRS[k] for k in rows),
            sum(rows[k] * TEMPLATES[k] * k for k in rows))


# The process's CPU set at start: the reference protocol's OMP_PROC_BIND makes
# the OpenMP runtime pin the thread that first runs a parallel region, and
# threads started later (the serving loop's) would inherit that one core.
START_AFFINITY = os.sched_getaffinity(0)


def restore_affinity() -> None:
    os.sched_setaffinity(0, START_AFFINITY)


def host_info() -> dict:
    """CPU model and the cores this process may use (the GPU host's share)."""
    model = "unknown"
    try:
        for line in pathlib.Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "usable_cores": len(START_AFFINITY),
            "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_DYNAMIC",
                                                     "OMP_PROC_BIND", "OMP_PLACES")}}


def timed(fn, budget_s: float, min_runs: int = 5, max_runs: int = 50):
    """One warm-up call, then timed calls until the budget is spent (at least
    min_runs): (median s, p95 s, runs, last result)."""
    result = fn()
    samples = []
    t_end = time.perf_counter() + budget_s
    while len(samples) < min_runs or (time.perf_counter() < t_end and len(samples) < max_runs):
        t0 = time.perf_counter()
        result = fn()
        samples.append(time.perf_counter() - t0)
    return statistics.median(samples), float(np.percentile(samples, 95)), len(samples), result


def cpu_baseline_kopt(wl, budget_s: float):
    """The reference's own CPU cost matrix (EvaluateKOptTemplateCosts with
    the CPU backend, OpenMP over rows, at most 8 threads), built from the
    reference sources into oracle/_ref; the oracle restatement otherwise."""
    import _oracle as O

    use_ref = O.ref_available()
    fn = O.ref_kopt_costs if use_ref else O.kopt_costs

    def full():
        outputs, threads = {}, 0
        for k in (3, 4, 5):
            out, th, _ = fn(wl.graph.x, wl.graph.y, wl.graph.distance_type, k, wl.bucket(k))
            threads = max(threads, th)
            outputs[k] = out
        return outputs, threads

    median, p95, runs, (outputs, threads) = timed(full, budget_s)
    restore_affinity()
    return {
        "value": wl.cells() / median,
        "unit": "cells/s",
        "cores": threads,
        "kind": "reference" if use_ref else "port",
        "sample": f"full cost matrix of the same workload ({wl.cells()} cells), one warm-up and "
                  f"{runs} timed repetitions, median {median * 1e3:.1f} ms, p95 {p95 * 1e3:.1f} ms",
        "p95_cells_per_s": wl.cells() / p95,
        "host": host_info(),
    }, outputs


def cpu_baseline_jv(graph):
    """Reference RunJvElimination with the CPU backend (serial), same input."""
    import ctypes as C

    import _oracle as O

    with tempfile.TemporaryDirectory() as tmp:
        tsp, edg = pathlib.Path(tmp) / "g.tsp", pathlib.Path(tmp) / "g.edg"
        graph.write(tsp, edg)
        buf = C.create_string_buffer(1 << 24)

        def run():
            if O.ref_available():
                size = O.ref().ref_jv_run_files(str(tsp).encode(), str(edg).encode(), 100, buf,
                                                len(buf))
                assert size > 0
                return buf.value.decode()
            return O.jv_run(O.OGraph.load(tsp, edg))

        median, p95, runs, text = timed(run, 2.0)
    restore_affinity()
    scanned = sum(int(line.split()[2]) for line in text.splitlines() if line.startswith("metrics "))
    return {"edges_per_s": scanned / median, "p95_edges_per_s": scanned / p95,
            "kind": "reference" if O.ref_available() else "port", "cores": 1,
            "median_ms": median * 1e3, "p95_ms": p95 * 1e3, "runs": runs}, text


def many_instance_leg(ctx, cpu: bool, sample: int = 16):
    """256 independent instances (n log-uniform in [200, 4000], uniform or
    8-cluster Gaussian coordinates, CEIL_2D for odd indices, 10-NN): JV epochs
    of all graphs in lockstep as one replayed HIP graph per epoch, then 4,096
    k-opt rows per instance (2,048 / 1,229 / 819) in one batched launch.  The
    first `sample` instances are checked against the reference CPU path,
    which is also timed on them."""
    import _oracle as O

    from cudaedgeelimination_amd.resident import JvFleet

    t0 = time.perf_counter()
    instances = [W.config4_instance(i) for i in range(256)]
    gen_s = time.perf_counter() - t0
    graphs = [g for g, _ in instances]
    out = {"instances": len(graphs), "nodes": sum(g.dimension for g in graphs),
           "edge_slots": sum(g.edge_count for g in graphs), "generate_s": gen_s}

    runs = []
    for _ in range(2):
        fleet = JvFleet(ctx, graphs)
        st = fleet.run(100, use_graph=True)
        runs.append((st, fleet))
    st, fleet = min(runs, key=lambda r: r[0]["total_ms"])
    jv = {"epochs": st["epochs"], "edges_scanned": st["edges_scanned"],
          "committed": st["committed"], "graph_launches": st["graph_launches"],
          "epoch_device_ms": st["device_ms"] / max(st["epochs"], 1),
          "device_edges_per_s": st["edges_scanned"] / (st["device_ms"] * 1e-3),
          "algorithm_edges_per_s": st["edges_scanned"] / (st["total_ms"] * 1e-3),
          "total_ms": st["total_ms"], "host_ms": st["host_ms"], "assemble_ms": st["assemble_ms"]}
    if cpu:
        t1 = time.perf_counter()
        want = [O.jv_run(O.OGraph.from_snapshot(g)) for g in graphs[:sample]]
        cpu_s = time.perf_counter() - t1
        scanned = sum(int(line.split()[2]) for text in want for line in text.splitlines()
                      if line.startswith("metrics "))
        jv["proofs_equal_cpu"] = f"{sum(O.canonical_proof(fleet.proof(i)) == O.canonical_proof(w) for i, w in enumerate(want))}/{sample}"
        jv["cpu_baseline"] = {"edges_per_s": scanned / cpu_s, "kind": "port", "cores": 1,
                              "sample": f"first {sample} instances, JV to fixed point"}
        restore_affinity()
    for _, f in runs:
        f.close()
    out["jv"] = jv

    batch = KoptBatch(ctx)
    segs, tasks = [], []
    for g, wl in instances:
        inst = batch.add_instance(g)
        for k in (3, 4, 5):
            rows = wl.bucket(k)
            segs.append((inst, k, len(rows)))
            tasks.append(rows)
    batch.set_work(segs, np.concatenate(tasks))
    tuning = batch.autotune(5)
    batch.run(3)
    ms = batch.run(10)
    graph_ms = batch.run(10, use_graph=True)
    cells = batch.cells
    kopt = {"rows": int(sum(len(t) for t in tasks)), "cells": cells, "tuning": tuning,
            "kernel_ms": ms, "graph_replay_kernel_ms": graph_ms,
            "cells_per_s": cells / (min(ms, graph_ms) * 1e-3)}
    if cpu:
        ok = 0
        for i in range(3 * sample):
            g = graphs[i // 3]
            k = segs[i][1]
            want, _, _ = O.kopt_costs(g.x, g.y, g.distance_type, k, tasks[i])
            ok += bool(np.array_equal(batch.segment_costs(i), want))
        kopt["segments_equal_cpu"] = f"{ok}/{3 * sample}"
        restore_affinity()
    batch.close()
    out["kopt"] = kopt
    return out


def serve_stream(ctx, graphs, pools, requests, is_jv, arrivals, mode: str, burst: bool, keep=False):
    """One request stream against a fresh server of `mode` ("batched",
    "plain", or "queue": the device work queue's persistent kernel): latency
    from each request's scheduled arrival to completion.  keep: also return
    the burst's results (key "outputs")."""
    import threading

    count = len(requests)
    kopt_cells = sum(len(t) * TEMPLATES[k] for _, k, t in requests if k)
    if mode == "queue":
        server = DeviceQueue(ctx, task_rows=1 << 20, cost_cells=1 << 25, max_nodes=1 << 16,
                             ring_entries=1 << 15)
    else:
        server = KoptServer(ctx, batching=(mode == "batched"), hbm_budget_bytes=2 << 30,
                            max_batch_cells=1 << 21)
    gids = [server.add_graph(g) for g in graphs]
    jids = [server.add_jv_graph(g) for g in graphs]
    warm = server.submit(gids[0], 3, pools[0][3][:8])
    server.wait(warm)
    latencies = np.zeros(count)
    late = np.zeros(count)
    outputs = None
    if burst:
        # the burst arrives as one message: queued in one call
        t0 = time.perf_counter()
        ids = server.submit_many([(gids[g], k, tasks) if k else (jids[g], 0, tasks)
                                  for g, k, tasks in requests])
        done = server.wait_many(ids)
        wall = time.perf_counter() - t0
        latencies[:] = [lat for _, lat in done]
        if keep:
            outputs = [out for out, _ in done]
    else:
        ids = [0] * count
        submitted = threading.Semaphore(0)

        def collect():
            for i in range(count):
                submitted.acquire()
                _, info = server.wait(ids[i])
                latencies[i] = info["latency_ms"] + late[i]

        waiter = threading.Thread(target=collect)
        waiter.start()
        t0 = time.perf_counter()
        for i, (g, k, tasks) in enumerate(requests):
            due = t0 + arrivals[i]
            while True:
                now = time.perf_counter()
                if now >= due:
                    break
                if due - now > 2e-4:
                    time.sleep(1e-4)
            ids[i] = server.submit(gids[g], k, tasks) if k else server.submit_jv(jids[g], tasks)
            late[i] = (time.perf_counter() - due) * 1e3  # late submit counts
            submitted.release()
        waiter.join()
        wall = time.perf_counter() - t0
    stats = server.stats()
    server.close()
    out = {
        "p50_ms": float(np.percentile(latencies, 50)),
        "p99_ms": float(np.percentile(latencies, 99)),
        "mean_ms": float(latencies.mean()),
        "jv_p99_ms": float(np.percentile(latencies[is_jv], 99)) if is_jv.any() else None,
        "cells_per_s": kopt_cells / wall,
        "jv_edges_per_s": stats.get("jv_edges", 0) / wall,
        # the queue launches once per busy period, the servers once per batch
        "launches": stats["launches"] if mode == "queue" else stats["batches"],
        "device_cells_per_s": (stats["cells"] / (stats["device_ms"] * 1e-3)
                               if stats.get("device_ms", 0) > 0 else None),
        "wall_s": wall,
        # how far behind schedule the client submitted (its own limit)
        "client_late_p50_ms": float(np.percentile(late, 50)) if not burst else None,
    }
    if outputs is not None:
        out["outputs"] = outputs
    return out


def serving_leg(ctx, qps_list, seconds: float, seed: int = 0xEE000005, jv_share: float = 0.2,
                burst: int = 20000, saturation_p99_ms: float = 0.5, saturation_rates=()):
    """Open-loop Poisson queries against four resident graphs (n = 1k, 5k,
    15k, 30k, uniform, 10-NN).  Leaf-cost queries: sizes lognormal (median
    2,896 cells, sigma 1.04) clipped to [128, 65,536] cells, k = 3 / 4 / 5 at
    50/30/20%.  A share `jv_share` of the requests are JV rechecks of random
    edge slots of the graph (count lognormal, median 256, sigma 1.0, clipped
    to [16, 4,096]).  Latency runs from the scheduled arrival to completion,
    so queueing counts (no coordinated omission).  Same request stream for the
    batched server and the plain per-query path.  A rate of 0 is a burst of
    `burst` requests all due at once, queued with one submit_many call: its
    cells/s is each path's capacity.  Then, per path, the saturation rate:
    the highest of `saturation_rates` (ascending) whose p99 stays within
    `saturation_p99_ms`, each rate run while the previous one held."""
    graphs = W.config5_graphs()
    pools = W.config5_row_pools(graphs, seed)
    out = {"graphs": [g.dimension for g in graphs], "points": []}

    def stream(qps, count_seconds, share):
        rng = np.random.default_rng(seed + qps)
        if qps == 0:
            # burst: every request is due at once, so the wall time measures
            # how fast each path drains a full queue (its capacity)
            count = burst
            arrivals = np.zeros(count)
        else:
            # at least 10,000 requests, so that p99 rests on 100 of them
            count = max(int(qps * count_seconds), 10000)
            arrivals = np.cumsum(rng.exponential(1.0 / qps, count))
        requests, is_jv = W.config5_requests(graphs, pools, count, rng, share)
        return requests, is_jv, arrivals

    def points(share, modes):
        pts = []
        for qps in qps_list:
            requests, is_jv, arrivals = stream(qps, seconds, share)
            kopt_cells = [len(t) * TEMPLATES[k] for _, k, t in requests if k]
            point = {"qps": qps if qps else "burst", "requests": len(requests), "jv_rechecks": int(is_jv.sum()),
                     "mean_cells": float(np.mean(kopt_cells)),
                     "mean_jv_edges": float(np.mean([len(t) for _, k, t in requests if not k] or [0]))}
            for mode in modes:
                point[mode] = serve_stream(ctx, graphs, pools, requests, is_jv, arrivals, mode, qps == 0,
                                           keep=(qps == 0))
            if qps == 0:
                # every path's burst results against the first path's
                outs = {mode: point[mode].pop("outputs") for mode in modes}
                differing = {mode: sum(not np.array_equal(a, b) for a, b in zip(outs[modes[0]], outs[mode]))
                             for mode in modes[1:]}
                point["results_equal_across_paths"] = not any(differing.values())
                if any(differing.values()):
                    point["requests_differing_from_" + modes[0]] = differing
            for mode in modes:
                if mode != "plain":
                    point[f"{mode}_over_plain_cells_per_s"] = point[mode]["cells_per_s"] / point["plain"]["cells_per_s"]
            pts.append(point)
        return pts

    def saturation(share, modes):
        sat = {"p99_bound_ms": saturation_p99_ms, "rates": list(saturation_rates)}
        for mode in modes:
            held, tried = None, []
            for qps in saturation_rates:
                requests, is_jv, arrivals = stream(qps, seconds, share)
                r = serve_stream(ctx, graphs, pools, requests, is_jv, arrivals, mode, False)
                tried.append({"qps": qps, "p50_ms": r["p50_ms"], "p99_ms": r["p99_ms"],
                              "client_late_p50_ms": r["client_late_p50_ms"], "launches": r["launches"]})
                log(f"saturation {mode} {qps} QPS: p99 {r['p99_ms']:.3f} ms")
                if r["p99_ms"] > saturation_p99_ms:
                    break
                held = qps
            sat[mode] = {"saturation_qps": held, "points": tried,
                         # the last rate held with nothing left to try: a lower bound
                         "at_last_rate": held == saturation_rates[-1]}
        for mode in modes:
            if mode != "plain" and sat[mode]["saturation_qps"] and sat["plain"]["saturation_qps"]:
                sat[f"{mode}_over_plain_qps"] = sat[mode]["saturation_qps"] / sat["plain"]["saturation_qps"]
        return sat

    out["points"] = points(jv_share, ("queue", "batched", "plain"))
    if saturation_rates:
        out["saturation"] = saturation(jv_share, ("queue", "batched", "plain"))
    # leaf-cost queries alone
    lo = {"jv_share": 0.0, "points": points(0.0, ("queue", "batched", "plain"))}
    if saturation_rates:
        lo["saturation"] = saturation(0.0, ("queue", "batched", "plain"))
    out["leaf_only"] = lo
    return out


def pack_improving(costs: np.ndarray, tasks: np.ndarray, T: int, words: int) -> np.ndarray:
    """Per row, bit j of word j / 32: cell (row, j) < the row's deleted_cost."""
    improving = costs.reshape(len(tasks), T) < tasks["deleted_cost"][:, None]
    out = np.zeros((len(tasks), words), np.uint32)
    for j in range(T):
        out[:, j // 32] |= improving[:, j].astype(np.uint32) << np.uint32(j % 32)
    return out


COPY_CELLS = 3348853  # int64 cells of the k-opt workload


def copy_calibration(launches: int) -> int:
    """Device-to-device copy of exactly COPY_CELLS int64 values, `launches`
    times: the reference kernel of known traffic (26,790,824 bytes read and
    written per launch) that scales the FETCH_SIZE / WRITE_SIZE counters."""
    import torch

    src = torch.arange(COPY_CELLS, dtype=torch.int64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(launches):
        dst.copy_(src)
    torch.cuda.synchronize()
    assert torch.equal(dst, src)
    print(json.dumps({"copy_bytes_per_launch": COPY_CELLS * 8, "launches": launches}), flush=True)
    return 0


def profile_figure(name: str, key: str):
    """A figure `key: value` from a committed rocprofv3 summary, or None."""
    path = ROOT / "profiles" / name
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        if line.startswith(key + ":"):
            return float(line.split(":", 1)[1].strip())
    return None


def measured_traffic():
    """HBM bytes per k-opt kernel launch from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE, each scaled by the copy calibration), or None."""
    v = profile_figure("kopt_kernel_config3_traffic.md", "traffic_bytes_per_launch")
    return None if v is None else int(v)


def cursor_leg(ctx, graph, blocks: int = 40) -> dict:
    """The reference's leaf search walks a tour context's deletion sets in
    blocks of cost_batch_size (4096) rows, one synchronous evaluation per
    block.  Times k = 5 blocks three ways: rows uploaded to the plain path
    (ee_kopt_eval_costs, the reference-shaped call; the rows are made
    beforehand and not timed), rows built on the device with the cells
    downloaded, and rows built on the device with only the improving-template
    masks downloaded."""
    tour = W.tour_context(0xEE0000C3, graph, 400, 120)
    table = backend.build_kopt_reconnect_table(5)
    total = backend.kopt_cursor_count(tour, 5)
    rows = 4096
    firsts = [i * rows for i in range(blocks)]
    pre = [backend.evaluate_kopt_cursor_cuda(graph, None, tour, 5, f, rows, output=None, ctx=ctx).tasks
           for f in firsts]
    cells = rows * len(table.templates)

    # C-ABI calls on preallocated buffers, so the figures are the library's
    from cudaedgeelimination_amd import _lib
    lib = _lib.lib()
    T = len(table.templates)
    templates = np.ascontiguousarray(table.templates, dtype=_lib.MATCHING_DTYPE)
    costs = np.empty(rows * T, dtype=np.int64)
    masks = np.empty(rows * ((T + 31) // 32), dtype=np.uint32)
    view = tour.view()
    usage, produced = _lib.KoptUsage(), C.c_uint32()
    gx, gy = _lib.ptr(graph.x), _lib.ptr(graph.y)

    def plain_call(f, t):
        _lib.check(lib.ee_kopt_eval_costs(ctx.handle, gx, gy, graph.dimension, graph.distance_type, 5,
                                          _lib.ptr(templates), T, table.generator_hash, _lib.ptr(t),
                                          len(t), _lib.ptr(costs), C.byref(usage)))

    def cursor_call(f, t, out_costs, out_masks):
        _lib.check(lib.ee_kopt_cursor_eval(ctx.handle, gx, gy, graph.dimension, graph.distance_type,
                                           C.byref(view), 5, _lib.ptr(templates), T,
                                           table.generator_hash, f, rows, None, None, out_costs,
                                           out_masks, C.byref(produced), C.byref(usage)))

    def timed(call):
        call(firsts[0], pre[0])  # warm the caches and workspaces
        parts = np.zeros(3)
        t0 = time.perf_counter()
        for f, tasks in zip(firsts, pre):
            call(f, tasks)
            parts += (usage.h2d_ms, usage.kernel_ms, usage.d2h_ms)
        wall = time.perf_counter() - t0
        return {"ms_per_block": wall * 1e3 / blocks, "cells_per_s": cells * blocks / wall,
                "sets_per_s": rows * blocks / wall, "rows_ms": parts[0] / blocks,
                "kernel_ms": parts[1] / blocks, "download_ms": parts[2] / blocks}

    plain = timed(plain_call)
    dev_costs = timed(lambda f, t: cursor_call(f, t, _lib.ptr(costs), None))
    dev_masks = timed(lambda f, t: cursor_call(f, t, None, _lib.ptr(masks)))
    check = backend.evaluate_kopt_cursor_cuda(graph, table, tour, 5, firsts[-1], rows, "costs", ctx)
    same = bool(np.array_equal(
        check.costs, backend.evaluate_kopt_template_costs_cuda(graph, table, pre[-1], ctx)[0]))
    return {"workload": "tour of 400 nodes of the n=15000 graph, 120 selectable positions, "
                        f"k=5: {total} deletion sets, {blocks} blocks of {rows} rows",
            # rows_ms: the rows' upload (plain) or enqueueing their kernel;
            # kernel_ms: launch to finish (the row kernel included)
            "plain_uploaded_rows": plain, "device_rows_costs": dev_costs,
            "device_rows_masks": dev_masks, "cells_equal_plain": same,
            "masks_speedup_over_plain": plain["ms_per_block"] / dev_masks["ms_per_block"]}


def certified_leaf_leg(ctx, wl, block_rows: int = 4096, depth: int = 3, reps: int = 5) -> dict:
    """The leaf search's certified cost loop (src/cpu/kopt_search.cpp:2651-2764)
    over the 65,536-row workload: blocks of `block_rows` rows per k (the
    reference's cost_batch_size), each evaluated on the device and certified
    on the host against the exact CPU matrix (every cell recomputed and
    compared) before its improving masks reach the consumer, with `depth`
    blocks in flight on the device.  Also the cursor form: the deletion sets of
    a 400-node tour built on the device."""
    blocks = [(k, wl.bucket(k)[s:s + block_rows]) for k in (3, 4, 5)
              for s in range(0, len(wl.bucket(k)), block_rows)]
    cells = sum(len(rows) * TEMPLATES[k] for k, rows in blocks)
    p = LeafPipeline(ctx, wl.graph, depth=depth, max_rows=block_rows)

    def run_host():
        submitted = improving = 0
        for i in range(len(blocks)):
            while submitted < len(blocks) and submitted < i + depth:
                k, rows = blocks[submitted]
                p.submit_tasks(k, rows, tag=submitted)
                submitted += 1
            improving += p.next(copy=False).improving_cells
        return improving

    def stats_delta(fn):
        before = p.stats()
        t0 = time.perf_counter()
        value = fn()
        wall = time.perf_counter() - t0
        after = p.stats()
        return wall, value, {key: after[key] - before[key] for key in
                             ("blocks", "cells", "kernel_ms", "copy_ms", "wait_ms", "certify_ms")}

    run_host()  # warm-up: host pool threads, pattern caches, pinned pages
    runs = sorted((stats_delta(run_host) for _ in range(reps)), key=lambda r: r[0])
    wall, improving, st = runs[len(runs) // 2]
    out = {
        "workload": f"{len(blocks)} blocks of <= {block_rows} rows per k over the 65536-row workload",
        "cells": cells, "blocks": len(blocks), "depth": depth,
        "certified_cells_per_s": cells / wall, "ms": wall * 1e3,
        "improving_cells": improving,
        # per run: device time per block summed (rows + launch; downloads),
        # host time waiting on the device, host time certifying
        "device_kernel_ms": st["kernel_ms"], "device_copy_ms": st["copy_ms"],
        "host_wait_ms": st["wait_ms"], "certify_ms": st["certify_ms"],
        "certify_threads": p.stats()["certify_threads"],
    }

    # cursor blocks: the first 40 blocks of k = 5 sets of a 400-node tour
    t = W.tour_context(5, wl.graph, 400, 120)
    p.set_tour(t)
    n_cursor = 40

    def run_cursor():
        submitted = improving = 0
        for i in range(n_cursor):
            while submitted < n_cursor and submitted < i + depth:
                p.submit_cursor(5, submitted * block_rows, block_rows, tag=submitted)
                submitted += 1
            improving += p.next(copy=False).improving_cells
        return improving

    run_cursor()
    runs = sorted((stats_delta(run_cursor) for _ in range(reps)), key=lambda r: r[0])
    wall, improving, st = runs[len(runs) // 2]
    c_cells = n_cursor * block_rows * TEMPLATES[5]
    out["cursor_rows"] = {
        "workload": f"{n_cursor} blocks of {block_rows} k=5 deletion sets of a 400-node tour, "
                    "rows built on the device and rebuilt on the host",
        "cells": c_cells, "certified_cells_per_s": c_cells / wall, "ms": wall * 1e3,
        "improving_cells": improving, "device_kernel_ms": st["kernel_ms"],
        "device_copy_ms": st["copy_ms"], "host_wait_ms": st["wait_ms"], "certify_ms": st["certify_ms"],
    }
    p.close()
    return out


def ht_leg(ctx, cpu: bool, reps: int = 20) -> dict:
    """HT Hamilton replies of one target edge against every other node of the
    30,000-node served graph (10-NN), and end replies of 30,000 tasks: device
    time (count, scan and write kernels) and whole calls, with the reference
    CPU path (EvaluateHtHamiltonReplies, CPU backend, oracle/_ref) beside it."""
    g = W.config5_graphs(sizes=(30000,))[0]
    e = int(np.flatnonzero(g.edge_active)[12345])
    target = (int(g.edge_u[e]), int(g.edge_v[e]))
    centers = np.setdiff1d(np.arange(g.dimension), target).astype(np.int32)
    deg = np.diff(g.row_offsets).astype(np.int64)
    pairs = int((deg[centers] * (deg[centers] - 1) // 2).sum())
    # exact distances: 5 per neighbour (center edge, two 2-opt orientations),
    # 1 per surviving-candidate pair, 3 per center
    dists = int(5 * deg[centers].sum() + pairs + 3 * len(centers))
    backend.evaluate_ht_hamilton_replies_cuda(g, target, centers, ctx)  # warm caches
    kernel, wall = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        off, rep, u = backend.evaluate_ht_hamilton_replies_cuda(g, target, centers, ctx)
        wall.append(time.perf_counter() - t0)
        kernel.append(u.kernel_ms)
    rng = np.random.default_rng(0xEE0000E1)
    act = np.flatnonzero(g.edge_active)
    pick = rng.choice(act, size=30000)
    tasks = np.stack([g.edge_u[pick], g.edge_v[pick]], axis=1).astype(np.int32)
    backend.evaluate_ht_end_replies_cuda(g, tasks, ctx)
    end_kernel, end_wall = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        eoff, erep, eu = backend.evaluate_ht_end_replies_cuda(g, tasks, ctx)
        end_wall.append(time.perf_counter() - t0)
        end_kernel.append(eu.kernel_ms)
    k_ms = statistics.median(kernel)
    out = {"workload": f"target edge {target} of the n=30000 10-NN graph, {len(centers)} centers, "
                       f"{pairs} neighbour pairs; 30000 end tasks",
           "replies": int(len(rep)), "kernel_ms": k_ms, "call_ms": statistics.median(wall) * 1e3,
           "pairs_per_s": pairs / (k_ms * 1e-3), "distance_evals_per_s": dists / (k_ms * 1e-3),
           "end_replies": int(len(erep)), "end_kernel_ms": statistics.median(end_kernel),
           "end_call_ms": statistics.median(end_wall) * 1e3}
    if cpu:
        import _oracle as O
        if O.ref_available():
            ref_args = (g.x, g.y, g.distance_type, g.edge_u, g.edge_v, g.edge_weight, g.edge_active)
            times = []
            for _ in range(5):
                t0 = time.perf_counter()
                roff, rrep = O.ref_ht_hamilton_replies(*ref_args, target, centers, pairs)
                times.append(time.perf_counter() - t0)
            out["cpu_baseline"] = {"ms": statistics.median(times) * 1e3, "kind": "reference",
                                   "cores": 1, "sample": "the same call, five runs (median)",
                                   "replies_equal": bool(np.array_equal(roff, off)
                                                         and np.array_equal(rrep, rep))}
    return out


def main() -> int:
    # the reference protocol for its OpenMP CPU path (set before it loads)
    for key, value in (("OMP_DYNAMIC", "FALSE"), ("OMP_PROC_BIND", "spread"),
                       ("OMP_PLACES", "cores")):
        os.environ.setdefault(key, value)
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # a step is one pass over the batch (about 13 us of GPU time): enough of
    # them that the timed call's fixed cost (event wait, host call) stays
    # below one percent of the timed region
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-jv", action="store_true", help="skip the JV workload")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-serve", action="store_true", help="skip the serving workload")
    ap.add_argument("--no-many", action="store_true", help="skip the many-instance workload")
    ap.add_argument("--no-cursor", action="store_true", help="skip the deletion-set cursor leg")
    ap.add_argument("--no-ht", action="store_true", help="skip the HT reply leg")
    ap.add_argument("--no-leaf", action="store_true", help="skip the certified leaf leg")
    ap.add_argument("--serve-qps", default="2000,8000,20000,0",
                    help="Poisson rates of the serving leg; 0 is a burst of --serve-burst requests")
    ap.add_argument("--serve-burst", type=int, default=20000)
    ap.add_argument("--saturation-qps", default="10000,14000,20000,28000,40000,56000,80000",
                    help="ascending Poisson rates of the saturation search (empty: none)")
    ap.add_argument("--saturation-p99-ms", type=float, default=0.5)
    ap.add_argument("--serve-seconds", type=float, default=1.0)
    ap.add_argument("--tuning", default=None,
                    help="r3,r4,r5,grid,block,nt: skip the autotuner and use this tuning")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device copy of the cost-matrix size (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the path is sharded by row range across GPUs, see DESIGN.md)", file=sys.stderr)
        return 2

    if args.copy_calibration:
        return copy_calibration(args.steps)

    log("building the k-opt workload (n=15000, 65536 rows)")
    wl = W.config3()
    ctx = backend.Context(0)
    batch = KoptBatch(ctx)
    inst = batch.add_instance(wl.graph)
    segments = [(inst, k, int((wl.ks == k).sum())) for k in (3, 4, 5)]
    batch.set_work(segments, np.concatenate([wl.bucket(k) for k in (3, 4, 5)]))
    if args.tuning:
        v = [int(t) for t in args.tuning.split(",")]
        batch.set_tuning(v[:3], v[3], v[4], v[5])
        serial_tuning = batch.tuning()
    else:
        log("autotuning (single-stream launches)")
        serial_tuning = batch.autotune(10)
    batch.run(max(args.warmup, 1))
    log(f"timing {args.steps} steps")
    ctx_sync = backend._lib.lib().ee_ctx_synchronize
    ctx_sync(ctx.handle)
    # serial launches on one stream at their own best shape: the kernel's own
    # duration (roofline)
    t0 = time.perf_counter()
    device_ms = batch.run(args.steps)
    ctx_sync(ctx.handle)
    wall = time.perf_counter() - t0
    graph_ms = batch.run(args.steps, use_graph=True)
    # the headline: launches dealt over four streams and four cost pools,
    # each still a full pass over the batch, so a launch's start (dependent
    # gathers) overlaps the previous ones' drain
    batch.set_pipeline_depth(PIPELINE_DEPTH)
    if args.tuning:
        tuning = serial_tuning
    else:
        log("autotuning (pipelined runs)")
        tuning = batch.autotune(40, pipelined=True)  # enough launches to fill the pipeline
    batch.run(max(args.warmup, 1), pipelined=True)
    ctx_sync(ctx.handle)
    t0 = time.perf_counter()
    pipelined_ms = batch.run(args.steps, pipelined=True)
    ctx_sync(ctx.handle)
    wall_pipelined = time.perf_counter() - t0
    # evidence of the overlap, from the device clock: every wavefront stamps
    # its start and end, so each launch spans its first start to its last
    # end; the same shape alone on one stream beside it
    n_stamped = min(max(args.steps, 64), 512)
    s0, e0 = batch.run_stamped(n_stamped, pipelined=False)
    s1, e1 = batch.run_stamped(n_stamped, pipelined=True)
    overlap = {
        "launches": n_stamped,
        "alone_median_launch_us": float(np.median(e0 - s0)),
        "alone_wall_us_per_launch": float((e0.max() - s0.min()) / n_stamped),
        "pipelined_median_launch_us": float(np.median(e1 - s1)),
        "pipelined_wall_us_per_launch": float((e1.max() - s1.min()) / n_stamped),
        # summed launch spans over the run's span: launches running at once
        "launches_in_flight": float((e1 - s1).sum() / (e1.max() - s1.min())),
        "fraction_starting_before_previous_ends": float(np.mean(s1[1:] < e1[:-1])),
        "clock": "s_memrealtime (constant rate, hipDeviceAttributeWallClockRate)",
    }
    cells = batch.cells
    ms_per_step = wall_pipelined * 1e3 / args.steps
    value = cells / (wall_pipelined / args.steps)
    gpu_out = {k: batch.segment_costs(i) for i, k in enumerate((3, 4, 5))}
    # every pool the pipelined launches wrote, compared with the CPU below
    pool_sha = [{k: sha(batch.segment_costs(i, pool)) for i, k in enumerate((3, 4, 5))}
                for pool in range(min(PIPELINE_DEPTH, args.steps))]
    algo_bytes = kopt_algorithmic_bytes(wl)
    achieved = algo_bytes / (device_ms * 1e-3) / 1e9
    pipelined_achieved = algo_bytes / (pipelined_ms * 1e-3) / 1e9
    dist_evals, per_cell_evals = kopt_distance_evals(wl)
    # consumer-side output: per row the mask of improving templates
    # (cell < deleted_cost) instead of the cells, evaluated and downloaded
    batch.set_output("masks")
    if args.tuning:
        batch.set_tuning(serial_tuning["rows_per_tile"], serial_tuning["grid"], serial_tuning["block"],
                         serial_tuning["non_temporal"])
        mask_tuning = batch.tuning()
    else:
        log("autotuning (improving-template masks)")
        mask_tuning = batch.autotune(10)  # its own store pattern, tuned on its own
    batch.run(args.warmup)
    mask_ms = batch.run(args.steps)
    t_dl = time.perf_counter()
    gpu_masks = batch.download_masks()
    mask_dl_s = time.perf_counter() - t_dl
    batch.set_output("costs")
    masks = {
        "words_per_row": int(gpu_masks.shape[1]), "kernel_ms": mask_ms, "tuning": mask_tuning,
        # against the cells' own single-stream launch
        "kernel_vs_cells": mask_ms / device_ms,
        "download_ms": mask_dl_s * 1e3, "bytes": int(gpu_masks.nbytes),
        "improving_cells": int(np.unpackbits(gpu_masks.view(np.uint8)).sum()),
        "cells_download_ms": None,
    }

    log("plain per-k path (synchronous upload, launch, download per k)")
    plain_samples = []
    for _ in range(5):
        t1 = time.perf_counter()
        for k in (3, 4, 5):
            backend.evaluate_kopt_template_costs_cuda(
                wl.graph, backend.build_kopt_reconnect_table(k), wl.bucket(k), ctx)
        plain_samples.append(time.perf_counter() - t1)
    plain_s = statistics.median(plain_samples)

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {
            "workload": "kopt-leaf-cost-matrix n=15000 uniform [0,1e6]^2 8-NN, 65536 rows "
                        "k3/k4/k5=32768/19661/13107, 3348853 int64 cells, EUC_2D",
            "rows": int(len(wl.tasks)), "cells": cells, "nodes": wl.graph.dimension,
            "tuning": tuning, "launches_per_step": 1,
            "streams": PIPELINE_DEPTH, "output_pools": PIPELINE_DEPTH,
        },
        # the same steps launched back to back on one stream, at the best
        # single-stream shape
        "serial": {"value": cells / (wall / args.steps), "ms_per_step": wall * 1e3 / args.steps,
                   "kernel_ms": device_ms, "tuning": serial_tuning},
        "pipelined_overlap": overlap,
        # the kernel's own duration comes from the serial launches: rocprof's
        # per-kernel durations, which the committed summary must match, are
        # defined only for launches that do not overlap.  The pipelined
        # region's share per launch and the HBM rate it implies sit beside it.
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
            # SQ_INSTS_VALU x 4 cycles over every SIMD's cycles per launch
            "valu_utilisation": profile_figure("kopt_kernel_config3_valu.md", "valu_utilisation"),
            "kernel": "KoptKernelResident", "algorithmic_bytes_per_launch": algo_bytes,
            "kernel_ms": device_ms, "graph_replay_kernel_ms": graph_ms,
            "pipelined_ms_per_launch": pipelined_ms, "pipelined_streams": PIPELINE_DEPTH,
            "pipelined_achieved": pipelined_achieved,
            "pipelined_frac": pipelined_achieved / HBM_PEAK_GBS,
            # exact distances per second: evaluated once per (row, port pair),
            # and as a per-cell evaluation (k per cell) would count them
            "distance_evals_per_s": dist_evals / (device_ms * 1e-3),
            "per_cell_equivalent_distance_evals_per_s": per_cell_evals / (device_ms * 1e-3),
        },
        "plain_path": {"cells_per_s": cells / plain_s, "ms": plain_s * 1e3,
                       "batched_speedup": value / (cells / plain_s)},
        "improving_masks": masks,
    }
    # download time of the full matrix for comparison (all three segments)
    t_dl = time.perf_counter()
    batch.download()
    masks["cells_download_ms"] = (time.perf_counter() - t_dl) * 1e3

    if not args.no_cpu:
        log(f"CPU baseline ({args.cpu_seconds:.0f} s budget)")
        cpu, cpu_out = cpu_baseline_kopt(wl, args.cpu_seconds)
        result["cpu_baseline"] = cpu
        want_masks = np.concatenate([pack_improving(cpu_out[k], wl.bucket(k), TEMPLATES[k], masks["words_per_row"])
                                     for k in (3, 4, 5)])
        masks["equal_cpu"] = bool(np.array_equal(gpu_masks, want_masks))
        cpu_sha = {k: sha(cpu_out[k]) for k in (3, 4, 5)}
        result["parity"] = {
            "kopt_cells_equal_cpu": all(np.array_equal(gpu_out[k], cpu_out[k]) for k in (3, 4, 5)),
            # each pool of the pipelined (headline) launches, by SHA-256
            "kopt_pools_checked": len(pool_sha),
            "kopt_pools_equal_cpu": all(p[k] == cpu_sha[k] for p in pool_sha for k in (3, 4, 5)),
        }
    else:
        result["cpu_baseline"] = None

    if not args.no_jv:
        log("JV multi-epoch workload (60x60 lattice, 10-NN)")
        graph = W.config2()
        runs = []
        for _ in range(3):
            session = JvSession(ctx, graph)
            runs.append(session.run(100))
            session.close()
        best = min(runs, key=lambda r: r.total_ms)
        jv = {
            "workload": f"drilling-grid lattice n={graph.dimension} m={graph.edge_count} 10-NN",
            "epochs": best.epochs, "edges_scanned": best.edges_scanned,
            "committed": best.committed,
            "kernel_edges_per_s": best.edges_scanned / (best.device_ms * 1e-3),
            "propose_edges_per_s": best.edges_scanned / (best.propose_ms * 1e-3),
            "algorithm_edges_per_s": best.edges_scanned / (best.total_ms * 1e-3),
            "kernel_ms": best.device_ms, "total_ms": best.total_ms,
            # where the rest of the run goes (host side of each epoch)
            "h2d_ms": best.h2d_ms, "d2h_ms": best.d2h_ms, "propose_ms": best.propose_ms,
            "verify_ms": best.verify_ms, "commit_ms": best.commit_ms,
            # device commit (deactivate + CSR compaction) and the end of the
            # run's wait for the helper thread's snapshot hashes
            "device_commit_ms": best.compact_ms, "hash_wait_ms": best.hash_wait_ms,
            "final_hash": f"{best.final_hash:016x}",
        }
        # every resident byte touched once per epoch: 24 B per edge slot (u,
        # v, weight, active, witness), 20 B per node + 4 (row offsets,
        # coordinates), 8 B per adjacency entry (2 per active edge)
        jv_bytes = best.epochs * (24 * graph.edge_count + 20 * graph.dimension + 4) + 16 * best.edges_scanned
        jv_achieved = jv_bytes / (best.device_ms * 1e-3) / 1e9
        jv["roofline"] = {
            "bound": "hbm", "achieved": jv_achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": jv_achieved / HBM_PEAK_GBS, "kernel": "JvKernel",
            "algorithmic_bytes": jv_bytes, "kernel_us_per_epoch": best.device_ms * 1e3 / best.epochs,
            # from the committed rocprof summary of the same workload
            "valu_utilisation": profile_figure("jv_kernel_lattice.md", "valu_utilisation"),
            "salu_per_edge": profile_figure("jv_kernel_lattice.md", "salu_per_edge"),
            "valu_per_edge": profile_figure("jv_kernel_lattice.md", "valu_per_edge"),
        }
        # where a run's time goes at the scale of a 15,000-node instance
        # (uniform points, 10-NN): device proposal and commit against the
        # host's re-verification, commit rule and snapshot hashes
        big = W.config5_graphs(sizes=(15000,))[0]
        split_runs = []
        for _ in range(3):
            session = JvSession(ctx, big)
            split_runs.append(session.run(100))
            session.close()
        sb = min(split_runs, key=lambda r: r.total_ms)
        jv["split_uniform_15000"] = {
            "workload": f"uniform n={big.dimension} m={big.edge_count} 10-NN",
            "epochs": sb.epochs, "edges_scanned": sb.edges_scanned, "committed": sb.committed,
            "total_ms": sb.total_ms, "algorithm_edges_per_s": sb.edges_scanned / (sb.total_ms * 1e-3),
            "first_upload_ms": sb.h2d_ms, "kernel_ms": sb.device_ms,
            "device_commit_ms": sb.compact_ms, "propose_ms": sb.propose_ms,
            "verify_ms": sb.verify_ms, "commit_ms": sb.commit_ms, "hash_wait_ms": sb.hash_wait_ms,
        }
        if not args.no_cpu:
            import _oracle as O

            cpu_jv, text = cpu_baseline_jv(graph)
            jv["cpu_baseline"] = cpu_jv
            jv["proof_equals_cpu"] = O.canonical_proof(text) == O.canonical_proof(best.proof)
            cpu_big, big_text = cpu_baseline_jv(big)
            jv["split_uniform_15000"]["cpu_baseline"] = cpu_big
            jv["split_uniform_15000"]["proof_equals_cpu"] = (O.canonical_proof(big_text)
                                                             == O.canonical_proof(sb.proof))
        result["jv"] = jv

    if not args.no_leaf:
        log("certified leaf pipeline (device blocks ahead of host certification)")
        leaf = certified_leaf_leg(ctx, wl)
        # the reference's synchronous shape of the same certified loop: the
        # plain device call, then its CPU matrix of the same rows, per block
        # (plain_path.ms + the CPU baseline's matrix time, summed over k)
        if result.get("cpu_baseline"):
            cpu_s = wl.cells() / result["cpu_baseline"]["value"]
            leaf["reference_synchronous_cells_per_s"] = wl.cells() / (plain_s + cpu_s)
            leaf["reference_cpu_matrix_cells_per_s"] = result["cpu_baseline"]["value"]
            leaf["speedup_over_reference_cpu_matrix"] = (leaf["certified_cells_per_s"]
                                                         / result["cpu_baseline"]["value"])
        result["certified_leaf"] = leaf

    if not args.no_ht:
        log("HT replies (30,000 centers of one target)")
        result["ht_replies"] = ht_leg(ctx, cpu=not args.no_cpu)

    if not args.no_cursor:
        log("deletion-set cursor blocks (device-built rows)")
        result["cursor"] = cursor_leg(ctx, wl.graph)

    if not args.no_many:
        log("many-instance workload (256 instances: JV fleet + k-opt batch)")
        result["many_instance"] = many_instance_leg(ctx, cpu=not args.no_cpu)

    if not args.no_serve:
        log("serving workload (open-loop Poisson queries)")
        result["serving"] = serving_leg(
            ctx, [int(q) for q in args.serve_qps.split(",")], args.serve_seconds, burst=args.serve_burst,
            saturation_p99_ms=args.saturation_p99_ms,
            saturation_rates=[int(q) for q in args.saturation_qps.split(",") if q])

    batch.close()
    ctx.close()
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
