"""Leaf-cost queries and JV rechecks through the device work queue: a
persistent kernel polls a request ring in coherent host memory and raises a
completion flag per request.  Every request's cells and witnesses equal the
CPU oracle's and the batched server's, alone and in the mix, with concurrent
submitters, with arenas and a ring small enough that submits wait for room,
across the kernel's idle exit and relaunch, and with graphs added while
serving."""
from __future__ import annotations

import sys
import threading
import time

import numpy as np
import pytest

import _oracle as O
from cudaedgeelimination_amd import _lib, workloads as W
from cudaedgeelimination_amd.resident import DeviceQueue, KoptServer, TEMPLATES_PER_K

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def graphs_and_rows():
    graphs, rows = [], []
    for i, n in enumerate((300, 1200, 2500)):
        wl = W.config3(seed=0xEE0005A0 + i, n=n, rows=600)
        graphs.append(wl.graph)
        rows.append({k: wl.bucket(k) for k in (3, 4, 5)})
    return graphs, rows


def make_requests(rows, count, seed, most=90):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        g = int(rng.integers(len(rows)))
        k = int(rng.choice([3, 4, 5], p=[0.5, 0.3, 0.2]))
        pool = rows[g][k]
        size = int(rng.integers(1, min(len(pool), most) + 1))
        start = int(rng.integers(0, len(pool) - size + 1))
        out.append((g, k, pool[start:start + size]))
    return out


def expected(graphs, g, k, tasks):
    graph = graphs[g]
    costs, _, _ = O.kopt_costs(graph.x, graph.y, graph.distance_type, k, tasks)
    return costs


def mismatch(got, want, k, info=None):
    """Where a result differs: wrong cells, the rows holding them, and the
    first wrong row's leading cells (got / want)."""
    T = TEMPLATES_PER_K[k]
    bad = np.nonzero(got != want)[0]
    rows = sorted(set((bad // T).tolist()))
    r = rows[0]
    return (f"{len(bad)}/{len(got)} cells wrong in rows {rows[:12]} (entry "
            f"{None if info is None else info['batch_id']}); row {r} got "
            f"{got[r * T:r * T + 4].tolist()} want {want[r * T:r * T + 4].tolist()}")


def test_queue_requests_equal_oracle_alone_and_in_the_mix(ctx, graphs_and_rows):
    """One request alone, then 300 in flight at once."""
    say("queue: alone, then 300 in flight")
    graphs, rows = graphs_and_rows
    queue = DeviceQueue(ctx)
    ids = [queue.add_graph(g) for g in graphs]
    t = rows[2][5][:17]
    costs, info = queue.wait(queue.submit(ids[2], 5, t), 60000)
    assert np.array_equal(costs, expected(graphs, 2, 5, t))
    assert info["latency_ms"] > 0
    reqs = make_requests(rows, 300, seed=1)
    submitted = [(queue.submit(ids[g], k, t), g, k, t) for g, k, t in reqs]
    for rid, g, k, t in submitted:
        costs, _ = queue.wait(rid, 60000)
        assert np.array_equal(costs, expected(graphs, g, k, t)), (g, k, len(t))
    stats = queue.stats()
    assert stats["requests"] == len(reqs) + 1
    assert stats["cells"] == sum(len(t) * TEMPLATES_PER_K[k] for _, k, t in reqs) + len(rows[2][5][:17]) * 208
    assert stats["launches"] >= 1
    queue.close()


def test_queue_concurrent_submitters(ctx, graphs_and_rows):
    say("queue: six concurrent submitters")
    graphs, rows = graphs_and_rows
    queue = DeviceQueue(ctx)
    ids = [queue.add_graph(g) for g in graphs]
    failures = []

    def client(seed):
        try:
            for g, k, t in make_requests(rows, 60, seed):
                costs, info = queue.wait(queue.submit(ids[g], k, t), 60000)
                want = expected(graphs, g, k, t)
                if not np.array_equal(costs, want):
                    failures.append((seed, g, k, len(t), mismatch(costs, want, k, info)))
        except Exception as e:  # noqa: BLE001
            failures.append(repr(e))

    threads = [threading.Thread(target=client, args=(s,)) for s in range(6)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(120)
    assert not failures
    assert queue.stats()["requests"] == 360
    queue.close()


def test_queue_waits_for_room(ctx, graphs_and_rows):
    """Arenas of 256 rows / 8,192 cells and a ring of 8 entries: submits
    wait for earlier requests (collecting them) and every result is right."""
    say("queue: small arenas and ring")
    graphs, rows = graphs_and_rows
    queue = DeviceQueue(ctx, task_rows=256, cost_cells=8192, ring_entries=8, workers=64)
    ids = [queue.add_graph(g) for g in graphs]
    reqs = make_requests(rows, 150, seed=3, most=39)  # k = 5: 39 rows = 8,112 cells at most
    submitted = [(queue.submit(ids[g], k, t), g, k, t) for g, k, t in reqs]
    for rid, g, k, t in submitted:
        costs, _ = queue.wait(rid, 60000)
        assert np.array_equal(costs, expected(graphs, g, k, t)), (g, k, len(t))
    assert queue.stats()["peak_inflight_cells"] <= 8192
    queue.close()


def test_queue_idle_exit_and_relaunch(ctx, graphs_and_rows):
    """The kernel ends after 2 ms without requests; the next submit starts
    it again."""
    say("queue: idle exit and relaunch")
    graphs, rows = graphs_and_rows
    queue = DeviceQueue(ctx, idle_exit_ms=2.0)
    g = queue.add_graph(graphs[1])
    for round_ in range(3):
        t = rows[1][4][round_ * 20:round_ * 20 + 20]
        costs, _ = queue.wait(queue.submit(g, 4, t), 60000)
        assert np.array_equal(costs, expected(graphs, 1, 4, t))
        time.sleep(0.05)
    assert queue.stats()["launches"] >= 3
    queue.close()


def test_queue_graph_added_while_serving(ctx, graphs_and_rows):
    graphs, rows = graphs_and_rows
    queue = DeviceQueue(ctx)
    g0 = queue.add_graph(graphs[0])
    rid = queue.submit(g0, 5, rows[0][5][:10])
    g1 = queue.add_graph(graphs[1])
    rid1 = queue.submit(g1, 3, rows[1][3][:10])
    c0, _ = queue.wait(rid)
    c1, _ = queue.wait(rid1)
    assert np.array_equal(c0, expected(graphs, 0, 5, rows[0][5][:10]))
    assert np.array_equal(c1, expected(graphs, 1, 3, rows[1][3][:10]))
    queue.close()


def test_queue_rejects_bad_requests(ctx, graphs_and_rows):
    graphs, rows = graphs_and_rows
    queue = DeviceQueue(ctx, task_rows=64, cost_cells=4096, max_nodes=2000)
    g = queue.add_graph(graphs[0])
    with pytest.raises(_lib.InvalidArgument):
        queue.submit(g, 5, rows[0][5][:40])  # 8,320 cells > the 4,096-cell arena
    bad = rows[0][3][:2].copy()
    bad["port_nodes"][0, 0] = graphs[0].dimension
    with pytest.raises(_lib.InvalidArgument):
        queue.submit(g, 3, bad)
    with pytest.raises(_lib.InvalidArgument):
        queue.submit(g + 7, 3, rows[0][3][:2])
    with pytest.raises(_lib.InvalidArgument):
        queue.add_graph(graphs[2])  # 300 + 2,500 nodes exceed the 2,000-node pool
    with pytest.raises(_lib.InvalidArgument):
        queue.submit_many([(g, 3, rows[0][3][:2]), (g + 9, 3, rows[0][3][:2])])
    rid = queue.submit(g, 3, rows[0][3][:4])
    costs, _ = queue.wait(rid)
    assert np.array_equal(costs, expected(graphs, 0, 3, rows[0][3][:4]))
    queue.close()


def test_queue_refused_entry_fails_closed(ctx, graphs_and_rows):
    """An entry pointing past the task arena (forced from the host) is
    refused by the kernel's bounds check: its wait raises, requests served
    before it are intact, and the queue takes no more work."""
    graphs, rows = graphs_and_rows
    queue = DeviceQueue(ctx)
    g = queue.add_graph(graphs[0])
    t = rows[0][3][:20]
    costs, _ = queue.wait(queue.submit(g, 3, t), 60000)
    assert np.array_equal(costs, expected(graphs, 0, 3, t))
    bad = queue.inject_refusal(g, rows[0][3][:1])
    with pytest.raises(_lib.EeError, match="refused"):
        queue.wait(bad, 60000)
    with pytest.raises(_lib.EeError, match="refused"):
        queue.submit(g, 3, t)
    queue.close()


def test_queue_jv_rechecks_mixed_with_leaf_costs(ctx, graphs_and_rows):
    """JV rechecks of edge-slot subsets on resident snapshots (EUC_2D and
    CEIL_2D, some edges inactive), mixed with leaf-cost queries and submitted
    one by one: every witness equals the oracle's on that snapshot, every
    cost row the oracle's; odd counts end in a half-filled pair."""
    say("queue: JV rechecks mixed with leaf costs")
    graphs, rows = graphs_and_rows
    ceil = W.config1(seed=0xEE0005C1, n=500)
    ceil = W.knn_graph(ceil.x, ceil.y, 1, 12)
    sparse = graphs[1].copy()
    rng = np.random.default_rng(7)
    sparse.edge_active[rng.random(sparse.edge_count) < 0.1] = 0
    sparse.rebuild_csr()
    jv_graphs = [graphs[0], ceil, sparse]
    want = [O.OGraph.from_snapshot(g).witnesses() for g in jv_graphs]
    queue = DeviceQueue(ctx)
    kids = [queue.add_graph(g) for g in graphs]
    jids = [queue.add_jv_graph(g) for g in jv_graphs]
    submitted = []
    kopt_reqs = make_requests(rows, 60, seed=5)
    for i in range(120):
        if i % 2 == 0:
            j = int(rng.integers(len(jv_graphs)))
            m = jv_graphs[j].edge_count
            size = 1 if i == 0 else int(rng.integers(1, min(m, 700)))
            edges = rng.choice(m, size=size, replace=False).astype(np.int32)
            submitted.append(("jv", queue.submit_jv(jids[j], edges), j, edges))
        else:
            g, k, t = kopt_reqs[i // 2]
            submitted.append(("kopt", queue.submit(kids[g], k, t), g, (k, t)))
    total_edges = 0
    for kind, rid, g, payload in submitted:
        out, _ = queue.wait(rid, 60000)
        if kind == "jv":
            assert np.array_equal(out, want[g][payload]), (g, len(payload))
            total_edges += len(payload)
        else:
            k, t = payload
            assert np.array_equal(out, expected(graphs, g, k, t))
    assert queue.stats()["jv_edges"] == total_edges
    assert any((w >= 0).any() for w in want)
    with pytest.raises(_lib.InvalidArgument):
        queue.submit_jv(jids[0], np.array([jv_graphs[0].edge_count], np.int32))
    with pytest.raises(_lib.InvalidArgument):
        queue.submit_jv(len(jv_graphs) + 3, np.array([0], np.int32))
    queue.close()


def test_queue_jv_graph_added_while_serving(ctx, graphs_and_rows):
    """A JV graph added while requests are in flight: the running launch
    serves them and ends, the next submit launches again, and every result
    (before and after) is the oracle's."""
    graphs, rows = graphs_and_rows
    queue = DeviceQueue(ctx)
    g0 = queue.add_graph(graphs[0])
    j0 = queue.add_jv_graph(graphs[0])
    first = [queue.submit(g0, 4, rows[0][4][:30]), queue.submit_jv(j0, np.arange(50, dtype=np.int32))]
    launches = queue.stats()["launches"]
    j1 = queue.add_jv_graph(graphs[2])
    edges = np.arange(graphs[2].edge_count - 77, graphs[2].edge_count, dtype=np.int32)
    later = queue.submit_jv(j1, edges)
    c, _ = queue.wait(first[0])
    w0, _ = queue.wait(first[1])
    w1, _ = queue.wait(later)
    assert np.array_equal(c, expected(graphs, 0, 4, rows[0][4][:30]))
    assert np.array_equal(w0, O.OGraph.from_snapshot(graphs[0]).witnesses()[:50])
    assert np.array_equal(w1, O.OGraph.from_snapshot(graphs[2]).witnesses()[edges])
    assert queue.stats()["launches"] > launches
    queue.close()


def test_queue_served_graphs_at_full_size(ctx):
    """The served workload at full size (graphs of 1k, 5k, 15k and 30k nodes;
    leaf-cost queries of 128 to 65,728 cells; JV rechecks of 16 to 4,096 edge
    slots) queued in one call: every result equals the oracle's and the
    batched server's."""
    say("queue at full size: graphs 1k/5k/15k/30k")
    graphs = W.config5_graphs()
    pools = W.config5_row_pools(graphs)
    rng = np.random.default_rng(0xEE0005F6)
    requests, _ = W.config5_requests(graphs, pools, 500, rng)
    requests.append((3, 5, pools[3][5][:316]))
    requests.append((2, 4, pools[2][4][:2621]))
    assert any(k == 0 for _, k, _ in requests)
    want_w = [O.OGraph.from_snapshot(g).witnesses() for g in graphs]
    want = [want_w[g][t] if k == 0 else expected(graphs, g, k, t) for g, k, t in requests]
    queue = DeviceQueue(ctx)
    qids = [queue.add_graph(g) for g in graphs]
    jids = [queue.add_jv_graph(g) for g in graphs]
    outs = [o for o, _ in queue.wait_many(
        queue.submit_many([(qids[g], k, t) if k else (jids[g], 0, t) for g, k, t in requests]), 120000)]
    bad = [i for i, (o, w) in enumerate(zip(outs, want)) if not np.array_equal(o, w)]
    assert not bad, f"queue: requests differ from the oracle: {bad[:10]}"
    queue.close()
    server = KoptServer(ctx, batching=True, max_batch_cells=32768)
    sids = [server.add_graph(g) for g in graphs]
    sjids = [server.add_jv_graph(g) for g in graphs]
    souts = [o for o, _ in server.wait_many(
        server.submit_many([(sids[g], k, t) if k else (sjids[g], 0, t) for g, k, t in requests]), 120000)]
    server.close()
    for a, b in zip(outs, souts):
        assert np.array_equal(a, b)
