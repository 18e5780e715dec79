"""Serving loop on the device: every request's cells equal the same request
served alone (plain per-query server) and the CPU oracle, under continuous
batching with concurrent submitters; admission limits are enforced."""
from __future__ import annotations

import sys
import threading

import numpy as np
import pytest

import _oracle as O
from cudaedgeelimination_amd import _lib, workloads as W
from cudaedgeelimination_amd.resident import KoptServer, TEMPLATES_PER_K

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


def make_requests(rows, count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        g = int(rng.integers(len(rows)))
        k = int(rng.choice([3, 4, 5], p=[0.5, 0.3, 0.2]))
        pool = rows[g][k]
        size = int(rng.integers(1, min(len(pool), 90) + 1))
        start = int(rng.integers(0, len(pool) - size + 1))
        out.append((g, k, pool[start:start + size]))
    return out


def expected(graphs, g, k, tasks):
    graph = graphs[g]
    costs, _, _ = O.kopt_costs(graph.x, graph.y, graph.distance_type, k, tasks)
    return costs


@pytest.mark.parametrize("batching", [True, False])
def test_requests_equal_oracle_alone_and_in_the_mix(ctx, graphs_and_rows, batching):
    say(f"serving batching={batching}")
    graphs, rows = graphs_and_rows
    server = KoptServer(ctx, batching=batching, max_batch_cells=40000)
    ids = [server.add_graph(g) for g in graphs]
    reqs = make_requests(rows, 300, seed=1)
    submitted = [(server.submit(ids[g], k, t), g, k, t) for g, k, t in reqs]
    sizes = []
    for rid, g, k, t in submitted:
        costs, info = server.wait(rid, 60000)
        sizes.append(info["batch_size"])
        assert np.array_equal(costs, expected(graphs, g, k, t)), (g, k, len(t))
    stats = server.stats()
    assert stats["requests"] == len(reqs)
    if batching:
        assert max(sizes) > 1  # requests really shared launches
    else:
        assert max(sizes) == 1
    server.close()


def test_concurrent_submitters(ctx, graphs_and_rows):
    say("serving concurrent submitters")
    graphs, rows = graphs_and_rows
    server = KoptServer(ctx, batching=True, max_batch_cells=20000)
    ids = [server.add_graph(g) for g in graphs]
    failures = []

    def client(seed):
        try:
            for g, k, t in make_requests(rows, 60, seed):
                rid = server.submit(ids[g], k, t)
                costs, _ = server.wait(rid, 60000)
                if not np.array_equal(costs, expected(graphs, g, k, t)):
                    failures.append((seed, g, k))
        except Exception as e:  # noqa: BLE001
            failures.append(repr(e))

    threads = [threading.Thread(target=client, args=(s,)) for s in range(6)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(120)
    assert not failures
    server.close()


def test_graph_added_while_serving(ctx, graphs_and_rows):
    graphs, rows = graphs_and_rows
    server = KoptServer(ctx, batching=True)
    g0 = server.add_graph(graphs[0])
    rid = server.submit(g0, 5, rows[0][5][:10])
    g1 = server.add_graph(graphs[1])
    rid1 = server.submit(g1, 4, rows[1][4][:10])
    c0, _ = server.wait(rid)
    c1, _ = server.wait(rid1)
    assert np.array_equal(c0, expected(graphs, 0, 5, rows[0][5][:10]))
    assert np.array_equal(c1, expected(graphs, 1, 4, rows[1][4][:10]))
    server.close()


def test_admission_rejects_bad_requests(ctx, graphs_and_rows):
    graphs, rows = graphs_and_rows
    server = KoptServer(ctx, hbm_budget_bytes=64 * 1024, batching=True)
    g = server.add_graph(graphs[0])
    big = np.repeat(rows[0][5][:1], 100, axis=0)  # 20800 cells > 32 KiB slot budget
    with pytest.raises(_lib.InvalidArgument):
        server.submit(g, 5, big)
    bad = rows[0][3][:2].copy()
    bad["port_nodes"][0, 0] = graphs[0].dimension
    with pytest.raises(_lib.InvalidArgument):
        server.submit(g, 3, bad)
    with pytest.raises(_lib.InvalidArgument):
        server.submit(g + 7, 3, rows[0][3][:2])
    rid = server.submit(g, 3, rows[0][3][:4])
    costs, _ = server.wait(rid)
    assert np.array_equal(costs, expected(graphs, 0, 3, rows[0][3][:4]))
    server.close()


@pytest.mark.parametrize("batching", [True, False])
def test_jv_rechecks_share_launches_with_leaf_costs(ctx, graphs_and_rows, batching):
    """JV rechecks of edge-slot subsets on resident snapshots (EUC_2D and
    CEIL_2D, some edges inactive) mixed with leaf-cost queries: every
    witness equals the oracle's on that snapshot, every cost row the oracle's."""
    say(f"serving JV rechecks batching={batching}")
    graphs, rows = graphs_and_rows
    ceil = W.config1(seed=0xEE0005C1, n=500)
    ceil = W.knn_graph(ceil.x, ceil.y, 1, 12)
    sparse = graphs[1].copy()
    rng = np.random.default_rng(7)
    sparse.edge_active[rng.random(sparse.edge_count) < 0.1] = 0
    sparse.rebuild_csr()
    jv_graphs = [graphs[0], ceil, sparse]
    want = [O.OGraph.from_snapshot(g).witnesses() for g in jv_graphs]
    server = KoptServer(ctx, batching=batching, max_batch_cells=60000)
    kids = [server.add_graph(g) for g in graphs]
    jids = [server.add_jv_graph(g) for g in jv_graphs]
    submitted = []
    kopt_reqs = make_requests(rows, 60, seed=5)
    for i in range(120):
        if i % 2 == 0:
            j = int(rng.integers(len(jv_graphs)))
            m = jv_graphs[j].edge_count
            edges = rng.choice(m, size=int(rng.integers(1, min(m, 700))), replace=False).astype(np.int32)
            submitted.append(("jv", server.submit_jv(jids[j], edges), j, edges))
        else:
            g, k, t = kopt_reqs[i // 2]
            submitted.append(("kopt", server.submit(kids[g], k, t), g, (k, t)))
    total_edges = 0
    for kind, rid, g, payload in submitted:
        out, _ = server.wait(rid, 60000)
        if kind == "jv":
            assert np.array_equal(out, want[g][payload]), (g, len(payload))
            total_edges += len(payload)
        else:
            k, t = payload
            assert np.array_equal(out, expected(graphs, g, k, t))
    assert server.stats()["jv_edges"] == total_edges
    assert any((w >= 0).any() for w in want)
    with pytest.raises(_lib.InvalidArgument):
        server.submit_jv(jids[0], np.array([jv_graphs[0].edge_count], np.int32))
    with pytest.raises(_lib.InvalidArgument):
        server.submit_jv(len(jv_graphs) + 3, np.array([0], np.int32))
    server.close()


@pytest.mark.parametrize("batching", [True, False])
def test_submit_many_equals_oracle(ctx, graphs_and_rows, batching):
    """Many leaf-cost queries and JV rechecks queued in one call (inputs
    borrowed, not copied) give the oracle's cells and witnesses; a bad
    request anywhere in the list queues none of them."""
    say(f"serving submit_many batching={batching}")
    graphs, rows = graphs_and_rows
    server = KoptServer(ctx, batching=batching, max_batch_cells=50000)
    kids = [server.add_graph(g) for g in graphs]
    jids = [server.add_jv_graph(g) for g in graphs]
    want_w = [O.OGraph.from_snapshot(g).witnesses() for g in graphs]
    rng = np.random.default_rng(11)
    reqs, meta = [], []
    for i, (g, k, t) in enumerate(make_requests(rows, 200, seed=9)):
        if i % 4 == 3:
            m = graphs[g].edge_count
            edges = rng.choice(m, size=int(rng.integers(1, min(m, 400))), replace=False).astype(np.int32)
            reqs.append((jids[g], 0, edges))
            meta.append(("jv", g, edges))
        else:
            reqs.append((kids[g], k, t))
            meta.append(("kopt", g, (k, t)))
    ids = server.submit_many(reqs)
    assert len(ids) == len(reqs) and len(set(ids)) == len(ids)
    for (out, latency), (kind, g, payload) in zip(server.wait_many(ids, 60000), meta):
        assert latency >= 0
        if kind == "jv":
            assert np.array_equal(out, want_w[g][payload])
        else:
            assert np.array_equal(out, expected(graphs, g, *payload))
    before = server.stats()["requests"]
    bad = [(kids[0], 3, rows[0][3][:5]), (99, 3, rows[0][3][:5])]  # graph 99 does not exist
    with pytest.raises(_lib.InvalidArgument):
        server.submit_many(bad)
    assert server.stats()["requests"] == before
    server.close()


def test_served_graphs_at_full_size(ctx):
    """The served workload at its full size: four resident graphs (n = 1k,
    5k, 15k, 30k, uniform, 10-NN), lognormal leaf-cost queries of 128 to
    65,536 cells with JV rechecks mixed in, one of them a 65,536-cell query
    above the batch cell limit (it runs in a launch of its own).  Every
    result equals the CPU oracle and the plain per-query server's."""
    say("serving at full size: graphs 1k/5k/15k/30k")
    graphs = W.config5_graphs()
    assert [g.dimension for g in graphs] == [1000, 5000, 15000, 30000]
    pools = W.config5_row_pools(graphs)
    rng = np.random.default_rng(0xEE0005F5)
    requests, _ = W.config5_requests(graphs, pools, 600, rng)
    big = pools[3][5][:316]  # 316 x 208 = 65,728 cells: the largest query size
    requests.append((3, 5, big))
    requests.append((2, 4, pools[2][4][:2621]))  # 65,525 cells of k = 4
    sizes = [len(t) * TEMPLATES_PER_K[k] for _, k, t in requests if k]
    assert max(sizes) > 32768 and min(sizes) >= 128
    oracles = [O.OGraph.from_snapshot(g) for g in graphs]
    want_w = [o.witnesses() for o in oracles]
    want = []
    for g, k, t in requests:
        want.append(want_w[g][t] if k == 0 else expected(graphs, g, k, t))
    results = {}
    for mode in ("batched", "plain"):
        server = KoptServer(ctx, batching=(mode == "batched"), max_batch_cells=32768,
                            hbm_budget_bytes=1 << 30)
        kids = [server.add_graph(g) for g in graphs]
        jids = [server.add_jv_graph(g) for g in graphs]
        ids = server.submit_many([(kids[g], k, t) if k else (jids[g], 0, t) for g, k, t in requests])
        outs = [out for out, _ in server.wait_many(ids, 120000)]
        results[mode] = outs
        stats = server.stats()
        assert stats["requests"] == len(requests)
        if mode == "batched":
            assert stats["batches"] < len(requests)  # requests shared launches
        server.close()
        bad = [i for i, (o, w) in enumerate(zip(outs, want)) if not np.array_equal(o, w)]
        assert not bad, f"{mode}: requests differ from the oracle: {bad[:10]}"
    for a, b in zip(results["batched"], results["plain"]):
        assert np.array_equal(a, b)
