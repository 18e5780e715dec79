"""HIP paths against the CPU oracle and the reference-generated fixtures.

Integer work, so every comparison is bit-exact.  Scenarios follow the
reference's own GPU tests (tests/unit/test_core.cpp:167-229,
tests/unit/test_kopt.cpp:158-328) plus seeded configurations, the
multi-pass top-ten merge (rows longer than a wavefront) and edge cases.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import math
import pathlib
import random
import sys

import numpy as np
import pytest

import _oracle as O
from cudaedgeelimination_amd import _lib, backend, workloads as W
from cudaedgeelimination_amd.graph import CEIL_2D, EUC_2D, GraphSnapshot, exact_distances
from cudaedgeelimination_amd.resident import KoptBatch

pytestmark = pytest.mark.gpu

ROOT = pathlib.Path(__file__).resolve().parent.parent
GOLDEN = ROOT / "tests" / "golden"
G = json.loads((GOLDEN / "goldens.json").read_text())
RP = GOLDEN / "recursive-point"
INT64_MAX = np.iinfo(np.int64).max
T_OF = {3: 4, 4: 25, 5: 208}


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


# ---------------------------------------------------------------- distance
def test_device_isqrt_is_exact(ctx):
    say("device isqrt: 2^24 consecutive + near squares + random 64-bit")
    rng = random.Random(5)
    extra = []
    for _ in range(200000):
        r = rng.randrange(1 << 32)
        extra += [max(r * r - 1, 0), r * r, (r * r + 2 * r) % (1 << 64), rng.randrange(1 << 64)]
    extra += [(1 << 64) - 1, ((1 << 32) - 1) ** 2, (1 << 53) - 1, 1 << 53, (1 << 53) + 1]
    s = np.concatenate([np.arange(1 << 24, dtype=np.uint64), np.array(extra, dtype=np.uint64)])
    dev = np.empty_like(s)
    host = np.empty_like(s)
    _lib.check(_lib.lib().ee_isqrt_device(ctx.handle, O.p(s), O.p(dev), len(s)))
    _lib.lib().ee_isqrt_host(O.p(s), O.p(host), len(s))
    assert np.array_equal(dev, host)
    tail = np.array([math.isqrt(int(v)) for v in s[-len(extra):]], dtype=np.uint64)
    assert np.array_equal(dev[-len(extra):], tail)


@pytest.mark.parametrize("dtype", [EUC_2D, CEIL_2D])
def test_device_distances_match_oracle(ctx, dtype):
    rng = np.random.default_rng(17)
    n = 4096
    # coordinates spread up to the distance-safe limit (squared diagonal < 2^64)
    x = rng.integers(-(1 << 30), 1 << 30, n).astype(np.int64)
    y = rng.integers(-(1 << 30), 1 << 30, n).astype(np.int64)
    a = rng.integers(0, n, 200000).astype(np.int32)
    b = rng.integers(0, n, 200000).astype(np.int32)
    out = np.empty(len(a), np.int64)
    _lib.check(_lib.lib().ee_distances_device(ctx.handle, O.p(x), O.p(y), n, dtype, O.p(a),
                                              O.p(b), len(a), O.p(out)))
    assert np.array_equal(out, exact_distances(x[a], y[a], x[b], y[b], dtype))
    lib = O.oracle()
    v = C.c_int64()
    for i in range(0, len(a), 997):
        assert lib.orc_exact_distance(int(x[a[i]]), int(y[a[i]]), int(x[b[i]]), int(y[b[i]]),
                                      dtype, C.byref(v)) == 0
        assert out[i] == v.value


@pytest.mark.parametrize("dtype", [EUC_2D, CEIL_2D])
def test_device_near_distances_are_exact(ctx, dtype):
    """Coordinate extents below 2^25 take the FP64 path: perfect squares,
    their neighbours and the EUC_2D half-way points at the top of the range."""
    say(f"device FP64 distances dtype={dtype}")
    lim = (1 << 25) - 1
    rng = np.random.default_rng(23)
    m = np.arange(lim - 4000, lim + 1, dtype=np.int64)
    # node 0 at the origin; nodes on both axes and random points in the box
    x = np.concatenate([[0], m, np.zeros_like(m), rng.integers(0, lim + 1, 20000)]).astype(np.int64)
    y = np.concatenate([[0], np.zeros_like(m), m, rng.integers(0, lim + 1, 20000)]).astype(np.int64)
    y[1:len(m) + 1:3] = 1  # s = m^2 + 1
    # s = (2t^2 + 1)^2 - 1 just below a square, s = u^4 + u^2 on the EUC_2D
    # half-way point (m = u^2, s - m^2 = m) and s = u^4 - u^2 + 1 just past it
    t = np.arange(3000, 4096, dtype=np.int64)
    u = np.arange(4000, 5793, dtype=np.int64)
    x = np.concatenate([x, 2 * t * t, u * u, u * u - 1])
    y = np.concatenate([y, 2 * t, u, u])
    assert x.max() < (1 << 25) and y.max() < (1 << 25)
    n = len(x)
    a = np.concatenate([np.zeros(n - 1, np.int32), rng.integers(0, n, 200000).astype(np.int32)])
    b = np.concatenate([np.arange(1, n, dtype=np.int32), rng.integers(0, n, 200000).astype(np.int32)])
    out = np.empty(len(a), np.int64)
    _lib.check(_lib.lib().ee_distances_device(ctx.handle, O.p(x), O.p(y), n, dtype, O.p(a),
                                              O.p(b), len(a), O.p(out)))
    assert np.array_equal(out, exact_distances(x[a], y[a], x[b], y[b], dtype))


# ------------------------------------------------------------------- k-opt
def seventeen_graph(dtype=EUC_2D):
    x = np.array([(7 * i) % 23 for i in range(17)], np.int64)
    y = np.array([(i * i + 3 * i) % 19 for i in range(17)], np.int64)
    return GraphSnapshot.from_edges(x, y, dtype, np.zeros(0, np.int32), np.zeros(0, np.int32))


def test_kopt_reference_unit_test_scenario(ctx):
    """tests/unit/test_kopt.cpp:158-328: matrices, cache-hit pattern, the
    40-row batch and the CEIL_2D snapshot refresh."""
    from golden.make_goldens import seventeen_node_tasks

    say("k-opt 17-node scenario")
    g = seventeen_graph()
    backend.clear_kopt_cost_cuda_cache(ctx)
    first_bytes = None
    for k in (3, 4, 5):
        table = backend.build_kopt_reconnect_table(k)
        tasks = seventeen_node_tasks(g.x, g.y, k)
        costs, device, usage = backend.evaluate_kopt_template_costs_cuda(g, table, tasks, ctx)
        assert device >= 0
        assert costs.tolist() == G["kopt_17"][str(k)]
        assert usage.snapshot_hit == (k != 3)
        assert not usage.template_hit and not usage.workspace_hit and usage.resident_bytes > 0
        if k == 3:
            first_bytes = usage.resident_bytes
            again, _, u2 = backend.evaluate_kopt_template_costs_cuda(g, table, tasks, ctx)
            assert np.array_equal(again, costs)
            assert u2.snapshot_hit and u2.template_hit and u2.workspace_hit
            assert u2.resident_bytes == first_bytes
    rep = np.zeros(40, _lib.KOPT_TASK_DTYPE)
    rep["port_nodes"] = np.arange(10)
    t5 = backend.build_kopt_reconnect_table(5)
    costs, _, _ = backend.evaluate_kopt_template_costs_cuda(g, t5, rep, ctx)
    assert all(costs[i * 208:(i + 1) * 208].tolist() == G["kopt_17_repeated"]["row"]
               for i in range(40))
    ceil = seventeen_graph(CEIL_2D)
    one = np.zeros(1, _lib.KOPT_TASK_DTYPE)
    one["port_nodes"] = np.arange(10)
    costs, _, usage = backend.evaluate_kopt_template_costs_cuda(ceil, t5, one, ctx)
    assert costs.tolist() == G["kopt_17_ceil"]
    assert not usage.snapshot_hit and usage.template_hit and usage.workspace_hit
    backend.clear_kopt_cost_cuda_cache(ctx)


@pytest.mark.parametrize("half", [5000, (1 << 24) - 1, 1 << 30])
@pytest.mark.parametrize("dtype", [EUC_2D, CEIL_2D])
@pytest.mark.parametrize("k", [3, 4, 5])
def test_kopt_random_rows_match_oracle(ctx, k, dtype, half):
    """Coordinates in [-half, half): the FP64 distance path up to extents of
    2^25 - 2, the 64-bit integer path beyond."""
    say(f"k-opt random rows k={k} dtype={dtype} half={half}")
    rng = random.Random(31 * k + dtype + half)
    n = 64
    x = np.array([rng.randrange(-half, half) for _ in range(n)], np.int64)
    y = np.array([rng.randrange(-half, half) for _ in range(n)], np.int64)
    g = GraphSnapshot.from_edges(x, y, dtype, np.zeros(0, np.int32), np.zeros(0, np.int32))
    tasks = np.zeros(777, _lib.KOPT_TASK_DTYPE)
    for t in tasks:
        mode = rng.random()
        if mode < 0.4:
            nodes = rng.sample(range(n), 2 * k)
        elif mode < 0.8:  # consecutive chain: shared endpoints
            walk = rng.sample(range(n), k + 1)
            nodes = [v for e in range(k) for v in (walk[e], walk[e + 1])]
        else:  # heavy repetition over few nodes
            pool = rng.sample(range(n), k + 1)
            nodes = []
            for _ in range(k):
                while True:
                    a, b = rng.sample(pool, 2)
                    key = (min(a, b), max(a, b))
                    if key not in [(min(nodes[i], nodes[i + 1]), max(nodes[i], nodes[i + 1]))
                                   for i in range(0, len(nodes), 2)]:
                        break
                nodes += [a, b]
        t["port_nodes"][:2 * k] = nodes
        t["port_nodes"][2 * k:] = rng.randrange(n)  # unused ports carry junk
    table = backend.build_kopt_reconnect_table(k)
    got, _, _ = backend.evaluate_kopt_template_costs_cuda(g, table, tasks, ctx)
    want, _, _ = O.kopt_costs(x, y, dtype, k, tasks)
    assert np.array_equal(got, want)


def random_partition_row(rng, k):
    """Port labels 0..2k-1 of one valid task (k distinct deleted edges, no self
    loop): all distinct, a chain, or heavy repetition over k + 1 labels."""
    mode = rng.random()
    if mode < 0.3:
        return list(range(2 * k))
    if mode < 0.6:
        return [v for e in range(k) for v in (e, e + 1)]
    labels = []
    for _ in range(k):
        while True:
            a, b = rng.sample(range(k + 1), 2)
            if {a, b} not in [{labels[i], labels[i + 1]} for i in range(0, len(labels), 2)]:
                break
        labels += [a, b]
    return labels


@pytest.mark.parametrize("dtype", [EUC_2D, CEIL_2D])
@pytest.mark.parametrize("k", [3, 4, 5])
def test_kopt_shared_partition_tiles_match_oracle(ctx, k, dtype):
    """Blocks of 64 rows that all have one port partition (the same labels
    under random relabelings: all distinct, a chain, or heavy repetition),
    some with one row of another partition; several tile sizes move the block
    edges inside tiles."""
    say(f"k-opt shared-partition tiles k={k} dtype={dtype}")
    rng = random.Random(1000 * k + dtype)
    n = 96
    x = np.array([rng.randrange(0, 1 << 24) for _ in range(n)], np.int64)
    y = np.array([rng.randrange(0, 1 << 24) for _ in range(n)], np.int64)
    g = GraphSnapshot.from_edges(x, y, dtype, np.zeros(0, np.int32), np.zeros(0, np.int32))
    tasks = np.zeros(64 * 12, _lib.KOPT_TASK_DTYPE)
    for blk in range(12):
        labels = random_partition_row(rng, k)
        odd = rng.randrange(64) if blk % 3 == 1 else -1
        for i in range(64):
            row = labels if i != odd else random_partition_row(rng, k)
            relabel = rng.sample(range(n), 2 * k)
            t = tasks[blk * 64 + i]
            t["port_nodes"][:2 * k] = [relabel[v] for v in row]
            t["port_nodes"][2 * k:] = rng.randrange(n)
    want, _, _ = O.kopt_costs(x, y, dtype, k, tasks)
    table = backend.build_kopt_reconnect_table(k)
    got, _, _ = backend.evaluate_kopt_template_costs_cuda(g, table, tasks, ctx)
    assert np.array_equal(got, want)
    b = KoptBatch(ctx)
    inst = b.add_instance(g)
    b.set_work([(inst, k, len(tasks))], tasks)
    for rows in (64, 24, 7, 1):
        tuning = [32, 16, 8]
        tuning[k - 3] = rows
        b.set_tuning(tuning, 0, 64, 0)
        b.run(1)
        assert np.array_equal(b.segment_costs(0), want), rows
    b.close()


def test_kopt_arbitrary_template_tables_match_edge_list_oracle(ctx):
    """Every perfect matching of the 10 ports (including ones that re-add a
    deleted edge), in a shuffled order, against the per-cell edge-list oracle."""
    say("k-opt all 945 matchings")
    k = 5
    rng = random.Random(99)

    def matchings(free):
        if not free:
            yield []
            return
        for i in range(1, len(free)):
            for m in matchings(free[1:i] + free[i + 1:]):
                yield [(free[0], free[i])] + m

    mates = []
    for m in matchings(list(range(10))):
        mate = [0xFF] * 14
        for a, b in m:
            mate[a], mate[b] = b, a
        mates.append(mate)
    rng.shuffle(mates)
    table_arr = np.zeros(len(mates), _lib.MATCHING_DTYPE)
    table_arr["endpoint_count"] = 10
    table_arr["mate"] = mates
    table = backend.KOptReconnectTable(5, table_arr, 0x1234)
    g = seventeen_graph()
    tasks = np.zeros(12, _lib.KOPT_TASK_DTYPE)
    for t in tasks:
        walk = rng.sample(range(17), 6)
        t["port_nodes"] = [v for e in range(5) for v in (walk[e], walk[e + 1])] \
            if rng.random() < 0.5 else rng.sample(range(17), 10)
    got, _, usage = backend.evaluate_kopt_template_costs_cuda(g, table, tasks, ctx)
    dist = lambda a, b: int(exact_distances(g.x[a], g.y[a], g.x[b], g.y[b], EUC_2D))  # noqa: E731
    T = len(mates)
    for r, t in enumerate(tasks):
        for j, mate in enumerate(mates):
            assert got[r * T + j] == O.edge_list_cost(dist, k, list(t["port_nodes"]), mate[:10])
    backend.clear_kopt_cost_cuda_cache(ctx)


@pytest.mark.parametrize("name", ["config3_small", "config3"])
def test_kopt_seeded_workloads_equal_reference(ctx, name):
    say(f"k-opt {name}")
    entry = G[name]
    wl = W.config3(seed=0xEE0000C3, n=2000, rows=4096) if name == "config3_small" else W.config3()
    assert sha(wl.tasks) == entry["tasks_sha256"]
    for k in (3, 4, 5):
        table = backend.build_kopt_reconnect_table(k)
        got, _, _ = backend.evaluate_kopt_template_costs_cuda(wl.graph, table, wl.bucket(k), ctx)
        assert sha(got) == entry[f"k{k}"]["sha256"], f"k={k}"
        assert got[:T_OF[k]].tolist() == entry[f"k{k}"]["first_row"]
        assert int((got == INT64_MAX).sum()) == entry[f"k{k}"]["invalid_cells"]


def test_kopt_rejects_invalid_input(ctx):
    g = seventeen_graph()
    t3 = backend.build_kopt_reconnect_table(3)
    bad = np.zeros(1, _lib.KOPT_TASK_DTYPE)
    bad["port_nodes"][0, :6] = [0, 1, 1, 0, 2, 3]
    with pytest.raises(_lib.InvalidArgument):
        backend.evaluate_kopt_template_costs_cuda(g, t3, bad, ctx)
    bad["port_nodes"][0, :6] = [0, 1, 2, 3, 4, 17]  # node out of range
    with pytest.raises(_lib.InvalidArgument):
        backend.evaluate_kopt_template_costs_cuda(g, t3, bad, ctx)
    broken = backend.KOptReconnectTable(3, t3.templates.copy(), t3.generator_hash ^ 1)
    broken.templates["mate"][0, 0] = 7  # not a perfect matching of 6 ports
    bad["port_nodes"][0, :6] = [0, 1, 2, 3, 4, 5]
    with pytest.raises(_lib.InvalidArgument):
        backend.evaluate_kopt_template_costs_cuda(g, broken, bad, ctx)
    costs, _, _ = backend.evaluate_kopt_template_costs_cuda(g, t3, bad[:0], ctx)
    assert len(costs) == 0


# ---------------------------------------------------------------------- JV
def test_jv_reference_unit_test_scenario(ctx):
    """tests/unit/test_core.cpp:167-229 on the reference fixture: candidates
    equal the CPU path, cache-hit pattern and resident byte formula."""
    say("JV recursive-point scenario")
    g = GraphSnapshot.load(RP / "recursive-point.tsp", RP / "recursive-point.edg")
    backend.clear_jv_cuda_cache(ctx)
    w, u1 = backend.find_jv_witnesses(g, ctx)
    assert w.tolist() == G["recursive_point_witnesses"]
    assert not u1.static_hit and not u1.workspace_hit
    assert u1.resident_bytes == 24 * g.edge_count + 20 * g.dimension + 4 + 8 * g.adjacency_count
    assert all(v >= 0 and math.isfinite(v) for v in (u1.h2d_ms, u1.kernel_ms, u1.d2h_ms))
    w2, u2 = backend.find_jv_witnesses(g, ctx)
    assert np.array_equal(w, w2) and u2.static_hit and u2.workspace_hit
    assert u2.resident_bytes == u1.resident_bytes
    g.edge_active[0] = 0
    g.rebuild_csr()
    w3, u3 = backend.find_jv_witnesses(g, ctx)
    assert np.array_equal(w3, O.OGraph.from_snapshot(g).witnesses())
    assert u3.static_hit and u3.workspace_hit
    g.x[0] += 1000
    g.edge_weight[:] = exact_distances(g.x[g.edge_u], g.y[g.edge_u], g.x[g.edge_v],
                                       g.y[g.edge_v], g.distance_type)
    g.rebuild_csr()
    w4, u4 = backend.find_jv_witnesses(g, ctx)
    assert np.array_equal(w4, O.OGraph.from_snapshot(g).witnesses())
    assert not u4.static_hit and u4.workspace_hit
    cands, device, _ = backend.find_jv_candidates_cuda(g, ctx)
    assert device >= 0 and [c[0] for c in cands] == sorted(c[0] for c in cands)
    backend.clear_jv_cuda_cache(ctx)


def test_jv_complete_graph_k4(ctx):
    # tests/unit/test_core.cpp:86-117
    x = np.array([0, 10, 0, 10], np.int64)
    y = np.array([0, 0, 10, 10], np.int64)
    u, v = zip(*[(a, b) for a in range(4) for b in range(a + 1, 4)])
    g = GraphSnapshot.from_edges(x, y, EUC_2D, u, v)
    w, _ = backend.find_jv_witnesses(g, ctx)
    og = O.OGraph.from_snapshot(g)
    assert np.array_equal(w, og.witnesses())
    assert all(og.verify(e, int(c)) for e, c in enumerate(w) if c >= 0)


@pytest.mark.parametrize("name", ["config1", "config2", "config1_ceil"])
def test_jv_seeded_first_epoch_equals_reference(ctx, name):
    say(f"JV {name}")
    if name == "config1_ceil":
        base = W.config1(seed=0xEE0000C1, n=400)
        g = W.knn_graph(base.x, base.y, CEIL_2D, 12)
    else:
        g = W.config1() if name == "config1" else W.config2()
    w, _ = backend.find_jv_witnesses(g, ctx)
    assert sha(w) == G[f"{name}_first_epoch"]["witness_sha256"]


def random_knn(seed, n, k, dtype, box=4000):
    rng = np.random.default_rng(seed)
    x = rng.integers(0, box, n).astype(np.int64)
    y = rng.integers(0, box, n).astype(np.int64)
    return W.knn_graph(x, y, dtype, k)


@pytest.mark.parametrize("seed,n,k", [(1, 500, 6), (2, 800, 12), (3, 400, 40), (4, 300, 70),
                                      (5, 2000, 20)])
def test_jv_random_graphs_with_inactive_edges(ctx, seed, n, k):
    """Dense rows (k = 40, 70) take the multi-pass top-ten merge."""
    say(f"JV random n={n} k={k}")
    g = random_knn(seed, n, k, seed % 2)
    rng = np.random.default_rng(seed)
    g.edge_active[rng.random(g.edge_count) < 0.15] = 0
    g.rebuild_csr()
    assert g.degrees().max() > (32 if k >= 40 else 0)
    w, _ = backend.find_jv_witnesses(g, ctx)
    assert np.array_equal(w, O.OGraph.from_snapshot(g).witnesses())


@pytest.mark.parametrize("box", [1 << 25, (1 << 25) + 1, (1 << 31) - 1])
def test_jv_coordinate_ranges(ctx, box):
    """Coordinate extents just inside and outside the FP64 distance path, and
    at the distance-safe limit."""
    say(f"JV box={box}")
    rng = np.random.default_rng(11)
    x = rng.integers(0, box, 700).astype(np.int64)
    y = rng.integers(0, box, 700).astype(np.int64)
    x[:2], y[:2] = (0, box - 1), (box - 1, 0)  # extents exactly box - 1
    g = W.knn_graph(x, y, EUC_2D, 9)
    w, _ = backend.find_jv_witnesses(g, ctx)
    assert np.array_equal(w, O.OGraph.from_snapshot(g).witnesses())


def test_jv_lattice_ties(ctx):
    """Exact lattice without jitter: many equal scores, tie-break by node."""
    side = 30
    x = np.repeat(np.arange(side) * 10, side).astype(np.int64)
    y = np.tile(np.arange(side) * 10, side).astype(np.int64)
    g = W.knn_graph(x, y, EUC_2D, 12)
    w, _ = backend.find_jv_witnesses(g, ctx)
    assert np.array_equal(w, O.OGraph.from_snapshot(g).witnesses())


def test_jv_edge_cases(ctx):
    # a path: every degree <= 2, so nothing is eligible
    x = np.arange(10, dtype=np.int64) * 7
    y = np.zeros(10, np.int64)
    path = GraphSnapshot.from_edges(x, y, EUC_2D, np.arange(9), np.arange(1, 10))
    w, _ = backend.find_jv_witnesses(path, ctx)
    assert (w == -1).all()
    # every edge inactive
    g = random_knn(8, 200, 8, EUC_2D)
    g.edge_active[:] = 0
    g.rebuild_csr()
    w, _ = backend.find_jv_witnesses(g, ctx)
    assert (w == -1).all()
    # no edges at all
    empty = GraphSnapshot.from_edges(x, y, EUC_2D, np.zeros(0, np.int32), np.zeros(0, np.int32))
    w, _ = backend.find_jv_witnesses(empty, ctx)
    assert len(w) == 0
    # malformed CSR is refused before any launch
    bad = random_knn(9, 100, 6, EUC_2D)
    bad.neighbors = bad.neighbors.copy()
    bad.neighbors[0] = 100
    with pytest.raises(_lib.InvalidArgument):
        backend.find_jv_witnesses(bad, ctx)


def test_jv_multi_epoch_matches_oracle_each_epoch(ctx):
    """Config 2 to a fixed point, proposing on the device each epoch and
    committing with the oracle; every epoch's witnesses must match."""
    say("JV config2 epochs")
    g = W.config2()
    og = O.OGraph.from_snapshot(g)
    for epoch in range(20):
        w, usage = backend.find_jv_witnesses(g, ctx)
        assert np.array_equal(w, og.witnesses()), f"epoch {epoch}"
        ids = np.flatnonzero(w >= 0).astype(np.int32)
        if len(ids) == 0:
            break
        cw = w[ids].astype(np.int32)
        done_e = np.empty(len(ids), np.int32)
        done_w = np.empty(len(ids), np.int32)
        got = O.oracle().orc_jv_commit(og.h, len(ids), O.p(ids), O.p(cw), og.hash(),
                                       O.p(done_e), O.p(done_w))
        assert got >= 0
        g.edge_active[done_e[:got]] = 0
        g.rebuild_csr()
        assert g.content_hash() == og.hash()
