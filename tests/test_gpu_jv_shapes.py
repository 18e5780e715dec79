"""The JV kernel's two shapes against the oracle.

By default a wavefront takes two edge slots, one per 32-lane half, when both
rows of an edge fit the half (deg_a + deg_b <= 32); an edge that does not fit
(a longer row, a key that does not pack into 64 bits, a candidate row longer
than 32) is redone on a whole wavefront.  `set_jv_edges_per_wave(1)` gives
every edge slot a whole wavefront.  Both shapes must give the oracle's
witnesses (jv_elimination.cpp:45-117) on graphs built to reach each of those
branches; a host-side census of each graph shows that it does.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import _oracle as O
from cudaedgeelimination_amd import _lib, backend, workloads as W
from cudaedgeelimination_amd.graph import CEIL_2D, EUC_2D, GraphSnapshot, exact_distances
from cudaedgeelimination_amd.resident import JvSession

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def random_knn(seed, n, k, dtype, box=4000):
    rng = np.random.default_rng(seed)
    x = rng.integers(0, box, n).astype(np.int64)
    y = rng.integers(0, box, n).astype(np.int64)
    return W.knn_graph(x, y, dtype, k)


def census(g):
    """(eligible edges that fit a half, eligible edges that do not)."""
    deg = g.degrees()
    u, v = g.edge_u, g.edge_v
    eligible = (g.edge_active != 0) & (deg[u] > 2) & (deg[v] > 2)
    total = deg[u] + deg[v]
    return int(np.count_nonzero(eligible & (total <= 32))), int(np.count_nonzero(eligible & (total > 32)))


def hub_graph():
    """A 5-NN graph with node 0 joined to its 40 nearest nodes: edges near
    the hub fit a half but have the hub (row > 32) as a candidate."""
    rng = np.random.default_rng(77)
    n = 400
    x = rng.integers(0, 3000, n).astype(np.int64)
    y = rng.integers(0, 3000, n).astype(np.int64)
    base = W.knn_graph(x, y, EUC_2D, 5)
    pairs = set(zip(base.edge_u.tolist(), base.edge_v.tolist()))
    d = (x - x[0]) ** 2 + (y - y[0]) ** 2
    for j in np.argsort(d)[1:41]:
        pairs.add((0, int(j)))
    u, v = map(np.array, zip(*sorted(pairs)))
    return GraphSnapshot.from_edges(x, y, EUC_2D, u, v)


def wide_graph():
    """An 8-NN graph whose every third edge weighs about 2^41: scores do not
    pack into 40 bits, so those edges leave the pair path after scoring."""
    g = random_knn(31, 600, 8, EUC_2D)
    w = exact_distances(g.x[g.edge_u], g.y[g.edge_u], g.x[g.edge_v], g.y[g.edge_v], EUC_2D)
    w[::3] += 1 << 41
    return GraphSnapshot.from_edges(g.x, g.y, EUC_2D, g.edge_u, g.edge_v, w)


def case(name):
    if name == "lattice":
        return W.config2()
    if name == "knn6_ceil":
        return random_knn(6, 501, 6, CEIL_2D)
    if name == "knn12":
        return random_knn(12, 800, 12, EUC_2D)
    if name == "knn20":
        return random_knn(20, 600, 20, EUC_2D)
    if name == "hub":
        return hub_graph()
    if name == "wide":
        return wide_graph()
    if name == "lattice_inactive":
        g = W.config2()
        rng = np.random.default_rng(5)
        g.edge_active[rng.random(g.edge_count) < 0.3] = 0
        g.rebuild_csr()
        return g
    raise KeyError(name)


CASES = ["lattice", "knn6_ceil", "knn12", "knn20", "hub", "wide", "lattice_inactive"]


@pytest.mark.parametrize("name", CASES)
def test_both_shapes_equal_oracle(ctx, name):
    g = case(name)
    fit, rest = census(g)
    say(f"JV shapes {name}: m={g.edge_count} fit={fit} rest={rest}")
    if name == "knn12":  # both paths in one launch
        assert fit > 0 and rest > 0
    if name == "knn20":  # every eligible edge on a whole wavefront
        assert fit == 0 and rest > 0
    if name == "hub":
        assert g.degrees()[0] > 32
    expected = O.OGraph.from_snapshot(g).witnesses()
    try:
        ctx.set_jv_edges_per_wave(1)
        one, _ = backend.find_jv_witnesses(g, ctx)
        ctx.set_jv_edges_per_wave(2)
        two, _ = backend.find_jv_witnesses(g, ctx)
    finally:
        ctx.set_jv_edges_per_wave(2)
    assert np.array_equal(one, expected)
    assert np.array_equal(two, expected)


def test_odd_and_even_edge_counts(ctx):
    """The last pair of an odd edge count has one empty half."""
    counts = set()
    for n in (37, 38, 45, 64):
        g = random_knn(n, n, 5, EUC_2D, box=500)
        counts.add(g.edge_count % 2)
        w, _ = backend.find_jv_witnesses(g, ctx)
        assert np.array_equal(w, O.OGraph.from_snapshot(g).witnesses()), n
    assert counts == {0, 1}


def test_shapes_give_the_same_session_proof(ctx):
    """A resident multi-epoch run gives the same proof in both shapes."""
    g = W.config2()
    proofs = []
    try:
        for edges in (1, 2):
            ctx.set_jv_edges_per_wave(edges)
            s = JvSession(ctx, g)
            proofs.append(O.canonical_proof(s.run(100).proof))
            s.close()
    finally:
        ctx.set_jv_edges_per_wave(2)
    assert proofs[0] == proofs[1]


def test_shape_setting_is_validated(ctx):
    from cudaedgeelimination_amd import _lib

    for bad in (0, 3):
        with pytest.raises(_lib.InvalidArgument):
            ctx.set_jv_edges_per_wave(bad)


def test_large_epoch_session_equals_oracle(ctx):
    """A 15,000-node 10-NN graph: the first epoch commits about 7,600 edges,
    so the host re-verifies on several threads; the device commits, compacts
    and hashes every generation.  The proof equals the oracle's."""
    g = W.config5_graphs(sizes=(15000,))[0]
    expected = O.jv_run(O.OGraph.from_snapshot(g), max_rounds=100)
    s = JvSession(ctx, g)
    run = s.run(100)
    s.close()
    say(f"JV 15000: epochs={run.epochs} committed={run.committed} total_ms={run.total_ms:.3f}")
    assert run.committed > 2048
    assert O.canonical_proof(run.proof) == O.canonical_proof(expected)


def test_corrupted_witness_stops_the_run_and_commits_nothing(ctx):
    """A device witness corrupted on the host before re-verification (the
    reference stops the run on a rejected candidate, jv_elimination.cpp:
    301-303): the run fails, nothing is committed on the host or the device,
    and the next run gives the oracle's proof."""
    g = W.config2()
    expected = O.jv_run(O.OGraph.from_snapshot(g), max_rounds=100)
    s = JvSession(ctx, g)
    before = s.active().copy()
    e = int(np.flatnonzero(g.degrees()[g.edge_u] > 2)[0])
    s.inject_witness(e, int(g.edge_u[e]))  # a candidate equal to an endpoint is never a witness
    with pytest.raises(_lib.EeError) as err:
        s.run(100)
    assert err.value.status == 3 and "re-verification" in err.value.message
    assert np.array_equal(s.active(), before)  # host snapshot
    dev = s.device_snapshot()
    assert np.array_equal(dev["edge_active"], before)  # the device committed nothing either
    assert np.array_equal(dev["row_offsets"], g.row_offsets)
    run = s.run(100)
    s.close()
    assert O.canonical_proof(run.proof) == O.canonical_proof(expected)
