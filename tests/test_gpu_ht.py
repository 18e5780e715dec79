"""HT reply enumerations on the device against the oracle (bit-exact).

The oracle's enumerations equal the reference CPU path's
(tests/test_oracle_ht.py).  Here the device's offsets and replies must
equal the oracle's on the same seeded cases: three graphs (uniform,
lattice with tied distances, CEIL_2D), a dense graph whose rows exceed the
64 neighbours a wavefront caches in LDS, a graph with coordinate extents
beyond 2^25 (the 64-bit distance path; the others take the FP64 one), all 30,000 nodes of a served-size
graph as centers, end replies, the cache-hit pattern of the reference's
HtReplyCudaCacheUsage, and the input checks of its CUDA path
(src/cuda/ht_replies.cu:521-580).
"""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

import _oracle as O
from cudaedgeelimination_amd import _lib, backend, workloads as W
from cudaedgeelimination_amd.graph import EUC_2D
from test_oracle_ht import end_tasks, graphs, ht_cases, pair_bound

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def snapshots():
    return graphs()  # includes dense90 (rows beyond the LDS cache) and wide (64-bit distances)


@pytest.mark.parametrize("name", ["uniform300", "lattice", "ceil400", "dense90", "wide"])
def test_hamilton_replies_match_oracle(ctx, snapshots, name):
    g = snapshots[name]
    say(f"ht hamilton {name} max degree {int(np.diff(g.row_offsets).max())}")
    og = O.OGraph.from_snapshot(g)
    for target, centers in ht_cases(g, 0xC3, targets=5, centers=64 if name == "dense90" else 257):
        off, rep, _ = backend.evaluate_ht_hamilton_replies_cuda(g, target, centers, ctx)
        woff, wrep = O.ht_hamilton_replies(og, target, centers, pair_bound(g, centers))
        assert np.array_equal(off, woff)
        assert np.array_equal(rep, wrep)


@pytest.mark.parametrize("name", ["uniform300", "lattice", "ceil400", "dense90", "wide"])
def test_end_replies_match_oracle(ctx, snapshots, name):
    g = snapshots[name]
    say(f"ht end {name}")
    tasks = end_tasks(g, 0xD4, count=1000)
    off, rep, _ = backend.evaluate_ht_end_replies_cuda(g, tasks, ctx)
    deg = np.diff(g.row_offsets).astype(np.int64)
    woff, wrep = O.ht_end_replies(O.OGraph.from_snapshot(g), tasks, int((deg[tasks[:, 0]] - 1).sum()))
    assert np.array_equal(off, woff) and np.array_equal(rep, wrep)


def test_every_node_of_a_30k_graph(ctx):
    g = W.config5_graphs(sizes=(30000,))[0]
    e = int(np.flatnonzero(g.edge_active)[12345])
    target = (int(g.edge_u[e]), int(g.edge_v[e]))
    centers = np.setdiff1d(np.arange(g.dimension), target).astype(np.int32)
    say(f"ht hamilton 30k centers={len(centers)}")
    off, rep, usage = backend.evaluate_ht_hamilton_replies_cuda(g, target, centers, ctx)
    woff, wrep = O.ht_hamilton_replies(O.OGraph.from_snapshot(g), target, centers,
                                       pair_bound(g, centers))
    assert np.array_equal(off, woff) and np.array_equal(rep, wrep)
    assert len(rep) > 0


def test_cache_usage_and_checks(ctx, snapshots):
    say("ht cache and checks")
    g = snapshots["uniform300"]
    backend.clear_ht_reply_cuda_cache(ctx)
    (target, centers), = ht_cases(g, 0xE5, targets=1, centers=64)
    _, _, u1 = backend.evaluate_ht_hamilton_replies_cuda(g, target, centers, ctx)
    _, _, u2 = backend.evaluate_ht_hamilton_replies_cuda(g, target, centers, ctx)
    assert not u1.graph_hit and u2.graph_hit and u2.workspace_hit
    assert u2.resident_bytes >= u1.resident_bytes > 0
    backend.clear_ht_reply_cuda_cache(ctx)
    _, _, u3 = backend.evaluate_ht_hamilton_replies_cuda(g, target, centers, ctx)
    assert not u3.graph_hit
    # empty center list
    off, rep, _ = backend.evaluate_ht_hamilton_replies_cuda(g, target, np.zeros(0, np.int32), ctx)
    assert list(off) == [0] and len(rep) == 0
    with pytest.raises(_lib.InvalidArgument):  # a center on the target
        backend.evaluate_ht_hamilton_replies_cuda(g, target, np.array([target[0]], np.int32), ctx)
    u, v = int(g.edge_u[0]), int(g.edge_v[0])
    with pytest.raises(_lib.InvalidArgument):  # target not an edge
        far = next(w for w in range(g.dimension)
                   if w != u and w not in g.neighbors[g.row_offsets[u]:g.row_offsets[u + 1]])
        backend.evaluate_ht_hamilton_replies_cuda(g, (u, far), centers, ctx)
    with pytest.raises(_lib.InvalidArgument):  # end task off the graph
        backend.evaluate_ht_end_replies_cuda(g, np.array([[u, u]], np.int32), ctx)
    # a CSR that still lists an edge marked inactive is refused
    active = g.edge_active.copy()
    active[0] = 0
    bad = dataclasses.replace(g, edge_active=active)
    assert np.array_equal(bad.row_offsets, g.row_offsets)
    with pytest.raises(_lib.InvalidArgument):
        backend.evaluate_ht_end_replies_cuda(bad, np.array([[u, v]], np.int32), ctx)
    # the context still answers after the refusals
    off, rep, _ = backend.evaluate_ht_hamilton_replies_cuda(g, target, centers, ctx)
    woff, wrep = O.ht_hamilton_replies(O.OGraph.from_snapshot(g), target, centers,
                                       pair_bound(g, centers))
    assert np.array_equal(off, woff) and np.array_equal(rep, wrep)
