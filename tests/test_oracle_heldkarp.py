"""Safety of JV elimination against exact optima on small complete graphs.

The reference's exhaustive check (tests/exhaustive/check_small_instances.py:
19-71, 117-145) builds, for n = 6..12, a complete EUC_2D graph on n distinct
points drawn by `random.Random(0xC0DA0000 + n)` from [0, 2000]^2, runs JV
elimination to a fixed point and requires that no deleted edge lies on any
optimal tour.  Here the optimal-tour edges come from a forward/backward
Held-Karp formulation (an edge (v, w) is on an optimal tour iff some partial
path through a node set ending at v, the edge, and the best completion from w
sum to the optimum), itself checked against brute force over all tours for
n <= 8.  The oracle's JV run is checked on every n (CPU); the device session
runs the same instances and must give the oracle's proofs (GPU).
"""
from __future__ import annotations

import itertools
import math
import random

import numpy as np
import pytest

import _oracle as O
from cudaedgeelimination_amd.graph import EUC_2D, GraphSnapshot

INF = float("inf")


def euc_2d(a, b) -> int:
    s = (a[0] - b[0]) ** 2 + (a[1] - b[1]) ** 2
    r = math.isqrt(s)
    return r + (1 if s - r * r > r else 0)


def instance(n: int):
    rng = random.Random(0xC0DA0000 + n)
    pts = []
    while len(pts) < n:
        p = (rng.randrange(0, 2001), rng.randrange(0, 2001))
        if p not in pts:
            pts.append(p)
    d = [[euc_2d(pts[u], pts[v]) for v in range(n)] for u in range(n)]
    return pts, d


def optimal_tour_edges(d):
    """(optimum, set of (u, v), u < v, on at least one optimal tour)."""
    n = len(d)
    full = (1 << n) - 1
    f = [[INF] * n for _ in range(1 << n)]  # 0 -> ... -> v through exactly `mask`
    f[1][0] = 0
    for mask in range(1, 1 << n, 2):
        for v in range(n):
            fv = f[mask][v]
            if fv == INF or not mask >> v & 1:
                continue
            for w in range(1, n):
                if not mask >> w & 1:
                    nm = mask | 1 << w
                    if fv + d[v][w] < f[nm][w]:
                        f[nm][w] = fv + d[v][w]
    g = [[INF] * n for _ in range(1 << n)]  # v -> the nodes outside `mask` -> 0
    for v in range(1, n):
        g[full][v] = d[v][0]
    for mask in range(full - 2, 0, -2):  # odd masks (containing 0), decreasing
        for v in range(n):
            if not mask >> v & 1:
                continue
            best = INF
            for w in range(1, n):
                if not mask >> w & 1:
                    best = min(best, d[v][w] + g[mask | 1 << w][w])
            g[mask][v] = best
    opt = min(f[full][v] + d[v][0] for v in range(1, n))
    edges = set()
    for v in range(1, n):
        if f[full][v] + d[v][0] == opt:
            edges.add((0, v))
    for mask in range(1, full, 2):
        for v in range(n):
            if f[mask][v] == INF or not mask >> v & 1:
                continue
            for w in range(1, n):
                if not mask >> w & 1 and f[mask][v] + d[v][w] + g[mask | 1 << w][w] == opt:
                    edges.add((min(v, w), max(v, w)))
    return opt, edges


def brute_force_edges(d):
    n = len(d)
    best, edges = None, set()
    for perm in itertools.permutations(range(1, n)):
        tour = (0,) + perm
        cost = sum(d[tour[i]][tour[(i + 1) % n]] for i in range(n))
        tour_edges = {tuple(sorted((tour[i], tour[(i + 1) % n]))) for i in range(n)}
        if best is None or cost < best:
            best, edges = cost, set(tour_edges)
        elif cost == best:
            edges |= tour_edges
    return best, edges


def complete_graph(pts, d):
    n = len(pts)
    u, v = zip(*[(a, b) for a in range(n) for b in range(a + 1, n)])
    w = [d[a][b] for a, b in zip(u, v)]
    x = np.array([p[0] for p in pts], np.int64)
    y = np.array([p[1] for p in pts], np.int64)
    return x, y, np.array(u, np.int32), np.array(v, np.int32), np.array(w, np.int64)


@pytest.mark.parametrize("n", [6, 7, 8])
def test_held_karp_edges_equal_brute_force(n):
    _, d = instance(n)
    assert optimal_tour_edges(d) == brute_force_edges(d)


def jv_removed_edges(x, y, u, v, w):
    og = O.OGraph.from_arrays(x, y, EUC_2D, u, v, w)
    proof = O.jv_run(og, max_rounds=100)
    final = og.clone()
    assert O.oracle().orc_jv_replay(final.h, proof.encode()) == 0
    active = final.arrays()["edge_active"]
    return proof, {(int(a), int(b)) for a, b, on in zip(u, v, active) if not on}


def test_jv_deletes_no_optimal_tour_edge():
    removed_total = 0
    for n in range(6, 13):
        pts, d = instance(n)
        _, optimal = optimal_tour_edges(d)
        _, removed = jv_removed_edges(*complete_graph(pts, d))
        assert not removed & optimal, (n, sorted(removed & optimal))
        removed_total += len(removed)
    # the instances do delete edges, so the assertion above has content
    assert removed_total > 0


@pytest.mark.gpu
def test_device_session_equals_oracle_on_small_instances(ctx):
    from cudaedgeelimination_amd.resident import JvSession

    for n in range(6, 13):
        pts, d = instance(n)
        x, y, u, v, w = complete_graph(pts, d)
        proof, _ = jv_removed_edges(x, y, u, v, w)
        g = GraphSnapshot.from_edges(x, y, EUC_2D, u, v, w)
        s = JvSession(ctx, g)
        run = s.run(100)
        s.close()
        assert O.canonical_proof(run.proof) == O.canonical_proof(proof), n
