"""The oracle's deletion-set cursor, pinned without the GPU.

Reference: KOptSearchCursor (src/cpu/kopt_search.cpp:2123-2156, 2258-2271),
AdvanceCombination (kopt_search.cpp:736-752) and BuildKOptCostTask
(kopt_search.cpp:561-584).  The reference cursor is in an unnamed namespace,
so its block sequence cannot be called from the reference build; the oracle
is pinned instead by

* an independent enumeration: itertools.combinations yields the (k - 1)-subsets
  of the indices 1 .. S-1 in lexicographic order, and AdvanceCombination
  never moves index 0 (its loop stops at position 1, kopt_search.cpp:741-746),
  so every set is index 0 plus one such subset;
* the closed form C(S - 1, k - 1) for the number of sets, and none for
  k >= S (PrepareNextK, kopt_search.cpp:2260-2262);
* the task rule on every row: ports (tour[p], tour[(p + 1) mod L]) per
  sorted position p and deleted_cost the exact sum;
* BuildKOptCostTask's failures (a negative cached cost, an int64 overflow).

A planted mistake (index 0 allowed to move, positions left unsorted, the
wrap at the tour's end dropped) fails the first test.
"""
from __future__ import annotations

import ctypes as C
import itertools
import math

import numpy as np
import pytest

import _oracle as O
from cudaedgeelimination_amd import _lib, backend, workloads as W


def lex_rank(labels, n):
    """Rank of the increasing label tuple among the (len)-subsets of 1..n in
    lexicographic order (combinatorial number system)."""
    rank, prev, r = 0, 0, len(labels)
    for s, v in enumerate(labels):
        rem = r - 1 - s
        for u in range(prev + 1, v):
            rank += math.comb(n - u, rem)
        prev = v
    return rank


def expected_rows(tour, costs, selectable, k):
    L, S = len(tour), len(selectable)
    if k >= S:
        return [], []
    tasks, positions = [], []
    for tail in itertools.combinations(range(1, S), k - 1):
        pos = sorted(int(selectable[i]) for i in (0,) + tail)
        ports = []
        for p in pos:
            ports += [int(tour[p]), int(tour[(p + 1) % L])]
        tasks.append((ports + [0] * (10 - len(ports)), sum(int(costs[p]) for p in pos)))
        positions.append(pos)
    return tasks, positions


@pytest.fixture(scope="module")
def graph():
    return W.config1()


@pytest.mark.parametrize("length,selectable,seed", [(7, 4, 1), (9, 6, 2), (14, 9, 3), (20, 20, 4),
                                                    (16, 12, 5)])
def test_cursor_matches_enumeration(graph, length, selectable, seed):
    t = W.tour_context(seed, graph, length, selectable)
    for k in (3, 4, 5):
        want_tasks, want_pos = expected_rows(t.tour, t.edge_costs, t.selectable, k)
        expect = math.comb(selectable - 1, k - 1) if k < selectable else 0
        assert len(want_tasks) == expect
        tasks, pos = O.cursor_tasks(t.tour, t.edge_costs, t.selectable, k, 0, 10**6)
        assert len(tasks) == expect
        for i, (ports, cost) in enumerate(want_tasks):
            assert list(tasks["port_nodes"][i]) == ports
            assert int(tasks["deleted_cost"][i]) == cost
            assert list(pos[i]) == want_pos[i]
        # windows of the same sequence, including ones that run off its end
        for first, count in ((0, 1), (1, 3), (expect // 2, 5), (max(expect - 2, 0), 10)):
            wt, wp = O.cursor_tasks(t.tour, t.edge_costs, t.selectable, k, first, count)
            assert len(wt) == max(0, min(count, expect - first))
            assert wt.tobytes() == tasks[first:first + len(wt)].tobytes()
            assert wp.tobytes() == pos[first:first + len(wp)].tobytes()


def test_every_set_holds_the_required_edge(graph):
    t = W.tour_context(11, graph, 30, 12)
    req = int(t.selectable[0])
    for k in (3, 4, 5):
        _, pos = O.cursor_tasks(t.tour, t.edge_costs, t.selectable, k, 0, 10**6)
        assert all(req in row for row in pos.tolist())


def test_lex_rank_counts_the_enumeration():
    # the ranking the GPU tests use at sizes the oracle cannot walk
    for n, r in ((8, 2), (9, 3), (10, 4)):
        ranks = [lex_rank(c, n) for c in itertools.combinations(range(1, n + 1), r)]
        assert ranks == list(range(math.comb(n, r)))


def test_cursor_failures(graph):
    t = W.tour_context(7, graph, 12, 8)
    bad = t.edge_costs.copy()
    bad[int(t.selectable[3])] = -1
    with pytest.raises(RuntimeError):
        O.cursor_tasks(t.tour, bad, t.selectable, 3, 0, 100)
    big = t.edge_costs.copy()
    big[t.selectable.astype(np.int64)] = np.iinfo(np.int64).max // 2
    with pytest.raises(RuntimeError):
        O.cursor_tasks(t.tour, big, t.selectable, 3, 0, 1)


def test_library_set_count_matches_closed_form(graph):
    # host-side arithmetic of the native library (no device call)
    for S in (0, 1, 3, 4, 5, 6, 50, 4000, 65536):
        t = backend.TourContext(np.arange(max(S, 3), dtype=np.int32),
                                np.zeros(max(S, 3), dtype=np.int64), np.arange(S, dtype=np.uint64))
        for k in (2, 3, 4, 5, 6):
            expect = math.comb(S - 1, k - 1) if 3 <= k <= 5 and k < S else 0
            assert backend.kopt_cursor_count(t, k) == expect
    t = backend.TourContext(np.arange(3, dtype=np.int32), np.zeros(3, dtype=np.int64),
                            np.arange(65537, dtype=np.uint64))
    with pytest.raises(_lib.InvalidArgument):
        backend.kopt_cursor_count(t, 3)
