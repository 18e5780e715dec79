"""Pins the CPU oracle against what the reference itself fixes and against
mathematics, independently of the oracle's own formulas.

* distance known answers: tests/unit/test_core.cpp:44-66
* template counts and generator hashes: configs/kopt_reconnect_hashes_v1.txt:3-5,
  tests/unit/test_kopt.cpp:330-352
* the reference fixture's protected tour costs 395: tests/CMakeLists.txt:13-19
* isqrt against Python's exact math.isqrt, exhaustively on small radicands and
  on near-square and random 64-bit radicands
* template structure: brute-force enumeration of proper reconnections
* k-opt cells against an independent per-cell edge-list oracle
  (tests/unit/test_kopt.cpp:48-74)
"""
from __future__ import annotations

import ctypes as C
import itertools
import math
import pathlib
import random

import numpy as np
import pytest

import _oracle as O

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"
INT64_MAX = (1 << 63) - 1


def odist(ax, ay, bx, by, t):
    out = C.c_int64()
    assert O.oracle().orc_exact_distance(ax, ay, bx, by, t, C.byref(out)) == 0
    return out.value


def test_distance_known_answers():
    assert odist(0, 0, 3, 4, 0) == 5
    assert odist(0, 0, 1, 1, 0) == 1
    assert odist(0, 0, 1, 1, 1) == 2
    lib = O.oracle()
    assert lib.orc_isqrt_floor(0) == 0
    assert lib.orc_isqrt_floor(15) == 3
    assert lib.orc_isqrt_floor(16) == 4


def test_isqrt_matches_math_isqrt():
    lib = O.oracle()
    for s in range(0, 1 << 16):
        assert lib.orc_isqrt_floor(s) == math.isqrt(s)
    rng = random.Random(7)
    for _ in range(20000):
        r = rng.randrange(1 << 32)
        for s in (r * r - 1, r * r, r * r + 1, r * r + r, r * r + r + 1, rng.randrange(1 << 64)):
            if 0 <= s < (1 << 64):
                assert lib.orc_isqrt_floor(s) == math.isqrt(s)
    assert lib.orc_isqrt_floor((1 << 64) - 1) == (1 << 32) - 1


def test_rounding_rules_match_real_square_root():
    # EUC_2D = nearest integer of sqrt(S) (never a tie for integer S),
    # CEIL_2D = ceiling; checked with exact integer arithmetic.
    rng = random.Random(11)
    for _ in range(5000):
        dx, dy = rng.randrange(1 << 31), rng.randrange(1 << 31)
        s = dx * dx + dy * dy
        r = math.isqrt(s)
        euc = r + 1 if 4 * s >= (2 * r + 1) ** 2 else r
        ceil = r if r * r == s else r + 1
        assert odist(0, 0, dx, dy, 0) == euc
        assert odist(0, 0, dx, dy, 1) == ceil


def test_distance_overflow_is_rejected():
    out = C.c_int64()
    assert O.oracle().orc_exact_distance(-(1 << 62), 0, 1 << 62, 1 << 62, 0, C.byref(out)) != 0


PINNED_TEMPLATES = {3: (4, 0xE58AF5E08D290D04), 4: (25, 0x03179E3CA191CE82),
                    5: (208, 0x6696DDE548591BCE)}


def test_template_counts_and_pinned_hashes():
    for k, (count, h) in PINNED_TEMPLATES.items():
        mates, got = O.templates(k)
        assert len(mates) == count
        assert got == h


def proper_reconnections(k):
    """Brute force: perfect matchings of 2k ports that re-add no deleted
    edge (2e, 2e+1) and, together with the tour's path segments
    (2e+1)-(2e+2) and (2k-1)-0, form a single cycle through every port."""
    ends = 2 * k
    result = []

    def matchings(free):
        if not free:
            yield []
            return
        a = free[0]
        for i in range(1, len(free)):
            rest = free[1:i] + free[i + 1:]
            for m in matchings(rest):
                yield [(a, free[i])] + m

    for m in matchings(list(range(ends))):
        mate = [0] * ends
        for a, b in m:
            mate[a], mate[b] = b, a
        if any(mate[2 * e] == 2 * e + 1 for e in range(k)):
            continue
        seg = {}
        for e in range(k - 1):
            seg[2 * e + 1], seg[2 * e + 2] = 2 * e + 2, 2 * e + 1
        seg[2 * k - 1], seg[0] = 0, 2 * k - 1
        seen, at, steps = set(), 0, 0
        while at not in seen:
            seen.add(at)
            at = mate[at] if steps % 2 == 0 else seg[at]
            steps += 1
        if steps == ends and at == 0:
            result.append(tuple(mate))
    return result


@pytest.mark.parametrize("k", [3, 4, 5])
def test_templates_equal_brute_force_in_order(k):
    mates, _ = O.templates(k)
    assert [tuple(row[:2 * k]) for row in mates] == proper_reconnections(k)


def test_reference_fixture_tour_cost_is_395():
    g = O.OGraph.load(GOLDEN / "recursive-point" / "recursive-point.tsp",
                      GOLDEN / "recursive-point" / "recursive-point.edg")
    tour = [int(t) - 1 for t in
            (GOLDEN / "recursive-point" / "recursive-point.tour").read_text().split("TOUR_SECTION")[1].split()
            if t.strip() not in ("-1", "EOF")]
    cost = sum(g.distance(tour[i], tour[(i + 1) % len(tour)]) for i in range(len(tour)))
    assert cost == 395


def seventeen():
    x = np.array([(7 * i) % 23 for i in range(17)], np.int64)
    y = np.array([(i * i + 3 * i) % 19 for i in range(17)], np.int64)
    return x, y


@pytest.mark.parametrize("dtype", [0, 1])
@pytest.mark.parametrize("k", [3, 4, 5])
def test_kopt_cells_match_edge_list_oracle(k, dtype):
    x, y = seventeen()
    rng = random.Random(100 * k + dtype)
    tasks = np.zeros(60, O.TASK_DTYPE)
    for t in tasks:
        # mix of distinct ports and shared-endpoint chains (repeated nodes)
        if rng.random() < 0.5:
            nodes = rng.sample(range(17), 2 * k)
        else:
            walk = rng.sample(range(17), k + 1)
            nodes = [v for e in range(k) for v in (walk[e], walk[e + 1])]
        t["port_nodes"][:2 * k] = nodes
    costs, _, _ = O.kopt_costs(x, y, dtype, k, tasks)
    mates, _ = O.templates(k)
    T = len(mates)
    dist = lambda a, b: odist(x[a], y[a], x[b], y[b], dtype)  # noqa: E731
    for r, t in enumerate(tasks):
        for j, mate in enumerate(mates):
            assert costs[r * T + j] == O.edge_list_cost(dist, k, list(t["port_nodes"]), list(mate))


def test_kopt_rejects_invalid_tasks():
    x, y = seventeen()
    bad = np.zeros(1, O.TASK_DTYPE)
    bad["port_nodes"][0, :6] = [0, 1, 1, 0, 2, 3]  # deleted edge (0,1) twice
    with pytest.raises(O.OracleError):
        O.kopt_costs(x, y, 0, 3, bad)
    bad["port_nodes"][0, :6] = [0, 0, 1, 2, 3, 4]  # self loop
    with pytest.raises(O.OracleError):
        O.kopt_costs(x, y, 0, 3, bad)
    bad["port_nodes"][0, :6] = [0, 1, 2, 3, 4, 5]
    bad["deleted_cost"] = -1
    with pytest.raises(O.OracleError):
        O.kopt_costs(x, y, 0, 3, bad)


def test_kopt_parallel_policy_and_batch_cache():
    # tests/unit/test_kopt.cpp:240-305: 40 identical 5-opt rows (8320 cells)
    # use the 10-node batch table; 52 rows over 520 nodes exceed its 512 cap.
    x, y = seventeen()
    rep = np.zeros(40, O.TASK_DTYPE)
    rep["port_nodes"] = np.arange(10)
    costs, threads, nodes = O.kopt_costs(x, y, 0, 5, rep)
    assert nodes == 10 and 1 <= threads <= 8
    assert all((costs[i * 208:(i + 1) * 208] == costs[:208]).all() for i in range(40))
    bx = np.arange(520, dtype=np.int64)
    by = bx % 17
    big = np.zeros(52, O.TASK_DTYPE)
    big["port_nodes"] = np.arange(520).reshape(52, 10)
    _, _, nodes = O.kopt_costs(bx, by, 0, 5, big)
    assert nodes == 0
