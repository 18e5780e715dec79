"""Host certification of k-opt cost matrices (ee_kopt_certify), on the CPU.

The certifier recomputes every cell by the reference CPU scorer's rules
(KOptCostTaskCpuScorer, src/cpu/kopt_search.cpp:962-1042) and compares it
with the matrix it is given, as EvaluateKOptTemplateCosts certifies a device
matrix (kopt_search.cpp:1455-1475).  Checks:

* the oracle's CPU matrix certifies, on random rows of every node pattern
  (all ports distinct, chains, heavy repetition), both distance types, near
  and wide coordinate extents; the improving masks equal the cells below
  each row's deleted cost;
* a matrix with planted wrong cells (a changed sum, a valid cell made
  INT64_MAX, an INT64_MAX cell made finite) fails with the first planted
  cell as the mismatch;
* arbitrary template tables (every perfect matching of the 2k ports, not
  only the proper ones) certify the per-cell edge-list oracle's matrix
  (tests/unit/test_kopt.cpp:48-74);
* the reference's own CPU matrix (oracle/_ref) certifies, where built.
"""
from __future__ import annotations

import itertools
import random

import numpy as np
import pytest

import _oracle as O
from conftest import require_ref
from cudaedgeelimination_amd import _lib, backend
from cudaedgeelimination_amd.graph import CEIL_2D, EUC_2D, GraphSnapshot

INVALID = np.iinfo(np.int64).max


def random_rows(rng, n, k, count):
    tasks = np.zeros(count, _lib.KOPT_TASK_DTYPE)
    for t in tasks:
        mode = rng.random()
        if mode < 0.3:
            nodes = rng.sample(range(n), 2 * k)
        elif mode < 0.6:  # a chain: consecutive deleted edges share a node
            walk = rng.sample(range(n), k + 1)
            nodes = [v for e in range(k) for v in (walk[e], walk[e + 1])]
        else:  # heavy repetition over k + 1 nodes
            pool = rng.sample(range(n), k + 1)
            nodes, seen = [], set()
            while len(nodes) < 2 * k:
                a, b = rng.sample(pool, 2)
                key = (min(a, b), max(a, b))
                if key not in seen:
                    seen.add(key)
                    nodes += [a, b]
        t["port_nodes"][:2 * k] = nodes
        t["port_nodes"][2 * k:] = rng.randrange(n)  # unused ports carry junk
        t["deleted_cost"] = rng.randrange(0, 3 * 10**6)
    return tasks


def graph(rng, n, dtype, half):
    x = np.array([rng.randrange(-half, half) for _ in range(n)], np.int64)
    y = np.array([rng.randrange(-half, half) for _ in range(n)], np.int64)
    return GraphSnapshot.from_edges(x, y, dtype, np.zeros(0, np.int32), np.zeros(0, np.int32))


def masks_of(cells, deleted, T):
    bits = cells.reshape(len(deleted), T) < deleted[:, None]
    words = np.zeros((len(deleted), (T + 31) // 32), np.uint32)
    for j in range(T):
        words[:, j // 32] |= bits[:, j].astype(np.uint32) << np.uint32(j % 32)
    return words


@pytest.mark.parametrize("half", [600_000, 1 << 30])
@pytest.mark.parametrize("dtype", [EUC_2D, CEIL_2D])
@pytest.mark.parametrize("k", [3, 4, 5])
def test_oracle_matrix_certifies(k, dtype, half):
    rng = random.Random(97 * k + 7 * dtype + half)
    g = graph(rng, 80, dtype, half)
    tasks = random_rows(rng, 80, k, 1500)
    table = backend.build_kopt_reconnect_table(k)
    cells, _, _ = O.kopt_costs(g.x, g.y, dtype, k, tasks)
    improving, count = backend.certify_kopt_costs(g, table, tasks, cells)
    T = len(table.templates)
    want = masks_of(cells, tasks["deleted_cost"], T)
    assert np.array_equal(improving, want)
    assert count == int((cells.reshape(-1, T) < tasks["deleted_cost"][:, None]).sum())
    assert (cells == INVALID).any() and (cells != INVALID).any()


@pytest.mark.parametrize("k", [3, 4, 5])
def test_planted_mismatches_are_found(k):
    rng = random.Random(500 + k)
    g = graph(rng, 60, EUC_2D, 600_000)
    tasks = random_rows(rng, 60, k, 700)
    table = backend.build_kopt_reconnect_table(k)
    cells, _, _ = O.kopt_costs(g.x, g.y, EUC_2D, k, tasks)
    valid = np.flatnonzero(cells != INVALID)
    invalid = np.flatnonzero(cells == INVALID)
    plants = [
        (int(valid[len(valid) // 2]), lambda v: v + 1),
        (int(valid[-1]), lambda v: INVALID),
        (int(invalid[len(invalid) // 3]), lambda v: 12345),
        (0, lambda v: v - 1 if v != INVALID else 0),
    ]
    for at, change in plants:
        bad = cells.copy()
        bad[at] = change(int(bad[at]))
        with pytest.raises(_lib.EeError) as err:
            backend.certify_kopt_costs(g, table, tasks, bad)
        assert err.value.status == 3 and err.value.mismatch == at
        assert f"cell={at}" in err.value.message
    # two wrong cells: the first one is reported
    bad = cells.copy()
    hi, lo = int(valid[-2]), int(valid[3])
    bad[hi] += 5
    bad[lo] += 5
    with pytest.raises(_lib.EeError) as err:
        backend.certify_kopt_costs(g, table, tasks, bad)
    assert err.value.mismatch == min(hi, lo)


def all_matchings(ports):
    if not ports:
        yield []
        return
    a = ports[0]
    for i in range(1, len(ports)):
        b = ports[i]
        rest = ports[1:i] + ports[i + 1:]
        for m in all_matchings(rest):
            yield [(a, b)] + m


@pytest.mark.parametrize("k", [3, 4, 5])
def test_arbitrary_tables_certify_edge_list_oracle(k):
    rng = random.Random(900 + k)
    g = graph(rng, 40, CEIL_2D, 5000)
    tasks = random_rows(rng, 40, k, 40)
    mates = []
    for m in itertools.islice(all_matchings(list(range(2 * k))), 0, 945):
        mate = [0] * 14
        for a, b in m:
            mate[a], mate[b] = b, a
        mates.append(mate)
    rng.shuffle(mates)
    mates = mates[:300]
    table_arr = np.zeros(len(mates), _lib.MATCHING_DTYPE)
    table_arr["endpoint_count"] = 2 * k
    table_arr["mate"] = mates
    table = backend.KOptReconnectTable(k, table_arr, 0)

    def dist(a, b):
        return O.OGraph.from_snapshot(g).distance(a, b)

    og = O.OGraph.from_snapshot(g)
    cells = np.array([O.edge_list_cost(og.distance, k, list(t["port_nodes"][:2 * k]), m)
                      for t in tasks for m in mates], np.int64)
    improving, _ = backend.certify_kopt_costs(g, table, tasks, cells)
    assert np.array_equal(improving, masks_of(cells, tasks["deleted_cost"], len(mates)))


def test_reference_matrix_certifies():
    require_ref()
    rng = random.Random(4242)
    g = graph(rng, 120, EUC_2D, 500_000)
    for k in (3, 4, 5):
        tasks = random_rows(rng, 120, k, 900)
        cells, _, _ = O.ref_kopt_costs(g.x, g.y, EUC_2D, k, tasks)
        backend.certify_kopt_costs(g, backend.build_kopt_reconnect_table(k), tasks, cells)


def test_certify_rejects_bad_input():
    rng = random.Random(7)
    g = graph(rng, 30, EUC_2D, 1000)
    table = backend.build_kopt_reconnect_table(3)
    tasks = random_rows(rng, 30, 3, 4)
    cells, _, _ = O.kopt_costs(g.x, g.y, EUC_2D, 3, tasks)
    bad = tasks.copy()
    bad["port_nodes"][1, 1] = bad["port_nodes"][1, 0]  # a deleted self loop
    with pytest.raises(_lib.InvalidArgument):
        backend.certify_kopt_costs(g, table, bad, cells)
    with pytest.raises(ValueError):
        backend.certify_kopt_costs(g, table, tasks, cells[:-1])
    empty, count = backend.certify_kopt_costs(g, table, tasks[:0], cells[:0])
    assert empty.shape == (0, 1) and count == 0
