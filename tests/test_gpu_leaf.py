"""Certified leaf pipeline (ee_leaf_*) on the device.

The pipeline keeps up to `depth` blocks of task rows in flight on the device
while the host certifies the oldest one, as the reference's leaf loop
certifies each device matrix against a full CPU matrix
(src/cpu/kopt_search.cpp:1455-1475, 2651-2764).  Bit-exact checks:

* host-row blocks of every k at depths 1-4: every block comes back in
  submission order with cells equal to the oracle's CPU matrix and improving
  masks equal to the cells below each row's deleted cost;
* cursor blocks (rows built on the device): rows and positions equal the
  oracle cursor's, cells the oracle's matrix, and both equal the plain
  ee_kopt_cursor_eval outputs; host and cursor blocks interleave;
* the reference's own CPU matrix (oracle/_ref) equals the certified cells;
* failures: a deleted-cost overflow of a cursor block fails its ee_leaf_next,
  too many outstanding blocks, nothing outstanding, invalid rows, cursor
  blocks without a tour; the pipeline keeps working afterwards.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import _oracle as O
from conftest import require_ref
from cudaedgeelimination_amd import _lib, backend, workloads as W
from cudaedgeelimination_amd.resident import LeafPipeline

pytestmark = pytest.mark.gpu

T_OF = {3: 4, 4: 25, 5: 208}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def wl():
    return W.config3(rows=6000)


def masks_of(cells, deleted, T):
    bits = cells.reshape(len(deleted), T) < deleted[:, None]
    words = np.zeros((len(deleted), (T + 31) // 32), np.uint32)
    for j in range(T):
        words[:, j // 32] |= bits[:, j].astype(np.uint32) << np.uint32(j % 32)
    return words


def check_block(g, b, tasks):
    T = T_OF[b.k]
    assert b.tasks.tobytes() == np.ascontiguousarray(tasks).tobytes()
    cells, _, _ = O.kopt_costs(g.x, g.y, g.distance_type, b.k, tasks)
    assert np.array_equal(b.costs.reshape(-1), cells)
    want = masks_of(cells, tasks["deleted_cost"], T)
    assert np.array_equal(b.improving, want)
    assert b.improving_cells == int(sum(bin(int(w)).count("1") for w in want.reshape(-1)))


@pytest.mark.parametrize("depth", [1, 2, 3, 4])
def test_host_blocks_in_order_equal_oracle(ctx, wl, depth):
    say(f"leaf host blocks depth={depth}")
    g = wl.graph
    blocks = []
    for k in (3, 4, 5):
        rows = wl.bucket(k)
        for start in range(0, len(rows), 700):
            blocks.append((k, rows[start:start + 700]))
    p = LeafPipeline(ctx, g, depth=depth, max_rows=1024)
    submitted = 0
    for i in range(len(blocks)):
        while submitted < len(blocks) and submitted < i + depth:
            k, rows = blocks[submitted]
            p.submit_tasks(k, rows, tag=1000 + submitted)
            submitted += 1
        b = p.next()
        assert b.tag == 1000 + i and b.k == blocks[i][0] and b.positions is None
        check_block(g, b, blocks[i][1])
    st = p.stats()
    assert st["blocks"] == len(blocks) and st["rows"] == 6000
    assert st["cells"] == wl.cells() and st["certify_threads"] >= 1
    p.close()


def test_cursor_blocks_equal_oracle_and_plain_path(ctx, wl):
    say("leaf cursor blocks")
    g = wl.graph
    t = W.tour_context(5, g, 400, 120)
    tables = {k: backend.build_kopt_reconnect_table(k) for k in (3, 4, 5)}
    p = LeafPipeline(ctx, g, depth=3, max_rows=4096)
    p.set_tour(t)
    plan = [(3, 0), (3, 4096), (4, 0), (4, 4096 * 20), (5, 0), (5, 4096 * 1000), (5, 9_000_000 - 100)]
    total = {k: backend.kopt_cursor_count(t, k) for k in (3, 4, 5)}
    queue = []
    for i, (k, first) in enumerate(plan):
        rows = p.submit_cursor(k, first, 4096, tag=i)
        assert rows == min(4096, max(total[k] - first, 0))
        if rows:
            queue.append((i, k, first, rows))
        if len(queue) == 3 or i == len(plan) - 1:
            while queue:
                tag, k, first, rows = queue.pop(0)
                b = p.next()
                assert (b.tag, b.k, b.first) == (tag, k, first) and len(b.tasks) == rows
                want_t, want_p = O.cursor_tasks(t.tour, t.edge_costs, t.selectable, k, first, rows)
                assert np.array_equal(b.positions.astype(np.uint64), want_p)
                check_block(g, b, want_t)
                plain = backend.evaluate_kopt_cursor_cuda(g, tables[k], t, k, first, rows, ctx=ctx)
                assert plain.tasks.tobytes() == b.tasks.tobytes()
                assert np.array_equal(plain.costs, b.costs.reshape(-1))
    p.close()


def test_host_and_cursor_blocks_interleave(ctx, wl):
    say("leaf interleaved blocks")
    g = wl.graph
    t = W.tour_context(11, g, 300, 90)
    p = LeafPipeline(ctx, g, depth=4, max_rows=2048)
    p.set_tour(t)
    host = wl.bucket(4)[:2048]
    expect = []
    for i in range(12):
        if i % 2:
            p.submit_tasks(4, host[:100 + 150 * i], tag=i)
            expect.append(("host", host[:100 + 150 * i]))
        else:
            k = 3 + (i // 2) % 3
            first = 37 * i
            rows = p.submit_cursor(k, first, 2048, tag=i)
            want_t, _ = O.cursor_tasks(t.tour, t.edge_costs, t.selectable, k, first, rows)
            expect.append(("cursor", want_t))
        if i >= 3:
            b = p.next()
            kind, want = expect[b.tag]
            check_block(g, b, want)
            assert (b.positions is None) == (kind == "host")
    for _ in range(3):
        b = p.next()
        check_block(g, b, expect[b.tag][1])
    p.close()


def test_reference_cpu_matrix_equals_certified_cells(ctx, wl):
    require_ref()
    say("leaf vs reference CPU matrix")
    g = wl.graph
    p = LeafPipeline(ctx, g, depth=2, max_rows=2048)
    for k in (3, 4, 5):
        rows = wl.bucket(k)[:2048]
        p.submit_tasks(k, rows, tag=k)
        b = p.next()
        ref, _, _ = O.ref_kopt_costs(g.x, g.y, g.distance_type, k, rows)
        assert np.array_equal(b.costs.reshape(-1), ref)
    p.close()


def test_failures_leave_the_pipeline_usable(ctx, wl):
    say("leaf failures")
    g = wl.graph
    t = W.tour_context(9, g, 40, 20)
    p = LeafPipeline(ctx, g, depth=2, max_rows=256)
    with pytest.raises(_lib.EeError) as err:  # no tour yet
        p.submit_cursor(3, 0, 10)
    assert err.value.status == 4
    with pytest.raises(_lib.EeError) as err:  # nothing outstanding
        p.next()
    assert err.value.status == 4
    rows = wl.bucket(3)[:200]
    bad = rows.copy()
    bad["port_nodes"][7, 0] = g.dimension
    with pytest.raises(_lib.InvalidArgument):
        p.submit_tasks(3, bad)
    with pytest.raises(_lib.InvalidArgument):
        p.submit_tasks(3, wl.bucket(3)[:257])  # beyond max_rows
    p.submit_tasks(3, rows, tag=1)
    p.submit_tasks(3, rows, tag=2)
    with pytest.raises(_lib.EeError) as err:  # depth 2 reached
        p.submit_tasks(3, rows, tag=3)
    assert err.value.status == 4
    assert p.next().tag == 1 and p.next().tag == 2
    # an overflowing deleted cost fails its block, and only its block
    big = t.edge_costs.copy()
    big[t.selectable.astype(np.int64)] = np.iinfo(np.int64).max // 2
    p.set_tour(backend.TourContext(t.tour, big, t.selectable))
    p.submit_cursor(3, 0, 64, tag=7)
    p.submit_tasks(3, rows, tag=8)
    with pytest.raises(_lib.EeError) as err:
        p.next()
    assert err.value.status == 3 and "overflow" in err.value.message
    b = p.next()
    assert b.tag == 8
    check_block(g, b, rows)
    p.set_tour(t)
    assert p.submit_cursor(3, 0, 64, tag=9) == 64
    b = p.next()
    want_t, _ = O.cursor_tasks(t.tour, t.edge_costs, t.selectable, 3, 0, 64)
    check_block(g, b, want_t)
    st = p.stats()
    assert st["blocks"] == 4  # the failed block is not counted
    p.close()


def test_corrupted_cell_fails_certification_and_only_its_block(ctx, wl):
    """A cell of a device block corrupted on the host before certification
    (the fail-closed path of the reference's certify-every-cell contract):
    ee_leaf_next refuses that block naming the cell; the next block, and a
    corrupted value equal to the true one, pass."""
    say("leaf certification fault")
    g = wl.graph
    p = LeafPipeline(ctx, g, depth=2, max_rows=256)
    rows = wl.bucket(5)[:100]
    cells, _, _ = O.kopt_costs(g.x, g.y, g.distance_type, 5, rows)
    cell = 37 * T_OF[5] + 101
    p.submit_tasks(5, rows, tag=1)
    p.submit_tasks(5, rows, tag=2)
    p.inject_cell(cell, int(cells[cell]) + 1)
    with pytest.raises(_lib.EeError) as err:
        p.next()
    assert err.value.status == 3 and "certification" in err.value.message
    assert f"cell={cell}" in err.value.message
    b = p.next()
    assert b.tag == 2
    check_block(g, b, rows)
    # the true value is no fault
    p.submit_tasks(5, rows, tag=3)
    p.inject_cell(cell, int(cells[cell]))
    check_block(g, p.next(), rows)
    p.close()
