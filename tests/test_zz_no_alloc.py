"""ee_kopt_eval_costs_noalloc: the plain k-opt evaluation staged in the
device's fixed scratch array, with no allocation.  Its cells must equal
ee_kopt_eval_costs's bit for bit, call after call on the shared array (the
second call smaller than the first, on another graph, k and distance path),
and a call whose staging exceeds the array must fail before any launch."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from cudaedgeelimination_amd import _lib, backend
from cudaedgeelimination_amd import workloads as W
from cudaedgeelimination_amd.graph import CEIL_2D, EUC_2D

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 32 << 20  # EE_NOALLOC_SCRATCH_BYTES


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def case(seed, n, k, rows, distance_type, hi):
    x, y = W._points_uniform(seed, n, 0, hi)
    g = W.knn_graph(x, y, distance_type, 8)
    tasks = W.kopt_tasks(seed + 1, g, np.full(rows, k, np.uint8))
    return g, backend.build_kopt_reconnect_table(k), tasks


def test_noalloc_equals_plain_path_on_consecutive_calls(ctx):
    # first: k = 5 leaf batch of 4,096 rows on the FP64 distance path
    # (extent 10^6); second: fewer rows, k = 3, CEIL_2D, on the 64-bit integer
    # path (extent 2^30), so every region it reads was last written by the
    # first call with other contents
    cases = [(0xA110C1, 3000, 5, 4096, EUC_2D, 1_000_000),
             (0xA110C2, 700, 3, 555, CEIL_2D, 1 << 30)]
    for seed, n, k, rows, dtype, hi in cases:
        say(f"no-alloc k={k} rows={rows} n={n} dtype={dtype}")
        g, table, tasks = case(seed, n, k, rows, dtype, hi)
        got, dev, usage = backend.evaluate_kopt_template_costs_cuda(g, table, tasks, ctx, no_alloc=True)
        want, _, _ = backend.evaluate_kopt_template_costs_cuda(g, table, tasks, ctx)
        assert dev == ctx.device
        assert got.shape == want.shape == (rows * len(table.templates),)
        assert np.array_equal(got, want)
        assert (got != np.iinfo(np.int64).max).any()
        assert not usage.snapshot_hit and not usage.template_hit and not usage.workspace_hit
        assert 0 < usage.resident_bytes <= SCRATCH_BYTES


def test_noalloc_rejects_work_beyond_the_scratch(ctx):
    g, table, tasks = case(0xA110C3, 500, 5, 64, EUC_2D, 1_000_000)
    # 208 cells x 8 B + 48 B per row: 20,000 rows need about 34 MB
    big = np.resize(tasks, 20_000)
    with pytest.raises(_lib.InvalidArgument, match="scratch"):
        backend.evaluate_kopt_template_costs_cuda(g, table, big, ctx, no_alloc=True)
    # the array is still usable after the refusal
    got, _, _ = backend.evaluate_kopt_template_costs_cuda(g, table, tasks, ctx, no_alloc=True)
    want, _, _ = backend.evaluate_kopt_template_costs_cuda(g, table, tasks, ctx)
    assert np.array_equal(got, want)


def test_noalloc_empty_and_invalid_input(ctx):
    g, table, tasks = case(0xA110C4, 300, 4, 8, EUC_2D, 1_000_000)
    got, _, usage = backend.evaluate_kopt_template_costs_cuda(g, table, tasks[:0], ctx, no_alloc=True)
    assert got.size == 0 and usage.resident_bytes == 0
    bad = tasks.copy()
    bad["port_nodes"][0, 0] = g.dimension  # out of range port
    with pytest.raises(_lib.InvalidArgument):
        backend.evaluate_kopt_template_costs_cuda(g, table, bad, ctx, no_alloc=True)
