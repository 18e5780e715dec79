"""The snapshot content hash evaluated on the device.

`ee_graph_content_hash_hip` splits the FNV-1a byte stream of the edge slots
into chunks, composes the permutations of the hash's low byte to find each
chunk's starting low byte, and sums each chunk's weighted contribution
mod 2^64.  It must equal the byte-serial hash `ee_graph_content_hash`
(src/common/graph.cpp:323-346, itself pinned to the reference's snapshot
hashes by the oracle tests) for every slot count, including those that cross
the chunk (8 slots), group and chunk-size (524,288 slots) boundaries, with
inactive slots, large weights and both distance types.
"""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest

from cudaedgeelimination_amd import _lib
from cudaedgeelimination_amd.graph import CEIL_2D, EUC_2D

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def hashes(ctx, n, dtype, x, y, u, v, w, active):
    p = _lib.ptr
    want = int(_lib.lib().ee_graph_content_hash(n, dtype, p(x), p(y), len(u), p(u), p(v), p(w), p(active)))
    got = C.c_uint64()
    _lib.check(_lib.lib().ee_graph_content_hash_hip(ctx.handle, n, dtype, p(x), p(y), len(u), p(u), p(v), p(w),
                                                    p(active), C.byref(got)))
    return want, got.value


@pytest.mark.parametrize("m", [0, 1, 2, 7, 8, 9, 65, 1000, 16203, 85710, 524288, 524289, 600001])
def test_device_hash_equals_byte_serial_hash(ctx, m):
    say(f"device content hash m={m}")
    rng = np.random.default_rng(m + 11)
    n = 5000
    x = rng.integers(-(1 << 40), 1 << 40, n).astype(np.int64)
    y = rng.integers(0, 1 << 20, n).astype(np.int64)
    u = rng.integers(0, n, m).astype(np.int32)
    v = rng.integers(0, n, m).astype(np.int32)
    w = (rng.integers(0, 1 << 62, m, dtype=np.int64) >> rng.integers(0, 60, m)).astype(np.int64)
    active = (rng.random(m) < 0.7).astype(np.uint8)
    for dtype in (EUC_2D, CEIL_2D):
        want, got = hashes(ctx, n, dtype, x, y, u, v, w, active)
        assert got == want, (m, dtype)


def test_device_hash_all_active_and_all_inactive(ctx):
    rng = np.random.default_rng(3)
    n, m = 300, 4097
    x = rng.integers(0, 10000, n).astype(np.int64)
    y = rng.integers(0, 10000, n).astype(np.int64)
    u = rng.integers(0, n, m).astype(np.int32)
    v = rng.integers(0, n, m).astype(np.int32)
    w = rng.integers(0, 20000, m).astype(np.int64)
    for active in (np.ones(m, np.uint8), np.zeros(m, np.uint8)):
        want, got = hashes(ctx, n, EUC_2D, x, y, u, v, w, active)
        assert got == want
