"""Host-side parts of the product library, on the CPU: the C ABI loads and
exports every declared symbol, the FP64-root-plus-correction square root is
exact, CSR / content hash / template generator equal the oracle's, the
seeded generator is deterministic and pinned, and device entry points fail
loudly without a device."""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import math
import pathlib
import random

import numpy as np
import pytest

import _oracle as O
from cudaedgeelimination_amd import _lib, backend, workloads as W
from cudaedgeelimination_amd.graph import CEIL_2D, EUC_2D, GraphSnapshot, exact_distances

ROOT = pathlib.Path(__file__).resolve().parent.parent
G = json.loads((ROOT / "tests" / "golden" / "goldens.json").read_text())


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    declared = []
    for header in sorted((ROOT / "include").glob("*.h")):
        declared += O.header_functions(header)
    assert declared, "no declarations found"
    missing = [name for name in declared if not hasattr(lib, name)]
    assert not missing, missing
    assert set(declared) <= set(_lib.PROTOTYPES), "bindings lag the headers"


def test_host_isqrt_is_exact():
    s = np.arange(0, 1 << 22, dtype=np.uint64)
    rng = random.Random(3)
    extra = []
    for _ in range(50000):
        r = rng.randrange(1 << 32)
        extra += [r * r - 1 if r else 0, r * r, (r * r + 2 * r) % (1 << 64), rng.randrange(1 << 64)]
    extra += [(1 << 64) - 1, (1 << 64) - 2, ((1 << 32) - 1) ** 2, (1 << 53) - 1, (1 << 53) + 1]
    s = np.concatenate([s, np.array(extra, dtype=np.uint64)])
    out = np.empty_like(s)
    _lib.lib().ee_isqrt_host(O.p(s), O.p(out), len(s))
    expect = np.array([math.isqrt(int(v)) for v in s[-len(extra):]], dtype=np.uint64)
    assert np.array_equal(out[-len(extra):], expect)
    small = np.floor(np.sqrt(s[:1 << 22].astype(np.float64))).astype(np.uint64)
    assert np.array_equal(out[:1 << 22], small)  # exact for s < 2^22 (floats exact there)


def test_vectorised_distances_match_oracle():
    rng = np.random.default_rng(5)
    ax, ay, bx, by = (rng.integers(-10**9, 10**9, 20000) for _ in range(4))
    for t in (EUC_2D, CEIL_2D):
        got = exact_distances(ax, ay, bx, by, t)
        lib = O.oracle()
        out = C.c_int64()
        for i in range(0, 20000, 97):
            assert lib.orc_exact_distance(int(ax[i]), int(ay[i]), int(bx[i]), int(by[i]), t,
                                          C.byref(out)) == 0
            assert got[i] == out.value


def test_templates_equal_pins():
    for k, (count, h) in {3: (4, 0xE58AF5E08D290D04), 4: (25, 0x03179E3CA191CE82),
                          5: (208, 0x6696DDE548591BCE)}.items():
        table = backend.build_kopt_reconnect_table(k)
        assert len(table.templates) == count and table.generator_hash == h
        mates, _ = O.templates(k)
        assert np.array_equal(table.templates["mate"], mates)
        assert (table.templates["endpoint_count"] == 2 * k).all()


@pytest.mark.parametrize("canonical", [True, False])
def test_csr_and_hash_equal_oracle(canonical):
    g = W.config1()
    u, v, w = g.edge_u.copy(), g.edge_v.copy(), g.edge_weight.copy()
    if not canonical:  # reversed order and swapped endpoints: the sort fallback
        u, v, w = v[::-1].copy(), u[::-1].copy(), w[::-1].copy()
    active = (np.arange(len(u)) % 7 != 3).astype(np.uint8)
    ours = GraphSnapshot.from_edges(g.x, g.y, EUC_2D, u, v, w, active)
    og = O.OGraph.from_arrays(g.x, g.y, EUC_2D, u, v, w, active)
    arr = og.arrays()
    for key in ("row_offsets", "neighbors", "csr_edge_ids"):
        assert np.array_equal(getattr(ours, key), arr[key]), key
    assert ours.content_hash() == og.hash()


def test_generated_inputs_are_pinned():
    assert f"{W.config1().content_hash():016x}" == G["config1_first_epoch"]["content_hash"]
    assert f"{W.config2().content_hash():016x}" == G["config2_first_epoch"]["content_hash"]
    wl = W.config3(seed=0xEE0000C3, n=2000, rows=4096)
    assert sha(wl.tasks) == G["config3_small"]["tasks_sha256"]
    assert sha(wl.ks) == G["config3_small"]["ks_sha256"]


def test_knn_graph_matches_brute_force():
    rng = np.random.default_rng(9)
    for trial in range(4):
        n, k = 120, 5 + trial
        # a coarse grid forces many exact distance ties
        x = rng.integers(0, 12, n).astype(np.int64) * 10
        y = rng.integers(0, 12, n).astype(np.int64) * 10
        x += np.arange(n) % 3  # keep points distinct
        g = W.knn_graph(x, y, EUC_2D, k)
        expect = set()
        for p in range(n):
            s = (x - x[p]) ** 2 + (y - y[p]) ** 2
            order = sorted((int(s[q]), q) for q in range(n) if q != p)[:k]
            for _, q in order:
                expect.add((min(p, q), max(p, q)))
        assert set(zip(g.edge_u.tolist(), g.edge_v.tolist())) == expect
        assert np.array_equal(g.edge_weight, exact_distances(x[g.edge_u], y[g.edge_u],
                                                             x[g.edge_v], y[g.edge_v], EUC_2D))


def test_generated_tasks_are_valid_walk_rows():
    wl = W.config3(seed=1234, n=3000, rows=2000)
    g = wl.graph
    edges = set(zip(g.edge_u.tolist(), g.edge_v.tolist()))
    for t, k in zip(wl.tasks, wl.ks):
        ports = t["port_nodes"]
        assert (ports[2 * k:] == 0).all()
        pairs = [(min(ports[2 * e], ports[2 * e + 1]), max(ports[2 * e], ports[2 * e + 1]))
                 for e in range(k)]
        assert len(set(pairs)) == k and all(p in edges for p in pairs)
        assert t["deleted_cost"] == sum(int(exact_distances(g.x[a], g.y[a], g.x[b], g.y[b], EUC_2D))
                                        for a, b in pairs)
    assert [int((wl.ks == k).sum()) for k in (3, 4, 5)] == [1000, 600, 400]


def test_config3_shape():
    assert W.kopt_quota(65536) == (32768, 19661, 13107)


def test_device_entry_points_fail_loudly_without_device():
    if _lib.lib().ee_device_count() > 0:
        pytest.skip("a HIP device is visible")
    handle = C.c_void_p()
    status = _lib.lib().ee_ctx_create(-1, C.byref(handle))
    assert status == 5  # EE_ERR_UNAVAILABLE
    with pytest.raises(_lib.Unavailable):
        backend.Context(-1)
    ok, reason = backend.cuda_backend_available()
    assert not ok and reason


def test_missing_library_raises(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "LIBRARY_PATH", tmp_path / "absent.so")
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(_lib.BackendMissing):
        _lib.lib()


def test_shim_rejects_malformed_ht_snapshots_before_any_device_call():
    """The drop-in shim's HT reply symbols throw std::invalid_argument for a
    snapshot whose containers disagree with its dimension or CSR (the
    reference's ValidateGraphShape, src/cuda/ht_replies.cu:521-531), ahead of
    device selection, so no GPU is needed (oracle/shim_shape_check.cpp)."""
    import subprocess

    exe = ROOT / "oracle" / "_ref" / "shim_shape_check"
    if not exe.exists():
        pytest.skip("the drop-in build (oracle/_ref) is not present")
    fixture = ROOT / "tests" / "golden" / "recursive-point"
    run = subprocess.run([str(exe), str(fixture / "recursive-point.tsp"),
                          str(fixture / "recursive-point.edg")], capture_output=True, text=True,
                         timeout=60)
    assert run.returncode == 0, run.stdout + run.stderr
    assert run.stdout.count("ok ") == 12


def test_host_units_under_address_and_undefined_behaviour_sanitizers():
    """The host-only units (generator, CSR and hash, template generator, the
    JV host run, the certifier) with ASan and UBSan, checked against the
    oracle (oracle/sanitize_harness.cpp; built by `make -C oracle sanitize`)."""
    import os
    import subprocess

    harness = ROOT / "oracle" / "_san" / "sanitize_harness"
    if not harness.exists():
        pytest.skip("sanitizer build not present (make -C oracle sanitize)")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    done = subprocess.run([str(harness)], capture_output=True, text=True, timeout=300, env=env)
    assert done.returncode == 0, done.stdout[-3000:] + done.stderr[-3000:]
    assert "sanitized host checks passed" in done.stdout
    assert "runtime error" not in done.stderr and "AddressSanitizer" not in done.stderr
