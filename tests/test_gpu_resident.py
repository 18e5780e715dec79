"""Resident forms on the device: multi-segment k-opt batches (every tuning,
graph replay, many instances) and the multi-epoch JV session, against the
plain path, the CPU oracle and the reference-generated fixtures."""
from __future__ import annotations

import hashlib
import json
import pathlib
import sys

import numpy as np
import pytest

import _oracle as O
from cudaedgeelimination_amd import _lib, backend, workloads as W
from cudaedgeelimination_amd.graph import CEIL_2D, EUC_2D, GraphSnapshot
from cudaedgeelimination_amd.resident import JvSession, KoptBatch

T_OF = {3: 4, 4: 25, 5: 208}

pytestmark = pytest.mark.gpu

ROOT = pathlib.Path(__file__).resolve().parent.parent
GOLDEN = ROOT / "tests" / "golden"
G = json.loads((GOLDEN / "goldens.json").read_text())


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def config3_small():
    return W.config3(seed=0xEE0000C3, n=2000, rows=4096)


def single_instance_batch(ctx, wl):
    b = KoptBatch(ctx)
    inst = b.add_instance(wl.graph)
    segs = [(inst, k, int((wl.ks == k).sum())) for k in (3, 4, 5)]
    b.set_work(segs, np.concatenate([wl.bucket(k) for k in (3, 4, 5)]))
    return b


# rows per wavefront tile for k = 3, 4, 5; grid; workgroup size; stores
TUNINGS = [
    ((32, 16, 8), 0, 256, 0),
    ((1, 1, 1), 0, 256, 0),
    ((7, 3, 5), 0, 64, 1),
    ((48, 24, 12), 512, 128, 0),
    ((33, 17, 2), 37, 192, 1),
]


@pytest.mark.parametrize("rows,grid,block,nt", TUNINGS)
def test_batch_equals_reference_for_every_tuning(ctx, config3_small, rows, grid, block, nt):
    say(f"batch tuning rows={rows} grid={grid} block={block} nt={nt}")
    b = single_instance_batch(ctx, config3_small)
    b.set_tuning(rows, grid, block, nt)
    b.run(2)
    for i, k in enumerate((3, 4, 5)):
        assert sha(b.segment_costs(i)) == G["config3_small"][f"k{k}"]["sha256"], f"k={k}"
    b.close()


def test_batch_full_config3_graph_replay(ctx):
    say("batch config3 graph replay")
    wl = W.config3()
    b = single_instance_batch(ctx, wl)
    for i, k in enumerate((3, 4, 5)):
        b.run(1)
        assert sha(b.segment_costs(i)) == G["config3"][f"k{k}"]["sha256"]
    # graph replays interleaved with eager work and downloads
    for rep in range(3):
        b.run(2, use_graph=True)
        for i, k in enumerate((3, 4, 5)):
            assert sha(b.segment_costs(i)) == G["config3"][f"k{k}"]["sha256"], f"replay {rep}"
        b.run(1, use_graph=False)
    b.close()


def test_batch_pipelined_runs_equal_reference(ctx):
    """Launches dealt over 2 to 4 streams and cost pools: every pool a
    pipelined run wrote holds the whole result after 1, 2, depth and 5
    launches, also between ordinary launches; a pool no launch of the current
    work wrote is refused; mask output and graph replay refuse pipelining."""
    say("batch config3 pipelined runs")
    wl = W.config3()
    b = single_instance_batch(ctx, wl)
    for depth in (2, 3, 4):
        b.set_pipeline_depth(depth)
        for launches in sorted({1, 2, depth, 5}):
            assert b.run(launches, pipelined=True) > 0
            for pool in range(min(launches, depth)):
                for i, k in enumerate((3, 4, 5)):
                    assert sha(b.segment_costs(i, pool)) == G["config3"][f"k{k}"]["sha256"], \
                        f"depth {depth}, {launches} launches, pool {pool}, k={k}"
            b.run(1)
    # new work forgets which pools hold results
    b.set_work(b.segments, np.concatenate([wl.bucket(k) for k in (3, 4, 5)]))
    with pytest.raises(_lib.EeError):
        b.segment_costs(0, 1)
    b.run(4, pipelined=True)
    for pool in range(4):
        assert sha(b.segment_costs(2, pool)) == G["config3"]["k5"]["sha256"], pool
    for bad in (1, 5):
        with pytest.raises(_lib.InvalidArgument):
            b.set_pipeline_depth(bad)
    with pytest.raises(_lib.InvalidArgument):
        b.run(2, use_graph=True, pipelined=True)
    b.set_output("masks")
    with pytest.raises(_lib.InvalidArgument):
        b.run(2, pipelined=True)
    b.close()


def test_stamped_runs_time_every_launch(ctx):
    """Device-clock stamps of every wavefront: each launch gets a positive
    span, launches on one stream follow each other, and the stamped runs
    leave every pool's cells unchanged."""
    say("batch stamped runs")
    wl = W.config3()
    b = single_instance_batch(ctx, wl)
    b.set_tuning([32, 16, 8], 0, 256, 1)
    b.set_pipeline_depth(4)
    for pipelined in (False, True):
        start, end = b.run_stamped(24, pipelined=pipelined)
        assert (end > start).all() and start.min() == 0.0
        assert (end - start).max() < 10_000.0  # a launch of this batch takes microseconds
        if not pipelined:
            assert (start[1:] >= end[:-1] - 0.05).all()  # one clock tick of slack
        for pool in range(4 if pipelined else 1):
            assert sha(b.segment_costs(2, pool)) == G["config3"]["k5"]["sha256"], (pipelined, pool)
    with pytest.raises(_lib.InvalidArgument):
        b.run_stamped(0)
    b.close()


def test_batch_autotune_and_cache(ctx, config3_small, tmp_path):
    say("batch autotune")
    b = single_instance_batch(ctx, config3_small)
    cache = tmp_path / "tuning.txt"
    chosen = b.autotune(3, str(cache))
    assert cache.exists() and cache.read_text().strip()
    b.run(1)
    for i, k in enumerate((3, 4, 5)):
        assert sha(b.segment_costs(i)) == G["config3_small"][f"k{k}"]["sha256"]
    b2 = single_instance_batch(ctx, config3_small)
    assert b2.autotune(3, str(cache)) == chosen  # read back, no re-tuning
    # pipelined runs are tuned and cached under their own key
    b.set_pipeline_depth(3)
    piped = b.autotune(6, str(cache), pipelined=True)
    assert len(cache.read_text().strip().splitlines()) == 2
    b.run(3, pipelined=True)
    for i, k in enumerate((3, 4, 5)):
        assert sha(b.segment_costs(i)) == G["config3_small"][f"k{k}"]["sha256"]
    b2.set_pipeline_depth(3)
    assert b2.autotune(6, str(cache), pipelined=True) == piped
    b.close()
    b2.close()


def test_batch_many_instances_match_oracle(ctx):
    say("batch many instances")
    graphs, wls = [], []
    for i in range(8):
        g, wl = W.config4_instance(i, rows=256)
        graphs.append(g)
        wls.append(wl)
    b = KoptBatch(ctx)
    ids = [b.add_instance(g) for g in graphs]
    segs, tasks = [], []
    for i, wl in enumerate(wls):
        for k in ((5, 3, 4) if i % 2 else (3, 4, 5)):
            rows = wl.bucket(k)
            if i == 3 and k == 4:
                rows = rows[:0]  # an empty segment
            segs.append((ids[i], k, len(rows)))
            tasks.append(rows)
    offsets = b.set_work(segs, np.concatenate(tasks))
    b.autotune(2)
    b.run(1)
    out = b.download()
    assert len(out) == int(offsets[-1]) and b.cells == sum(len(t) * T_OF[k] for t, (_, k, _) in zip(tasks, segs))
    for s, (inst, k, count) in enumerate(segs):
        g = graphs[inst]
        want, _, _ = O.kopt_costs(g.x, g.y, g.distance_type, k, tasks[s])
        assert int(offsets[s]) % 16 == 0  # 128-byte aligned segments
        assert np.array_equal(out[int(offsets[s]):int(offsets[s]) + len(want)], want), (inst, k)
        assert np.array_equal(b.segment_costs(s), want)
    b.close()


def test_batch_matches_plain_path(ctx, config3_small):
    b = single_instance_batch(ctx, config3_small)
    b.run(1)
    for i, k in enumerate((3, 4, 5)):
        plain, _, _ = backend.evaluate_kopt_template_costs_cuda(
            config3_small.graph, backend.build_kopt_reconnect_table(k), config3_small.bucket(k), ctx)
        assert np.array_equal(b.segment_costs(i), plain)
    b.close()


def test_batch_rejects_bad_work(ctx):
    from cudaedgeelimination_amd import _lib

    wl = W.config3(seed=77, n=300, rows=30)
    b = KoptBatch(ctx)
    inst = b.add_instance(wl.graph)
    with pytest.raises(_lib.InvalidArgument):
        b.set_work([(inst + 1, 3, 0)], wl.tasks[:0])
    bad = wl.bucket(3).copy()
    bad["port_nodes"][0, 0] = wl.graph.dimension
    with pytest.raises(_lib.InvalidArgument):
        b.set_work([(inst, 3, len(bad))], bad)
    with pytest.raises(_lib.InvalidArgument):
        b.set_tuning((4096, 8, 8))
    b.close()


# ---------------------------------------------------------------- sessions
def assert_device_snapshot(session, og):
    """The session's device-resident active flags and CSR (compacted on the
    device after every commit) equal the oracle snapshot's RebuildCsr."""
    dev = session.device_snapshot()
    want = og.arrays()
    for key in ("edge_active", "row_offsets", "neighbors", "csr_edge_ids"):
        assert np.array_equal(dev[key], want[key]), key


@pytest.mark.parametrize("name", ["config1", "config2", "config1_ceil"])
def test_session_proof_equals_reference(ctx, name):
    say(f"session {name}")
    if name == "config1_ceil":
        base = W.config1(seed=0xEE0000C1, n=400)
        g = W.knn_graph(base.x, base.y, CEIL_2D, 12)
    else:
        g = W.config1() if name == "config1" else W.config2()
    s = JvSession(ctx, g)
    run = s.run(100)
    assert O.canonical_proof(run.proof) == (GOLDEN / f"{name}.ref.proof").read_text()
    og = O.OGraph.from_snapshot(g)
    O.jv_replay(og, run.proof)
    assert run.device_ms > 0 and run.edges_scanned > 0
    # the final snapshot equals the oracle's after the same run
    final = og.clone()
    O.oracle().orc_jv_replay(final.h, run.proof.encode())
    assert np.array_equal(s.active(), final.arrays()["edge_active"])
    # the last epoch committed nothing, so the device holds the final snapshot
    assert_device_snapshot(s, final)
    assert run.epochs >= 2 and run.compact_ms > 0
    # a second run starts from the final snapshot: one epoch, nothing to commit
    again = s.run(100)
    assert again.epochs == 1 and again.committed == 0
    assert_device_snapshot(s, final)
    s.close()


def test_session_rounds_limit_leaves_device_on_the_proposed_snapshot(ctx):
    """max_rounds cuts a run after an epoch that committed: the host applied
    that commit, the device still holds the snapshot the epoch proposed on
    (its next run uploads the host's); the next run continues from there and
    together the two runs commit what one uncut run commits."""
    g = W.config2()
    og = O.OGraph.from_snapshot(g)
    s = JvSession(ctx, g)
    first = s.run(1)
    assert first.epochs == 1 and first.committed > 0
    assert_device_snapshot(s, og)  # not compacted: the cut run proposes no more
    rest = s.run(100)
    full = O.jv_run(og)
    final = og.clone()
    O.oracle().orc_jv_replay(final.h, full.encode())
    assert np.array_equal(s.active(), final.arrays()["edge_active"])
    assert_device_snapshot(s, final)
    assert first.committed + rest.committed == int(
        sum(1 for line in full.splitlines() if line.startswith("record ")))
    s.close()


def test_session_reference_fixture_and_rounds_limit(ctx):
    rp = GOLDEN / "recursive-point"
    g = GraphSnapshot.load(rp / "recursive-point.tsp", rp / "recursive-point.edg")
    s = JvSession(ctx, g)
    run = s.run(100)
    assert O.canonical_proof(run.proof) == O.canonical_proof((rp / "recursive-point.ref.proof").read_text())
    assert run.epochs == 2
    s2 = JvSession(ctx, g)
    one = s2.run(1)
    assert one.epochs == 1
    og = O.OGraph.from_snapshot(g)
    assert O.canonical_proof(one.proof) == O.canonical_proof(O.jv_run(og, max_rounds=1))
    s.close()
    s2.close()


@pytest.mark.parametrize("seed", range(3))
def test_session_random_graphs(ctx, seed):
    rng = np.random.default_rng(seed)
    n = 600
    x = rng.integers(0, 2000, n).astype(np.int64)
    y = rng.integers(0, 2000, n).astype(np.int64)
    g = W.knn_graph(x, y, seed % 2, 8 + 4 * seed)
    s = JvSession(ctx, g)
    run = s.run(100)
    og = O.OGraph.from_snapshot(g)
    want = O.jv_run(og)
    assert O.canonical_proof(run.proof) == O.canonical_proof(want)
    final = og.clone()
    O.oracle().orc_jv_replay(final.h, want.encode())
    assert_device_snapshot(s, final)
    s.close()


def pack_masks(costs: np.ndarray, tasks: np.ndarray, T: int, words: int) -> np.ndarray:
    """Bit j of word j / 32 of each row: cell (row, j) < the row's deleted_cost."""
    improving = costs.reshape(len(tasks), T) < tasks["deleted_cost"][:, None]
    out = np.zeros((len(tasks), words), np.uint32)
    for j in range(T):
        out[:, j // 32] |= improving[:, j].astype(np.uint32) << np.uint32(j % 32)
    return out


@pytest.mark.parametrize("rows", [(32, 16, 8), (48, 24, 6), (64, 64, 64), (1, 1, 1)])
def test_batch_improving_masks_match_oracle(ctx, config3_small, rows):
    """Mask output: exactly the templates TryReconnectFromCostRow rechecks
    (cell < deleted_cost), per row, for every tile shape; switching back to
    cost output restores the cells."""
    say(f"batch improving masks rows={rows}")
    wl = config3_small
    g = wl.graph
    b = KoptBatch(ctx)
    inst = b.add_instance(g)
    segs = [(inst, k, len(wl.bucket(k))) for k in (3, 4, 5)]
    tasks = np.concatenate([wl.bucket(k) for k in (3, 4, 5)])
    b.set_work(segs, tasks)
    b.set_tuning(list(rows), 0, 64, 0)
    b.set_output("masks")
    assert b.mask_words() == 7
    b.run(1)
    got = b.download_masks()
    row = 0
    total_bits = 0
    for k in (3, 4, 5):
        t = wl.bucket(k)
        want_costs, _, _ = O.kopt_costs(g.x, g.y, g.distance_type, k, t)
        want = pack_masks(want_costs, t, T_OF[k], 7)
        assert np.array_equal(got[row:row + len(t)], want), k
        total_bits += int(np.unpackbits(want.view(np.uint8)).sum())
        row += len(t)
    assert total_bits > 0  # the workload has improving cells
    b.run(2, use_graph=True)
    assert np.array_equal(b.download_masks(), got)
    b.set_output("costs")
    b.run(1)
    for i, k in enumerate((3, 4, 5)):
        assert sha(b.segment_costs(i)) == G["config3_small"][f"k{k}"]["sha256"]
    b.close()


def test_batch_masks_single_k_and_many_instances(ctx):
    """One k only (one mask word per row) and many instances with mixed k."""
    say("batch masks single k / many instances")
    g, wl = W.config4_instance(5, rows=300)
    b = KoptBatch(ctx)
    inst = b.add_instance(g)
    t3 = wl.bucket(3)
    b.set_work([(inst, 3, len(t3))], t3)
    b.set_output("masks")
    assert b.mask_words() == 1
    b.run(1)
    want_costs, _, _ = O.kopt_costs(g.x, g.y, g.distance_type, 3, t3)
    assert np.array_equal(b.download_masks(), pack_masks(want_costs, t3, 4, 1))
    b.close()
    graphs = [W.config4_instance(i, rows=200) for i in range(4)]
    b = KoptBatch(ctx)
    ids = [b.add_instance(gg) for gg, _ in graphs]
    segs, tasks = [], []
    for i, (gg, w) in enumerate(graphs):
        for k in (5, 4, 3):
            segs.append((ids[i], k, len(w.bucket(k))))
            tasks.append(w.bucket(k))
    b.set_work(segs, np.concatenate(tasks))
    b.set_output("masks")
    b.run(1)
    got = b.download_masks()
    row = 0
    for (inst, k, count), t in zip(segs, tasks):
        gg = graphs[inst][0]
        want_costs, _, _ = O.kopt_costs(gg.x, gg.y, gg.distance_type, k, t)
        assert np.array_equal(got[row:row + count], pack_masks(want_costs, t, T_OF[k], 7)), (inst, k)
        row += count
    b.close()
