"""Many JV graphs in lockstep (the many-instance batching configuration):
every graph's proof and final snapshot equal its standalone session's and the
CPU oracle's, with and without HIP graph replay of the epochs; and the
many-instance k-opt batch over the same graphs equals the oracle."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import _oracle as O
from cudaedgeelimination_amd import workloads as W
from cudaedgeelimination_amd.graph import CEIL_2D, EUC_2D
from cudaedgeelimination_amd.resident import JvFleet, JvSession, KoptBatch

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def instances():
    """Twelve small instances of the many-instance generator (n log-uniform,
    uniform or clustered coordinates, CEIL_2D for odd indices), 10-NN."""
    return [W.config4_instance(i, n_range=(150, 700), rows=512) for i in range(12)]


@pytest.mark.parametrize("use_graph", [False, True])
def test_fleet_proofs_equal_standalone_and_oracle(ctx, instances, use_graph):
    say(f"fleet use_graph={use_graph}")
    graphs = [g for g, _ in instances]
    assert {g.distance_type for g in graphs} == {EUC_2D, CEIL_2D}
    fleet = JvFleet(ctx, graphs)
    st = fleet.run(100, use_graph=use_graph)
    assert st["instances"] == len(graphs)
    assert st["committed"] > 0
    if use_graph:
        # the first epoch uploads the union eagerly; every later one (device
        # commit, launch, download) is a replayed graph
        assert st["graph_launches"] == st["epochs"] - 1 and st["epochs"] >= 2
    for i, g in enumerate(graphs):
        og = O.OGraph.from_snapshot(g)
        want = O.jv_run(og)
        got = fleet.proof(i)
        assert O.canonical_proof(got) == O.canonical_proof(want), i
        final = og.clone()
        O.oracle().orc_jv_replay(final.h, got.encode())
        assert np.array_equal(fleet.active(i), final.arrays()["edge_active"]), i
    # one standalone session gives the same bytes (minus timings)
    s = JvSession(ctx, graphs[3])
    run = s.run(100)
    assert O.canonical_proof(run.proof) == O.canonical_proof(fleet.proof(3))
    assert fleet.instance_stats(3)["epochs"] == run.epochs
    s.close()
    fleet.close()


def test_fleet_rounds_limit_and_rerun(ctx, instances):
    graphs = [g for g, _ in instances[:4]]
    fleet = JvFleet(ctx, graphs)
    st = fleet.run(1)
    assert st["epochs"] == 1
    for i, g in enumerate(graphs):
        og = O.OGraph.from_snapshot(g)
        assert O.canonical_proof(fleet.proof(i)) == O.canonical_proof(O.jv_run(og, max_rounds=1))
    fleet.close()


def test_many_instance_kopt_batch_equals_oracle(ctx, instances):
    say("many-instance k-opt batch")
    b = KoptBatch(ctx)
    segs, tasks, expect = [], [], []
    for g, wl in instances:
        inst = b.add_instance(g)
        for k in (3, 4, 5):
            rows = wl.bucket(k)
            segs.append((inst, k, len(rows)))
            tasks.append(rows)
            expect.append(O.kopt_costs(g.x, g.y, g.distance_type, k, rows)[0])
    b.set_work(segs, np.concatenate(tasks))
    b.run(1)
    for i, want in enumerate(expect):
        assert np.array_equal(b.segment_costs(i), want), i
    b.run(3, use_graph=True)
    for i, want in enumerate(expect):
        assert np.array_equal(b.segment_costs(i), want), i
    b.close()


@pytest.fixture(scope="module")
def full_instances():
    """The many-instance workload at its full size: 256 instances, n
    log-uniform in [200, 4000], 4,096 rows each (2,048 / 1,229 / 819)."""
    return [W.config4_instance(i) for i in range(256)]


def test_full_scale_fleet_and_batch_equal_oracle(ctx, full_instances):
    """Every one of the 256 proofs and all 768 cost segments of the
    many-instance workload equal the CPU oracle (JV epochs replayed as HIP
    graphs, as the bench runs them)."""
    from concurrent.futures import ThreadPoolExecutor

    graphs = [g for g, _ in full_instances]
    assert len(graphs) == 256
    assert min(g.dimension for g in graphs) >= 200 and max(g.dimension for g in graphs) <= 4000
    say("full-scale fleet: 256 graphs")
    fleet = JvFleet(ctx, graphs)
    st = fleet.run(100, use_graph=True)
    assert st["instances"] == 256 and st["committed"] > 0
    got = [fleet.proof(i) for i in range(256)]
    with ThreadPoolExecutor(8) as pool:
        want = list(pool.map(lambda g: O.jv_run(O.OGraph.from_snapshot(g)), graphs))
    bad = [i for i in range(256) if O.canonical_proof(got[i]) != O.canonical_proof(want[i])]
    assert not bad, f"proofs differ for instances {bad[:10]}"
    for i in (0, 1, 2, 3, 127, 255):  # final snapshots, by replaying the device proof
        final = O.OGraph.from_snapshot(graphs[i])
        O.oracle().orc_jv_replay(final.h, got[i].encode())
        assert np.array_equal(fleet.active(i), final.arrays()["edge_active"]), i
    fleet.close()

    say("full-scale k-opt batch: 768 segments, 1,048,576 rows")
    b = KoptBatch(ctx)
    segs, tasks, expect = [], [], []
    for g, wl in full_instances:
        inst = b.add_instance(g)
        for k in (3, 4, 5):
            rows = wl.bucket(k)
            segs.append((inst, k, len(rows)))
            tasks.append(rows)
            expect.append((g, k, rows))
    b.set_work(segs, np.concatenate(tasks))
    assert sum(c for _, _, c in segs) == 256 * 4096
    assert b.cells == 53_572_864
    b.run(1, use_graph=True)
    bad = []
    for i, (g, k, rows) in enumerate(expect):
        if not np.array_equal(b.segment_costs(i), O.kopt_costs(g.x, g.y, g.distance_type, k, rows)[0]):
            bad.append(i)
    assert not bad, f"segments differ: {bad[:10]}"
    b.close()
