"""Writes the golden fixtures under tests/golden/ from the reference's own CPU
path (oracle/_ref/libcudaee_ref.so, built from the reference sources by
oracle/Makefile).  Run in the container that holds the reference tree:

    make -C oracle ref && python tests/golden/make_goldens.py

Fixtures are data only: inputs (or the seed and digest that regenerate them)
and the reference's outputs.
"""
from __future__ import annotations

import hashlib
import json
import pathlib
import shutil
import sys
import tempfile

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import _oracle as O  # noqa: E402
from cudaedgeelimination_amd import workloads as W  # noqa: E402
from cudaedgeelimination_amd.graph import CEIL_2D, EUC_2D, GraphSnapshot  # noqa: E402

REFERENCE_DATA = pathlib.Path("/root/reference/tests/data")


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def ref_run(tsp, edg, rounds=100) -> str:
    import ctypes as C

    buf = C.create_string_buffer(1 << 26)
    size = O.ref().ref_jv_run_files(str(tsp).encode(), str(edg).encode(), rounds, buf, len(buf))
    assert 0 <= size < len(buf), O.ref().ref_last_error()
    return buf.value.decode()


def ref_witnesses(tsp, edg, m):
    w = np.empty(m, np.int32)
    assert O.ref().ref_jv_candidates_files(str(tsp).encode(), str(edg).encode(), O.p(w), m) >= 0
    return w


def seventeen_node_graph():
    # tests/unit/test_kopt.cpp:158-166
    x = np.array([(7 * i) % 23 for i in range(17)], np.int64)
    y = np.array([(i * i + 3 * i) % 19 for i in range(17)], np.int64)
    return x, y


def seventeen_node_tasks(x, y, k, dtype=EUC_2D):
    # tests/unit/test_kopt.cpp:177-191
    tasks = np.zeros(3, O.TASK_DTYPE)
    for port in range(2 * k):
        tasks[0]["port_nodes"][port] = port
        tasks[2]["port_nodes"][port] = (3 * port + 2) % 17
    tasks[1]["port_nodes"] = [0, 1, 1, 2, 3, 4, 5, 6, 7, 8]
    from cudaedgeelimination_amd.graph import exact_distances

    for t in tasks:
        pn = t["port_nodes"]
        t["deleted_cost"] = int(sum(exact_distances(x[pn[2 * e]], y[pn[2 * e]], x[pn[2 * e + 1]],
                                                    y[pn[2 * e + 1]], dtype) for e in range(k)))
    return tasks


def main() -> None:
    assert O.ref_available(), "build the reference first: make -C oracle ref"
    out: dict = {"generator": "tests/golden/make_goldens.py", "source": "reference CPU path"}

    # 1. the reference's own fixture and the JV proof the reference writes for it
    fixture = HERE / "recursive-point"
    fixture.mkdir(exist_ok=True)
    for name in ("recursive-point.tsp", "recursive-point.edg", "recursive-point.tour"):
        shutil.copyfile(REFERENCE_DATA / name, fixture / name)
    proof = ref_run(fixture / "recursive-point.tsp", fixture / "recursive-point.edg")
    (fixture / "recursive-point.ref.proof").write_text(proof)
    out["recursive_point_witnesses"] = ref_witnesses(
        fixture / "recursive-point.tsp", fixture / "recursive-point.edg", 28).tolist()

    # 2. templates
    out["templates"] = {}
    for k in (3, 4, 5):
        mates, h = O.templates(k, "ref")
        out["templates"][str(k)] = {"hash": f"{h:016x}",
                                    "mates": [bytes(row[:2 * k]).hex() for row in mates]}

    # 3. seeded JV instances: configs 1 and 2 (full size) as the reference runs them
    with tempfile.TemporaryDirectory() as tmp:
        tmp = pathlib.Path(tmp)
        for name, graph in (("config1", W.config1()), ("config2", W.config2()),
                            ("config1_ceil", _ceil_copy(W.config1(seed=0xEE0000C1, n=400)))):
            tsp, edg = tmp / f"{name}.tsp", tmp / f"{name}.edg"
            graph.write(tsp, edg, name)
            text = ref_run(tsp, edg)
            (HERE / f"{name}.ref.proof").write_text(O.canonical_proof(text))
            w = ref_witnesses(tsp, edg, graph.edge_count)
            out[f"{name}_first_epoch"] = {"n": graph.dimension, "m": graph.edge_count,
                                          "content_hash": f"{graph.content_hash():016x}",
                                          "witness_sha256": sha(w),
                                          "candidates": int((w >= 0).sum())}

    # 4. k-opt matrices: the reference test graph and tasks, then seeded workloads
    x, y = seventeen_node_graph()
    out["kopt_17"] = {}
    for k in (3, 4, 5):
        tasks = seventeen_node_tasks(x, y, k)
        costs, _, _ = O.ref_kopt_costs(x, y, EUC_2D, k, tasks)
        out["kopt_17"][str(k)] = costs.tolist()
    rep = np.zeros(40, O.TASK_DTYPE)
    rep["port_nodes"] = np.arange(10)
    from cudaedgeelimination_amd.graph import exact_distances
    rep["deleted_cost"] = int(sum(exact_distances(x[2 * e], y[2 * e], x[2 * e + 1], y[2 * e + 1],
                                                  EUC_2D) for e in range(5)))
    costs, threads, nodes = O.ref_kopt_costs(x, y, EUC_2D, 5, rep)
    out["kopt_17_repeated"] = {"row": costs[:208].tolist(), "cache_nodes": nodes}
    ceil_task = np.zeros(1, O.TASK_DTYPE)
    ceil_task["port_nodes"] = np.arange(10)
    costs, _, _ = O.ref_kopt_costs(x, y, CEIL_2D, 5, ceil_task)
    out["kopt_17_ceil"] = costs.tolist()

    for name, wl in (("config3_small", W.config3(seed=0xEE0000C3, n=2000, rows=4096)),
                     ("config3", W.config3())):
        entry = {"n": wl.graph.dimension, "rows": int(len(wl.tasks)),
                 "content_hash": f"{wl.graph.content_hash():016x}",
                 "tasks_sha256": sha(wl.tasks), "ks_sha256": sha(wl.ks)}
        for k in (3, 4, 5):
            tasks = wl.bucket(k)
            costs, threads, nodes = O.ref_kopt_costs(wl.graph.x, wl.graph.y, EUC_2D, k, tasks)
            entry[f"k{k}"] = {"rows": int(len(tasks)), "sha256": sha(costs),
                              "first_row": costs[:{3: 4, 4: 25, 5: 208}[k]].tolist(),
                              "invalid_cells": int((costs == np.iinfo(np.int64).max).sum())}
        out[name] = entry

    (HERE / "goldens.json").write_text(json.dumps(out, indent=1, sort_keys=True) + "\n")
    print("wrote", HERE / "goldens.json")


def _ceil_copy(g: GraphSnapshot) -> GraphSnapshot:
    return W.knn_graph(g.x, g.y, CEIL_2D, 12)


if __name__ == "__main__":
    main()
