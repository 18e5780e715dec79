"""Drop-in check: the reference's own CLI and unit test, built from the
reference sources with cudaedgeelimination_amd/shim in place of its two CUDA
translation units (oracle/Makefile, target dropin), run on the device.

* test_core's GPU test (tests/unit/test_core.cpp:167-229: candidates equal the
  CPU path, cache-hit pattern, resident bytes 24m + 20n + 4 + 8A);
* `gpu-eliminate --backend cuda` proofs equal `--backend cpu` proofs and
  replay with `verify` (the reference's CLI path, src/cli/main.cpp:249-275);
* the reference's leaf-search scenarios (tests/unit/test_kopt.cpp:575-707,
  842-897), restated in oracle/leaf_scenarios.cpp over the reference API:
  the collinear 7-node exhaustion (25 sets, 1,330 templates), device-backed
  proof bytes equal to CPU, the fused cursor batch counters and cache-hit
  pattern, the budget boundary, and 12 seeded differential trials;
* the reference's hybrid leaf scenario `ht-prove --leaf-backend cuda`
  (tests/CMakeLists.txt:146-184): every device k-opt matrix is certified cell
  by cell against the reference CPU matrix inside the reference code
  (src/cpu/kopt_search.cpp:1455-1475), then `ht-verify` replays the proof.
"""
from __future__ import annotations

import pathlib
import subprocess

import pytest

import _oracle as O
from cudaedgeelimination_amd import workloads as W

pytestmark = pytest.mark.gpu

ROOT = pathlib.Path(__file__).resolve().parent.parent
CLI = ROOT / "oracle" / "_ref" / "cudaee_hip"
TEST_CORE = ROOT / "oracle" / "_ref" / "test_core_hip"
TEST_HT = ROOT / "oracle" / "_ref" / "test_ht_hip"
LEAF = ROOT / "oracle" / "_ref" / "leaf_scenarios_hip"
RP = ROOT / "tests" / "golden" / "recursive-point"


def need(path):
    if not path.exists():
        pytest.skip(f"{path.name} not built (make -C oracle dropin, needs the reference tree)")


def run(args, cwd):
    print("run:", " ".join(map(str, args)), flush=True)
    done = subprocess.run([str(a) for a in args], cwd=cwd, capture_output=True, text=True,
                          timeout=120)
    print(done.stdout[-2000:], done.stderr[-2000:], flush=True)
    return done


def test_reference_unit_test_core_passes_on_device(tmp_path):
    need(TEST_CORE)
    done = run([TEST_CORE], tmp_path)
    assert done.returncode == 0 and "unit tests passed" in done.stdout


def test_reference_leaf_search_scenarios_on_device(tmp_path):
    need(LEAF)
    done = run([LEAF], tmp_path)
    assert done.returncode == 0 and "all leaf scenarios passed" in done.stdout
    assert "FAIL" not in done.stdout


def test_reference_unit_test_hamilton_tutte_passes_on_device(tmp_path):
    # the reference's HT unit test with its device sections on: its Hamilton
    # and end reply checks (tests/unit/test_hamilton_tutte.cpp:233-265,
    # 345-380) compare the shim's EvaluateHt*RepliesCuda with the CPU batch,
    # including the cache-hit pattern; its other device paths are absent and
    # report themselves skipped
    need(TEST_HT)
    done = run([TEST_HT], tmp_path)
    assert done.returncode == 0 and "Hamilton-Tutte tests passed" in done.stdout


@pytest.mark.parametrize("name", ["config1", "config2", "recursive"])
def test_reference_cli_cuda_proof_equals_cpu_proof(name):
    need(CLI)
    work = pathlib.Path("/tmp/ee_dropin_cli")
    work.mkdir(exist_ok=True)
    if name == "recursive":
        tsp, edg = RP / "recursive-point.tsp", RP / "recursive-point.edg"
    else:
        graph = W.config1() if name == "config1" else W.config2()
        tsp, edg = work / f"{name}.tsp", work / f"{name}.edg"
        graph.write(tsp, edg, name)
    proofs = {}
    for backend in ("cpu", "cuda"):
        out, proof = work / f"{name}.{backend}.edg", work / f"{name}.{backend}.proof"
        done = run([CLI, "gpu-eliminate", "--tsp", tsp, "--edges", edg, "--output", out,
                    "--proof", proof, "--backend", backend, "--max-rounds", "100"], work)
        assert done.returncode == 0 and f"status=OK backend={backend}" in done.stdout
        proofs[backend] = proof.read_text()
    assert O.canonical_proof(proofs["cuda"]) == O.canonical_proof(proofs["cpu"])
    done = run([CLI, "verify", "--tsp", tsp, "--edges", edg, "--proof",
                work / f"{name}.cuda.proof"], work)
    assert done.returncode == 0


def test_reference_hybrid_leaf_scenario_on_device():
    need(CLI)
    work = pathlib.Path("/tmp/ee_dropin_cli")
    work.mkdir(exist_ok=True)
    proof = work / "recursive-point.wavefront.hybrid.ht-proof"
    done = run([CLI, "ht-prove", "--tsp", RP / "recursive-point.tsp",
                "--edges", RP / "recursive-point.edg", "--u", "2", "--v", "4",
                "--proof", proof, "--scheduler", "wavefront", "--backend", "cpu",
                "--leaf-backend", "cuda", "--fuse-leaf-buckets", "0", "--reply-backend", "cpu",
                "--path-append-backend", "cpu", "--propagation-backend", "cpu",
                "--max-neighborhood", "0", "--max-cd-candidates", "8",
                "--max-candidate-degree", "0", "--max-root-replies", "1000", "--max-k", "3",
                "--max-deletion-sets", "1", "--max-depth", "1", "--max-states", "20000",
                "--max-total-replies", "200000", "--max-replies-per-move", "1000",
                "--max-point-candidates", "5", "--max-end-candidates", "5"], work)
    assert done.returncode == 0
    assert "fuse_leaf_buckets=0 leaf_cost_backend=cuda" in done.stdout
    done = run([CLI, "ht-verify", "--tsp", RP / "recursive-point.tsp",
                "--edges", RP / "recursive-point.edg", "--proof", proof], work)
    assert done.returncode == 0
