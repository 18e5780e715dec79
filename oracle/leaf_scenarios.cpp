// Test harness, not part of the product: the reference's leaf-search
// scenarios driven through the reference's own API, with its k-opt cost
// backend resolved to the HIP shim (cudaedgeelimination_amd/shim) at link
// time.  Built by oracle/Makefile (target leaf_scenarios_hip) from the
// reference sources where they lie; run by tests/test_gpu_dropin.py.
//
// Scenarios (restated, not copied):
//  1. The collinear 7-node tour of tests/unit/test_kopt.cpp:575-707: the
//     exhaustive 3/4/5-opt leaf search tests 25 anchored deletion sets and
//     1,330 proper templates and finds no improvement; the device-backed
//     proof's bytes equal the CPU proof's; two cursors over the same path
//     system fuse into 13 cost batches of 50 rows and 2,660 cells, every cell
//     certified, with one snapshot upload and three template and workspace
//     uploads on the device; a deletion budget of 3 ends inside the second
//     block (2 batches, 6 rows, 24 cells).
//  2. TestLeafCursorDifferential (test_kopt.cpp:842-897): 12 seeded random
//     7-node instances, three path systems each, budgets and batch sizes
//     varying by trial: the auto-backend batch, the CPU batch and the
//     per-system scalar searches give byte-identical proofs.  Run twice:
//     with the default device threshold (128 cells) and with every batch on
//     the device.
#include <cstdint>
#include <iostream>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "cuda_edge_elimination/local_search.hpp"
#include "cuda_edge_elimination/path_system.hpp"

namespace {

int g_failures = 0;

void Expect(bool ok, const std::string& what) {
  std::cout << (ok ? "ok   " : "FAIL ") << what << '\n';
  if (!ok) ++g_failures;
}

cudaee::GraphSnapshot PointsGraph(const std::vector<std::pair<int64_t, int64_t>>& xy) {
  cudaee::GraphSnapshot g;
  g.dimension = static_cast<int32_t>(xy.size());
  g.distance_type = cudaee::DistanceType::kEuc2D;
  g.integer_coordinates = true;
  g.integer_distance_safe = true;
  for (const auto& [x, y] : xy) {
    cudaee::Point p;
    p.x = static_cast<double>(x);
    p.y = static_cast<double>(y);
    p.integer_x = x;
    p.integer_y = y;
    g.points.push_back(p);
  }
  return g;
}

cudaee::KOptSearchOptions Options(uint32_t max_k, cudaee::PathCompatibilityBackend backend, uint32_t batch,
                                  uint64_t budget, uint64_t min_device_cells) {
  cudaee::KOptSearchOptions o;
  o.max_k = max_k;
  o.max_deletion_sets = budget;
  o.cost_backend = backend;
  o.cost_batch_size = batch;
  o.cuda_min_cost_cells = min_device_cells;
  return o;
}

std::string Bytes(const cudaee::PathSystemKOptProof& p) { return cudaee::SerializePathSystemKOptProof(p); }

void CollinearScenario() {
  std::vector<std::pair<int64_t, int64_t>> xy;
  for (int64_t i = 0; i < 7; ++i) xy.emplace_back(i, 0);
  const cudaee::GraphSnapshot g = PointsGraph(xy);
  const cudaee::NormalizedPathSystem paths = cudaee::NormalizePathSystem({{0, 1, 2, 3, 4, 5, 6}}, g.dimension);
  const cudaee::NodeEdge required{0, 1};
  using B = cudaee::PathCompatibilityBackend;

  const auto cpu = cudaee::ProvePathSystemByKOpt(g, paths, required, Options(5, B::kCpu, 4096, 0, 128));
  Expect(!cpu.proven && cpu.deletion_sets_tested == 25 && cpu.reconnect_matchings_tested == 1330,
         "collinear tour: 25 deletion sets and 1330 templates, no improvement (CPU)");
  const auto dev = cudaee::ProvePathSystemByKOpt(g, paths, required, Options(5, B::kCuda, 2, 0, 128));
  Expect(Bytes(dev) == Bytes(cpu), "collinear tour: device-backed proof bytes equal the CPU proof's");

  const auto cursor = Options(5, B::kAuto, 2, 0, 0);
  const auto scalar = cudaee::ProvePathSystemByKOpt(g, paths, required, cursor);
  cudaee::detail::ClearKOptCostCudaCache();
  const auto batch = cudaee::ProvePathSystemsByKOpt(g, {paths, paths}, required, cursor);
  Expect(batch.cpu_verified && batch.scalar_searches == 0 && batch.cursor_searches_started == 2 &&
             batch.cost_batches == 13 && batch.cost_tasks == 50 && batch.cost_cells == 2660 &&
             batch.cost_rows_consumed == 50 && batch.cpu_certified_cost_cells == batch.cost_cells,
         "two cursors: 13 batches, 50 rows, 2660 cells, all certified");
  Expect(batch.proofs.size() == 2 && Bytes(batch.proofs[0]) == Bytes(scalar) && Bytes(batch.proofs[1]) == Bytes(scalar),
         "two cursors: proof bytes equal the scalar search's");
  Expect(batch.cuda_cost_batches == batch.cost_batches && batch.snapshot_cache_hits + 1 == batch.cuda_cost_batches &&
             batch.template_cache_hits + 3 == batch.cuda_cost_batches &&
             batch.workspace_cache_hits + 3 == batch.cuda_cost_batches && batch.peak_device_cache_bytes > 0 &&
             batch.cpu_completeness_rows == 0 && batch.cpu_completeness_templates == 0 &&
             batch.cost_cpu_certify_ms > 0.0,
         "two cursors: every batch on the device, one snapshot and three template/workspace uploads");

  const auto budget = Options(5, B::kAuto, 2, 3, 0);
  const auto scalar_budget = cudaee::ProvePathSystemByKOpt(g, paths, required, budget);
  const auto batch_budget = cudaee::ProvePathSystemsByKOpt(g, {paths, paths}, required, budget);
  Expect(batch_budget.cost_batches == 2 && batch_budget.cost_tasks == 6 && batch_budget.cost_cells == 24 &&
             Bytes(batch_budget.proofs[0]) == Bytes(scalar_budget) &&
             Bytes(batch_budget.proofs[1]) == Bytes(scalar_budget),
         "deletion budget 3: 2 batches, 6 rows, 24 cells, proofs equal the scalar search's");
}

void DifferentialScenario(uint64_t min_device_cells) {
  using B = cudaee::PathCompatibilityBackend;
  std::mt19937 rng(24092026U);
  std::uniform_int_distribution<int64_t> coordinate(-40, 40);
  uint64_t device_batches = 0, mismatches = 0;
  for (uint32_t trial = 0; trial < 12; ++trial) {
    std::vector<std::pair<int64_t, int64_t>> xy;
    for (int node = 0; node < 7; ++node) {
      const int64_t x = coordinate(rng);
      const int64_t y = coordinate(rng);
      xy.emplace_back(x, y);
    }
    const cudaee::GraphSnapshot g = PointsGraph(xy);
    const std::vector<cudaee::NormalizedPathSystem> systems = {
        cudaee::NormalizePathSystem({{0, 1, 2, 3, 4, 5, 6}}, g.dimension),
        cudaee::NormalizePathSystem({{0, 1, 3, 5, 2, 6, 4}}, g.dimension),
        cudaee::NormalizePathSystem({{0, 1, 4}, {2, 5, 3, 6}}, g.dimension)};
    const uint64_t budget = trial % 4 == 0 ? 0 : 1 + trial % 7;
    const auto options = Options(3 + trial % 3, B::kAuto, 1 + trial % 4, budget, min_device_cells);
    auto cpu_options = options;
    cpu_options.cost_backend = B::kCpu;
    const cudaee::NodeEdge required{0, 1};
    const auto batch = cudaee::ProvePathSystemsByKOpt(g, systems, required, options);
    const auto cpu_batch = cudaee::ProvePathSystemsByKOpt(g, systems, required, cpu_options);
    device_batches += batch.cuda_cost_batches;
    if (!batch.cpu_verified || batch.proofs.size() != systems.size() || !cpu_batch.cpu_verified ||
        cpu_batch.cpu_certified_cost_cells != cpu_batch.cost_cells) {
      ++mismatches;
      continue;
    }
    for (size_t i = 0; i < systems.size(); ++i) {
      const std::string scalar = Bytes(cudaee::ProvePathSystemByKOpt(g, systems[i], required, cpu_options));
      const std::string direct = Bytes(cudaee::ProvePathSystemByKOpt(g, systems[i], required, options));
      if (direct != scalar || Bytes(batch.proofs[i]) != scalar || Bytes(cpu_batch.proofs[i]) != scalar) ++mismatches;
    }
  }
  Expect(mismatches == 0, "12 seeded trials x 3 path systems, device threshold " + std::to_string(min_device_cells) +
                              " cells: auto, CPU and scalar proofs byte-identical");
  // with the default threshold these 7-node leaves may stay on the CPU; the
  // threshold-0 run puts every batch on the device
  if (min_device_cells == 0) {
    Expect(device_batches > 0, "the auto backend ran " + std::to_string(device_batches) + " batches on the device");
  } else {
    std::cout << "info the auto backend ran " << device_batches << " batches on the device\n";
  }
}

}  // namespace

int main() {
  std::string reason;
  if (!cudaee::detail::KOptCostCudaAvailable(&reason)) {
    std::cout << "no device k-opt backend: " << reason << '\n';
    return 2;
  }
  try {
    CollinearScenario();
    DifferentialScenario(128);
    DifferentialScenario(0);
  } catch (const std::exception& e) {
    std::cout << "FAIL exception: " << e.what() << '\n';
    return 1;
  }
  std::cout << (g_failures == 0 ? "all leaf scenarios passed" : "leaf scenarios failed") << '\n';
  return g_failures == 0 ? 0 : 1;
}
