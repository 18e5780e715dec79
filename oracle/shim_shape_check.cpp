// Test harness (never shipped): the drop-in shim's HT reply entry points
// must reject a malformed snapshot with std::invalid_argument before they
// touch a device, as the reference's ValidateGraphShape does
// (src/cuda/ht_replies.cu:521-531, called ahead of SelectDevice at :593 and
// :675).  Runs without a GPU.  Usage: shim_shape_check <tsp> <edg>
#include "cuda_edge_elimination/graph.hpp"
#include "cuda_edge_elimination/hamilton_tutte.hpp"

#include <cstdio>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

int failures = 0;

void Expect(const char* what, const std::function<void(cudaee::GraphSnapshot&)>& spoil,
            const cudaee::GraphSnapshot& base) {
  for (int kind = 0; kind < 2; ++kind) {
    cudaee::GraphSnapshot g = base;
    spoil(g);
    try {
      if (kind == 0) {
        const cudaee::NodeEdge target{base.edges[0].u, base.edges[0].v};
        static_cast<void>(cudaee::detail::EvaluateHtHamiltonRepliesCuda(g, target, {2, 3}, nullptr, nullptr));
      } else {
        const cudaee::HtEndReplyTask task{base.edges[0].u, base.edges[0].v};
        static_cast<void>(cudaee::detail::EvaluateHtEndRepliesCuda(g, {task}, nullptr, nullptr));
      }
      std::printf("FAIL %s (%s): no exception\n", what, kind == 0 ? "hamilton" : "end");
      ++failures;
    } catch (const std::invalid_argument&) {
      std::printf("ok   %s (%s)\n", what, kind == 0 ? "hamilton" : "end");
    } catch (const std::exception& e) {
      std::printf("FAIL %s (%s): %s\n", what, kind == 0 ? "hamilton" : "end", e.what());
      ++failures;
    }
  }
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s <tsp> <edg>\n", argv[0]);
    return 2;
  }
  const cudaee::GraphSnapshot base = cudaee::GraphSnapshot::Load(argv[1], argv[2]);
  Expect("truncated row_offsets", [](cudaee::GraphSnapshot& g) { g.row_offsets.pop_back(); }, base);
  Expect("row_offsets beyond neighbours", [](cudaee::GraphSnapshot& g) { g.neighbors.pop_back(); }, base);
  Expect("missing points", [](cudaee::GraphSnapshot& g) { g.points.pop_back(); }, base);
  Expect("row_offsets[0] != 0", [](cudaee::GraphSnapshot& g) { g.row_offsets[0] = 1; }, base);
  Expect("not distance safe", [](cudaee::GraphSnapshot& g) { g.integer_distance_safe = false; }, base);
  Expect("non-integer coordinates", [](cudaee::GraphSnapshot& g) { g.integer_coordinates = false; }, base);
  return failures == 0 ? 0 : 1;
}
