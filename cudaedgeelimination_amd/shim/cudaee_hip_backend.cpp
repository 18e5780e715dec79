// Drop-in C++ backend for cudaEdgeElimination over the gfx950 C ABI.
//
// This translation unit takes the place of the reference's CUDA units for
// the candidate-generation paths (src/cuda/jv_cuda.cu, src/cuda/kopt_cost.cu
// and src/cuda/ht_replies.cu): it defines the same C++ symbols, declared in the
// reference headers, so the reference's callers (RunJvElimination,
// EvaluateKOptTemplateCosts, the HT leaf batches, the CLI, the unit tests)
// link unchanged against libee_hip.so:
//
//   cudaee::CudaBackendAvailable / FindJvCandidatesCuda / ClearJvCudaCache
//     (include/cuda_edge_elimination/elimination.hpp:333-338)
//   cudaee::detail::KOptCostCudaAvailable / EvaluateKOptTemplateCostsCuda /
//     ClearKOptCostCudaCache (include/cuda_edge_elimination/local_search.hpp:235-242)
//   cudaee::detail::HtHamiltonReplyCudaAvailable / EvaluateHtHamiltonRepliesCuda /
//     HtEndReplyCudaAvailable / EvaluateHtEndRepliesCuda / ClearHtReplyCudaCache
//     (include/cuda_edge_elimination/hamilton_tutte.hpp:533-545, src/cuda/ht_replies.cu)
//
// It is compiled against the reference headers, so it builds only where the
// reference tree is present.  Device caches live in one ee_ctx per host
// thread (the reference keeps thread_local per-device caches,
// src/cuda/jv_cuda.cu:138, src/cuda/kopt_cost.cu:133); the device is the
// thread's forced ordinal when set (CudaDevicePreferenceForCurrentThread),
// otherwise the device with the most free memory, kept for later calls.
// Status codes become the exception types the reference throws.
#include "cuda_edge_elimination/cuda_device_affinity.hpp"
#include "cuda_edge_elimination/elimination.hpp"
#include "cuda_edge_elimination/hamilton_tutte.hpp"
#include "cuda_edge_elimination/local_search.hpp"

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "ee_hip.h"

static_assert(sizeof(cudaee::KOptCostTask) == sizeof(ee_kopt_task), "task layout");
static_assert(offsetof(cudaee::KOptCostTask, deleted_cost) == offsetof(ee_kopt_task, deleted_cost),
              "task layout");
static_assert(sizeof(cudaee::EndpointMatching) == sizeof(ee_endpoint_matching), "template layout");

namespace cudaee {
namespace {

void Throw(ee_status status, const std::string& what) {
  const std::string message = what + ": " + ee_last_error();
  switch (status) {
    case EE_OK: return;
    case EE_ERR_INVALID_ARGUMENT: throw std::invalid_argument(message);
    case EE_ERR_OVERFLOW: throw std::overflow_error(message);
    case EE_ERR_LOGIC: throw std::logic_error(message);
    default: throw std::runtime_error(message);
  }
}

struct ThreadContext {
  ee_ctx* ctx = nullptr;
  int device = -1;
  ~ThreadContext() { Reset(); }
  void Reset() {
    if (ctx != nullptr) ee_ctx_destroy(ctx);
    ctx = nullptr;
    device = -1;
  }
};

thread_local ThreadContext t_jv;
thread_local ThreadContext t_kopt;

// SelectDevice (src/cuda/jv_cuda.cu:290-347): forced > sticky > most free.
ee_ctx* Acquire(ThreadContext* slot, std::string* reason) {
  const int forced = detail::CudaDevicePreferenceForCurrentThread();
  if (slot->ctx != nullptr && (forced < 0 || forced == slot->device)) return slot->ctx;
  slot->Reset();
  ee_ctx* ctx = nullptr;
  const ee_status status = ee_ctx_create(forced >= 0 ? forced : -1, &ctx);
  if (status != EE_OK) {
    if (reason != nullptr) *reason = ee_last_error();
    return nullptr;
  }
  slot->ctx = ctx;
  slot->device = ee_ctx_device(ctx);
  return ctx;
}

struct Coordinates {
  std::vector<std::int64_t> x, y;
  explicit Coordinates(const GraphSnapshot& graph) {
    x.reserve(graph.points.size());
    y.reserve(graph.points.size());
    for (const Point& p : graph.points) {
      x.push_back(p.integer_x);
      y.push_back(p.integer_y);
    }
  }
};

}  // namespace

bool CudaBackendAvailable(std::string* const reason) {
  return Acquire(&t_jv, reason) != nullptr;
}

std::vector<Candidate> FindJvCandidatesCuda(const GraphSnapshot& graph, int* const selected_device,
                                            JvCudaCacheUsage* const cache_usage) {
  if (cache_usage != nullptr) *cache_usage = {};
  if (!graph.integer_coordinates || !graph.integer_distance_safe) {
    throw std::runtime_error("HIP JV supports distance-safe integer coordinates only");
  }
  if (graph.edges.size() > static_cast<std::size_t>(std::numeric_limits<std::int32_t>::max())) {
    throw std::overflow_error("HIP JV edge count exceeds int32 kernel indexing");
  }
  std::string reason;
  ee_ctx* ctx = Acquire(&t_jv, &reason);
  if (ctx == nullptr) throw std::runtime_error("HIP device unavailable: " + reason);
  if (selected_device != nullptr) *selected_device = t_jv.device;
  if (graph.edges.empty()) return {};

  const Coordinates xy(graph);
  const std::size_t m = graph.edges.size();
  std::vector<std::int32_t> u(m), v(m), witnesses(m, -1);
  std::vector<std::int64_t> w(m);
  std::vector<std::uint8_t> active(m);
  for (std::size_t e = 0; e < m; ++e) {
    u[e] = graph.edges[e].u;
    v[e] = graph.edges[e].v;
    w[e] = graph.edges[e].weight;
    active[e] = graph.edges[e].active ? 1 : 0;
  }
  ee_graph_view view{};
  view.dimension = graph.dimension;
  view.distance_type = static_cast<std::uint8_t>(graph.distance_type);
  view.integer_coordinates = 1;
  view.x = xy.x.data();
  view.y = xy.y.data();
  view.edge_count = static_cast<std::int32_t>(m);
  view.edge_u = u.data();
  view.edge_v = v.data();
  view.edge_weight = w.data();
  view.edge_active = active.data();
  view.row_offsets = graph.row_offsets.data();
  view.neighbors = graph.neighbors.data();
  view.csr_edge_ids = graph.csr_edge_ids.data();
  ee_jv_usage usage{};
  Throw(ee_jv_find_candidates(ctx, &view, witnesses.data(), &usage), "HIP JV propose");
  if (cache_usage != nullptr) {
    cache_usage->static_hit = usage.static_hit != 0;
    cache_usage->workspace_hit = usage.workspace_hit != 0;
    cache_usage->resident_bytes = usage.resident_bytes;
    cache_usage->h2d_ms = usage.h2d_ms;
    cache_usage->kernel_ms = usage.kernel_ms;
    cache_usage->d2h_ms = usage.d2h_ms;
  }
  std::vector<Candidate> candidates;
  for (std::size_t e = 0; e < m; ++e) {
    if (witnesses[e] >= 0) {
      candidates.push_back({static_cast<std::int32_t>(e), witnesses[e], EliminationMethod::kJv});
    }
  }
  return candidates;
}

void ClearJvCudaCache() { t_jv.Reset(); }

namespace detail {

bool KOptCostCudaAvailable(std::string* const reason) {
  return Acquire(&t_kopt, reason) != nullptr;
}

std::vector<std::int64_t> EvaluateKOptTemplateCostsCuda(const GraphSnapshot& graph,
                                                        const KOptReconnectTable& table,
                                                        const std::vector<KOptCostTask>& tasks,
                                                        int* const selected_device,
                                                        KOptCudaCacheUsage* const cache_usage) {
  if (cache_usage != nullptr) *cache_usage = {};
  std::string reason;
  ee_ctx* ctx = Acquire(&t_kopt, &reason);
  if (ctx == nullptr) throw std::runtime_error("HIP k-opt cost backend unavailable: " + reason);
  if (selected_device != nullptr) *selected_device = t_kopt.device;
  if (tasks.empty()) return {};
  if (table.k < 3U || table.k > 5U || table.templates.empty()) {
    throw std::invalid_argument("HIP k-opt reconnect table is invalid");
  }
  const Coordinates xy(graph);
  std::vector<std::int64_t> costs(tasks.size() * table.templates.size());
  ee_kopt_usage usage{};
  Throw(ee_kopt_eval_costs(ctx, xy.x.data(), xy.y.data(), graph.dimension,
                           static_cast<std::uint8_t>(graph.distance_type), table.k,
                           reinterpret_cast<const ee_endpoint_matching*>(table.templates.data()),
                           static_cast<std::uint32_t>(table.templates.size()), table.generator_hash,
                           reinterpret_cast<const ee_kopt_task*>(tasks.data()), tasks.size(),
                           costs.data(), &usage),
        "HIP k-opt cost matrix");
  if (cache_usage != nullptr) {
    cache_usage->snapshot_hit = usage.snapshot_hit != 0;
    cache_usage->template_hit = usage.template_hit != 0;
    cache_usage->workspace_hit = usage.workspace_hit != 0;
    cache_usage->resident_bytes = usage.resident_bytes;
  }
  return costs;
}

void ClearKOptCostCudaCache() { t_kopt.Reset(); }

// HT reply enumerations (include/cuda_edge_elimination/hamilton_tutte.hpp:529-545)
namespace {

thread_local ThreadContext t_ht;

// The snapshot's edge arrays and a view over them and its CSR.
struct HtGraphView {
  explicit HtGraphView(const GraphSnapshot& graph) : xy(CheckedShape(graph)) {
    const std::size_t m = graph.edges.size();
    if (m > static_cast<std::size_t>(std::numeric_limits<std::int32_t>::max())) {
      throw std::overflow_error("HIP HT reply edge count exceeds int32 indexing");
    }
    u.resize(m);
    v.resize(m);
    w.resize(m);
    active.resize(m);
    for (std::size_t e = 0; e < m; ++e) {
      u[e] = graph.edges[e].u;
      v[e] = graph.edges[e].v;
      w[e] = graph.edges[e].weight;
      active[e] = graph.edges[e].active ? 1 : 0;
    }
    view.dimension = graph.dimension;
    view.distance_type = static_cast<std::uint8_t>(graph.distance_type);
    view.integer_coordinates = graph.integer_coordinates && graph.integer_distance_safe ? 1 : 0;
    view.x = xy.x.data();
    view.y = xy.y.data();
    view.edge_count = static_cast<std::int32_t>(m);
    view.edge_u = u.data();
    view.edge_v = v.data();
    view.edge_weight = w.data();
    view.edge_active = active.data();
    view.row_offsets = graph.row_offsets.data();
    view.neighbors = graph.neighbors.data();
    view.csr_edge_ids = graph.csr_edge_ids.data();
  }
  Coordinates xy;
  std::vector<std::int32_t> u, v;
  std::vector<std::int64_t> w;
  std::vector<std::uint8_t> active;
  ee_graph_view view{};

 private:
  // The container-size and coordinate parts of the reference's
  // ValidateGraphShape (src/cuda/ht_replies.cu:521-531), checked before any
  // row is read; the C ABI checks the CSR contents.  The reference throws
  // invalid_argument for every one of them.
  static const GraphSnapshot& CheckedShape(const GraphSnapshot& graph) {
    const auto n = static_cast<std::size_t>(std::max(graph.dimension, 0));
    if (!graph.integer_coordinates || !graph.integer_distance_safe || graph.dimension < 4 ||
        graph.points.size() != n || graph.row_offsets.size() != n + 1 ||
        graph.row_offsets.front() != 0 || graph.row_offsets.back() < 0 ||
        static_cast<std::size_t>(graph.row_offsets.back()) != graph.neighbors.size() ||
        graph.csr_edge_ids.size() != graph.neighbors.size()) {
      throw std::invalid_argument("HIP HT reply snapshot shape is invalid");
    }
    return graph;
  }
};

void FillUsage(const ee_ht_usage& usage, HtReplyCudaCacheUsage* const cache_usage) {
  if (cache_usage == nullptr) return;
  cache_usage->graph_hit = usage.graph_hit != 0;
  cache_usage->workspace_hit = usage.workspace_hit != 0;
  cache_usage->resident_bytes = usage.resident_bytes;
}

ee_ctx* HtContext(int* const selected_device) {
  std::string reason;
  ee_ctx* ctx = Acquire(&t_ht, &reason);
  if (ctx == nullptr) throw std::runtime_error("HIP HT reply backend unavailable: " + reason);
  if (selected_device != nullptr) *selected_device = t_ht.device;
  return ctx;
}

}  // namespace

static_assert(sizeof(HtNeighborPair) == sizeof(ee_ht_neighbor_pair), "reply layout");
static_assert(offsetof(HtNeighborPair, center) == offsetof(ee_ht_neighbor_pair, center) &&
                  offsetof(HtNeighborPair, first) == offsetof(ee_ht_neighbor_pair, first) &&
                  offsetof(HtNeighborPair, second) == offsetof(ee_ht_neighbor_pair, second),
              "reply layout");
static_assert(sizeof(NodeEdge) == sizeof(ee_node_edge), "edge layout");
static_assert(offsetof(NodeEdge, u) == offsetof(ee_node_edge, u) &&
                  offsetof(NodeEdge, v) == offsetof(ee_node_edge, v),
              "edge layout");
static_assert(sizeof(HtEndReplyTask) == sizeof(ee_ht_end_task), "task layout");

bool HtHamiltonReplyCudaAvailable(std::string* const reason) {
  return Acquire(&t_ht, reason) != nullptr;
}

HtHamiltonReplyDeviceBatch EvaluateHtHamiltonRepliesCuda(const GraphSnapshot& graph,
                                                         const NodeEdge target_edge,
                                                         const std::vector<std::int32_t>& centers,
                                                         int* const selected_device,
                                                         HtReplyCudaCacheUsage* const cache_usage) {
  if (cache_usage != nullptr) *cache_usage = {};
  const HtGraphView g(graph);  // shape checks first, as the reference validates before SelectDevice
  ee_ctx* ctx = HtContext(selected_device);
  HtHamiltonReplyDeviceBatch out;
  out.offsets.assign(centers.size() + 1, 0);
  std::uint64_t total = 0;
  ee_ht_usage usage{};
  // the reply count comes from the device's count pass; the vector is then
  // sized to it and filled in place (same record layout)
  Throw(ee_ht_hamilton_replies(ctx, &g.view, target_edge.u, target_edge.v, centers.data(),
                               centers.size(), out.offsets.data(), nullptr, 0, &total, &usage),
        "HIP HT Hamilton replies");
  out.replies.resize(total);
  Throw(ee_ht_copy_replies(ctx, 0, total, out.replies.data()), "HIP HT Hamilton replies");
  FillUsage(usage, cache_usage);
  return out;
}

bool HtEndReplyCudaAvailable(std::string* const reason) { return Acquire(&t_ht, reason) != nullptr; }

HtEndReplyDeviceBatch EvaluateHtEndRepliesCuda(const GraphSnapshot& graph,
                                               const std::vector<HtEndReplyTask>& tasks,
                                               int* const selected_device,
                                               HtReplyCudaCacheUsage* const cache_usage) {
  if (cache_usage != nullptr) *cache_usage = {};
  const HtGraphView g(graph);  // shape checks first, as the reference validates before SelectDevice
  ee_ctx* ctx = HtContext(selected_device);
  HtEndReplyDeviceBatch out;
  out.offsets.assign(tasks.size() + 1, 0);
  std::uint64_t total = 0;
  ee_ht_usage usage{};
  Throw(ee_ht_end_replies(ctx, &g.view, reinterpret_cast<const ee_ht_end_task*>(tasks.data()),
                          tasks.size(), out.offsets.data(), nullptr, 0, &total, &usage),
        "HIP HT end replies");
  out.replies.resize(total);
  Throw(ee_ht_copy_replies(ctx, 0, total, out.replies.data()), "HIP HT end replies");
  FillUsage(usage, cache_usage);
  return out;
}

void ClearHtReplyCudaCache() { t_ht.Reset(); }

}  // namespace detail
}  // namespace cudaee
