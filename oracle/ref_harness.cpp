// C entry points over the reference's own CPU path, for checking the oracle.
//
// TEST INFRASTRUCTURE ONLY.  This file is compiled against the reference's
// public headers (include/cuda_edge_elimination/*.hpp) and linked with the
// reference CPU sources built where they lie (see the Makefile next to it);
// the result lands in oracle/_ref/ and is never part of the product.
#include "cuda_edge_elimination/distance.hpp"
#include "cuda_edge_elimination/elimination.hpp"
#include "cuda_edge_elimination/graph.hpp"
#include "cuda_edge_elimination/local_search.hpp"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <string>
#include <unistd.h>
#include <vector>

namespace {

thread_local std::string g_error;

template <typename F> int Guard(F&& body) {
  try {
    body();
    return 0;
  } catch (const std::exception& e) {
    g_error = e.what();
    return -1;
  }
}

// Same construction as the reference unit tests' hand-built snapshots
// (tests/unit/test_core.cpp:86-98): integer coordinates, edges as given.
cudaee::GraphSnapshot FromArrays(int32_t n, int type, const int64_t* x, const int64_t* y,
                                 int32_t m, const int32_t* u, const int32_t* v, const int64_t* w,
                                 const uint8_t* active) {
  cudaee::GraphSnapshot g;
  g.dimension = n;
  g.distance_type = type == 0 ? cudaee::DistanceType::kEuc2D : cudaee::DistanceType::kCeil2D;
  g.integer_coordinates = true;
  g.integer_distance_safe = true;
  for (int32_t i = 0; i < n; ++i) {
    g.points.push_back({static_cast<double>(x[i]), static_cast<double>(y[i]), x[i], y[i]});
  }
  for (int32_t i = 0; i < m; ++i) {
    g.edges.push_back({u[i], v[i], w[i], active == nullptr || active[i] != 0});
  }
  g.RebuildCsr();
  return g;
}

std::string TempPath(const char* tag) {
  static int counter = 0;
  return "/tmp/ee_ref_" + std::to_string(::getpid()) + "_" + std::to_string(counter++) + tag;
}

std::string Slurp(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  std::ostringstream s;
  s << in.rdbuf();
  return s.str();
}

int64_t CopyOut(const std::string& text, char* out, size_t cap) {
  if (out != nullptr && cap > 0) {
    const size_t n = std::min(text.size(), cap - 1);
    std::memcpy(out, text.data(), n);
    out[n] = '\0';
  }
  return static_cast<int64_t>(text.size());
}

}  // namespace

extern "C" {

const char* ref_last_error(void) { return g_error.c_str(); }

uint64_t ref_isqrt_floor(uint64_t value) { return cudaee::IntegerSqrtFloor(value); }

int ref_exact_distance(int64_t ax, int64_t ay, int64_t bx, int64_t by, int type, int64_t* out) {
  cudaee::Point a{static_cast<double>(ax), static_cast<double>(ay), ax, ay};
  cudaee::Point b{static_cast<double>(bx), static_cast<double>(by), bx, by};
  std::string error;
  if (!cudaee::ExactIntegerDistance(
          a, b, type == 0 ? cudaee::DistanceType::kEuc2D : cudaee::DistanceType::kCeil2D, out,
          &error)) {
    g_error = error;
    return -1;
  }
  return 0;
}

uint64_t ref_graph_hash(int32_t n, int type, const int64_t* x, const int64_t* y, int32_t m,
                        const int32_t* u, const int32_t* v, const int64_t* w,
                        const uint8_t* active) {
  return FromArrays(n, type, x, y, m, u, v, w, active).ContentHash();
}

// GraphSnapshot::Load then FindJvCandidatesCpu; witnesses[m] = -1 or witness.
int64_t ref_jv_candidates_files(const char* tsp, const char* edg, int32_t* witnesses,
                                int32_t capacity) {
  int64_t found = -1;
  Guard([&] {
    const auto g = cudaee::GraphSnapshot::Load(tsp, edg);
    if (static_cast<size_t>(capacity) < g.edges.size()) throw std::runtime_error("capacity");
    for (size_t i = 0; i < g.edges.size(); ++i) witnesses[i] = -1;
    const auto c = cudaee::FindJvCandidatesCpu(g);
    for (const auto& cand : c) witnesses[cand.edge_id] = cand.witness;
    found = static_cast<int64_t>(c.size());
  });
  return found;
}

int64_t ref_jv_candidates(int32_t n, int type, const int64_t* x, const int64_t* y, int32_t m,
                          const int32_t* u, const int32_t* v, const int64_t* w,
                          const uint8_t* active, int32_t* witnesses) {
  int64_t found = -1;
  Guard([&] {
    const auto g = FromArrays(n, type, x, y, m, u, v, w, active);
    for (int32_t i = 0; i < m; ++i) witnesses[i] = -1;
    const auto c = cudaee::FindJvCandidatesCpu(g);
    for (const auto& cand : c) witnesses[cand.edge_id] = cand.witness;
    found = static_cast<int64_t>(c.size());
  });
  return found;
}

// GraphSnapshot::Load, RunJvElimination(kCpu), WriteProof: the proof text.
int64_t ref_jv_run_files(const char* tsp, const char* edg, uint32_t max_rounds, char* proof,
                         size_t cap) {
  int64_t size = -1;
  Guard([&] {
    auto g = cudaee::GraphSnapshot::Load(tsp, edg);
    const auto result = cudaee::RunJvElimination(&g, cudaee::Backend::kCpu, max_rounds);
    const std::string path = TempPath(".proof");
    cudaee::WriteProof(path, result);
    size = CopyOut(Slurp(path), proof, cap);
    std::filesystem::remove(path);
  });
  return size;
}

// ReadProof + ReplayProof on the graph loaded from files.  0 = accepted.
int ref_jv_replay_files(const char* tsp, const char* edg, const char* proof_text) {
  return Guard([&] {
    auto g = cudaee::GraphSnapshot::Load(tsp, edg);
    const std::string path = TempPath(".proof");
    {
      std::ofstream out(path, std::ios::binary);
      out << proof_text;
    }
    try {
      const auto expected = cudaee::ReadProof(path);
      static_cast<void>(cudaee::ReplayProof(&g, expected));
    } catch (...) {
      std::filesystem::remove(path);
      throw;
    }
    std::filesystem::remove(path);
  });
}

int ref_kopt_templates(uint32_t k, uint8_t* mates, uint32_t capacity, uint32_t* count,
                       uint64_t* hash) {
  return Guard([&] {
    const auto table = cudaee::BuildKOptReconnectTable(k);
    *count = static_cast<uint32_t>(table.templates.size());
    *hash = table.generator_hash;
    if (mates != nullptr) {
      if (capacity < table.templates.size()) throw std::runtime_error("capacity");
      for (size_t i = 0; i < table.templates.size(); ++i) {
        std::memcpy(mates + 14 * i, table.templates[i].mate.data(), 14);
      }
    }
  });
}

// EvaluateKOptTemplateCosts(..., kCpu): the reference CPU exact cost matrix.
int ref_kopt_costs(const int64_t* x, const int64_t* y, int32_t n, int type, uint32_t k,
                   const void* tasks, size_t task_count, int64_t* costs, uint32_t* threads_used,
                   uint32_t* cache_nodes) {
  static_assert(sizeof(cudaee::KOptCostTask) == 48, "task layout");
  return Guard([&] {
    cudaee::GraphSnapshot g;
    g.dimension = n;
    g.distance_type = type == 0 ? cudaee::DistanceType::kEuc2D : cudaee::DistanceType::kCeil2D;
    g.integer_coordinates = true;
    g.integer_distance_safe = true;
    for (int32_t i = 0; i < n; ++i) {
      g.points.push_back({static_cast<double>(x[i]), static_cast<double>(y[i]), x[i], y[i]});
    }
    std::vector<cudaee::KOptCostTask> t(task_count);
    std::memcpy(t.data(), tasks, task_count * sizeof(cudaee::KOptCostTask));
    const auto r = cudaee::EvaluateKOptTemplateCosts(g, k, t, cudaee::PathCompatibilityBackend::kCpu);
    std::memcpy(costs, r.added_costs.data(), r.added_costs.size() * sizeof(int64_t));
    if (threads_used != nullptr) *threads_used = r.cpu_threads_used;
    if (cache_nodes != nullptr) *cache_nodes = r.cpu_distance_cache_nodes;
  });
}


// The reference's CPU HT replies (EvaluateHtHamiltonReplies /
// EvaluateHtEndReplies with PathCompatibilityBackend::kCpu, include/
// cuda_edge_elimination/hamilton_tutte.hpp:409-420), flattened like the
// oracle's outputs.  Returns the total reply count, negative on failure.
int64_t ref_ht_hamilton_replies(int32_t n, int type, const int64_t* x, const int64_t* y, int32_t m,
                                const int32_t* u, const int32_t* v, const int64_t* w,
                                const uint8_t* active, int32_t tu, int32_t tv,
                                const int32_t* centers, size_t count, uint64_t* offsets,
                                int32_t* replies, uint64_t capacity) {
  int64_t total = 0;
  const int status = Guard([&] {
    const cudaee::GraphSnapshot g = FromArrays(n, type, x, y, m, u, v, w, active);
    const std::vector<int32_t> c(centers, centers + count);
    const auto r = cudaee::EvaluateHtHamiltonReplies(g, cudaee::NodeEdge{tu, tv}, c,
                                                     cudaee::PathCompatibilityBackend::kCpu);
    std::copy(r.offsets.begin(), r.offsets.end(), offsets);
    for (size_t i = 0; i < r.replies.size() && i < capacity; ++i) {
      replies[3 * i] = r.replies[i].center;
      replies[3 * i + 1] = r.replies[i].first;
      replies[3 * i + 2] = r.replies[i].second;
    }
    total = static_cast<int64_t>(r.replies.size());
  });
  return status == 0 ? total : status;
}

int64_t ref_ht_end_replies(int32_t n, int type, const int64_t* x, const int64_t* y, int32_t m,
                           const int32_t* u, const int32_t* v, const int64_t* w,
                           const uint8_t* active, const int32_t* tasks, size_t count,
                           uint64_t* offsets, int32_t* replies, uint64_t capacity) {
  int64_t total = 0;
  const int status = Guard([&] {
    const cudaee::GraphSnapshot g = FromArrays(n, type, x, y, m, u, v, w, active);
    std::vector<cudaee::HtEndReplyTask> t(count);
    for (size_t i = 0; i < count; ++i) t[i] = {tasks[2 * i], tasks[2 * i + 1]};
    const auto r = cudaee::EvaluateHtEndReplies(g, t, cudaee::PathCompatibilityBackend::kCpu);
    std::copy(r.offsets.begin(), r.offsets.end(), offsets);
    for (size_t i = 0; i < r.replies.size() && i < capacity; ++i) {
      replies[2 * i] = r.replies[i].u;
      replies[2 * i + 1] = r.replies[i].v;
    }
    total = static_cast<int64_t>(r.replies.size());
  });
  return status == 0 ? total : status;
}

}  // extern "C"
