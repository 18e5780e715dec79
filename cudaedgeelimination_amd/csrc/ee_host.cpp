// Host-only parts of the gfx950 backend: the error state of the C entry
// points, input validation shared by every path, and the canonical k-opt
// reconnect template generator.  No HIP runtime call is made here, so the
// unit also builds with a plain host compiler (the sanitizer build in
// oracle/Makefile).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "ee_exact.h"
#include "ee_internal.h"
#include "ee_runtime.h"

namespace ee {

namespace {
thread_local std::string g_last_error;
}  // namespace

void SetLastError(const std::string& message) { g_last_error = message; }
const char* LastError() { return g_last_error.c_str(); }

ee_status StatusFromException() {
  try {
    throw;
  } catch (const Unavailable& e) {
    SetLastError(e.what());
    return EE_ERR_UNAVAILABLE;
  } catch (const std::invalid_argument& e) {
    SetLastError(e.what());
    return EE_ERR_INVALID_ARGUMENT;
  } catch (const std::overflow_error& e) {
    SetLastError(e.what());
    return EE_ERR_OVERFLOW;
  } catch (const std::logic_error& e) {
    SetLastError(e.what());
    return EE_ERR_LOGIC;
  } catch (const std::bad_alloc&) {
    SetLastError("host allocation failed");
    return EE_ERR_RUNTIME;
  } catch (const std::exception& e) {
    SetLastError(e.what());
    return EE_ERR_RUNTIME;
  } catch (...) {
    SetLastError("unknown error");
    return EE_ERR_RUNTIME;
  }
}

double MsSince(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

size_t Grow(size_t current, size_t required) {
  if (current >= required) return current;
  if (current == 0 || current > SIZE_MAX / 2) return required;
  return std::max(required, current * 2);
}

// Bounding box diagonal squared <= UINT64_MAX (src/common/graph.cpp:170-181).
bool DistanceSafe(const int64_t* x, const int64_t* y, int32_t n) {
  if (n <= 0) return true;
  int64_t lx = x[0], hx = x[0], ly = y[0], hy = y[0];
  for (int32_t i = 1; i < n; ++i) {
    lx = std::min(lx, x[i]);
    hx = std::max(hx, x[i]);
    ly = std::min(ly, y[i]);
    hy = std::max(hy, y[i]);
  }
  const unsigned __int128 dx = static_cast<unsigned __int128>(static_cast<__int128>(hx) - lx);
  const unsigned __int128 dy = static_cast<unsigned __int128>(static_cast<__int128>(hy) - ly);
  return dx * dx + dy * dy <= static_cast<unsigned __int128>(UINT64_MAX);
}

// Per-axis coordinate extent below 2^25: every distance of the instance can
// take the FP64 path (ExactDistanceNear).
bool NearExtent(const int64_t* x, const int64_t* y, int32_t n) {
  if (n <= 0) return true;
  int64_t lx = x[0], hx = x[0], ly = y[0], hy = y[0];
  for (int32_t i = 1; i < n; ++i) {
    lx = std::min(lx, x[i]);
    hx = std::max(hx, x[i]);
    ly = std::min(ly, y[i]);
    hy = std::max(hy, y[i]);
  }
  return static_cast<__int128>(hx) - lx < kNearExtent && static_cast<__int128>(hy) - ly < kNearExtent;
}

// ------------------------------------------------------------------- k-opt
namespace {

// Reference generator (src/cpu/kopt_search.cpp:1143-1193 with
// src/cpu/path_system.cpp:47-71, 404-421): all perfect matchings of the 2k
// endpoints in depth-first order (lowest free endpoint paired with each later
// free endpoint, ascending); keep those that form one alternating cycle with
// the component matching {2e+1 - 2e+2, 2k-1 - 0} and never pair 2e with 2e+1.
void Matchings(uint8_t* mate, uint32_t ends, std::vector<std::vector<uint8_t>>* out) {
  uint32_t first = ends;
  for (uint32_t i = 0; i < ends; ++i) {
    if (mate[i] == 0xff) {
      first = i;
      break;
    }
  }
  if (first == ends) {
    out->emplace_back(mate, mate + ends);
    return;
  }
  for (uint32_t second = first + 1; second < ends; ++second) {
    if (mate[second] != 0xff) continue;
    mate[first] = static_cast<uint8_t>(second);
    mate[second] = static_cast<uint8_t>(first);
    Matchings(mate, ends, out);
    mate[first] = mate[second] = 0xff;
  }
}

bool OneAlternatingCycle(const std::vector<uint8_t>& inside, uint32_t k) {
  const uint32_t ends = 2 * k;
  std::vector<uint8_t> outside(ends);
  for (uint32_t e = 0; e + 1 < k; ++e) {
    outside[2 * e + 1] = static_cast<uint8_t>(2 * e + 2);
    outside[2 * e + 2] = static_cast<uint8_t>(2 * e + 1);
  }
  outside[2 * k - 1] = 0;
  outside[0] = static_cast<uint8_t>(2 * k - 1);
  std::vector<bool> seen(ends, false);
  uint32_t at = 0, steps = 0;
  while (!seen[at]) {
    seen[at] = true;
    at = (steps % 2 == 0) ? inside[at] : outside[at];
    ++steps;
  }
  return steps == ends && at == 0;
}

}  // namespace

std::vector<ee_endpoint_matching> CanonicalTemplates(uint32_t k, uint64_t* hash) {
  if (k < 3 || k > 5) throw std::invalid_argument("reconnect template k must be in [3,5]");
  const uint32_t ends = 2 * k;
  std::vector<std::vector<uint8_t>> all;
  std::vector<uint8_t> mate(ends, 0xff);
  Matchings(mate.data(), ends, &all);
  std::vector<ee_endpoint_matching> out;
  for (const auto& m : all) {
    bool proper = OneAlternatingCycle(m, k);
    for (uint32_t e = 0; e < k && proper; ++e) proper = m[2 * e] != 2 * e + 1;
    if (!proper) continue;
    ee_endpoint_matching t{};
    t.endpoint_count = static_cast<uint8_t>(ends);
    std::memset(t.mate, 0xff, sizeof(t.mate));
    std::memcpy(t.mate, m.data(), ends);
    out.push_back(t);
  }
  static const size_t kCounts[3] = {4, 25, 208};
  if (out.size() != kCounts[k - 3]) throw std::logic_error("proper reconnect template count");
  uint64_t h = 14695981039346656037ULL;
  const auto byte = [&h](uint8_t b) {
    h ^= b;
    h *= 1099511628211ULL;
  };
  for (const char ch : std::string("CUDAEE_KOPT_RECONNECT_TEMPLATES_V1")) byte(static_cast<uint8_t>(ch));
  for (const uint32_t v : {k, static_cast<uint32_t>(out.size())}) {
    for (int s = 0; s < 32; s += 8) byte(static_cast<uint8_t>((v >> s) & 0xff));
  }
  for (const auto& t : out) {
    for (uint32_t i = 0; i < t.endpoint_count; ++i) byte(t.mate[i]);
  }
  if (hash) *hash = h;
  return out;
}

// Validates a template table and derives its pair plan.
KoptPlan BuildPlan(uint32_t k, const ee_endpoint_matching* templates, uint32_t count) {
  if (k < 3 || k > 5) throw std::invalid_argument("k-opt reconnect table k must be in [3,5]");
  if (templates == nullptr || count == 0 || count > static_cast<uint32_t>(kMaxTemplates)) {
    throw std::invalid_argument("k-opt reconnect table is empty or too large");
  }
  const uint32_t ends = 2 * k;
  int pair_index[kMaxPorts][kMaxPorts];
  for (auto& row : pair_index) std::fill(row, row + kMaxPorts, -1);
  // at most C(10, 2) = 45 distinct pairs; fixed storage, no heap
  std::pair<uint8_t, uint8_t> pairs[kMaxPairs];
  uint32_t pair_count = 0;
  KoptPlan plan;
  plan.k = k;
  plan.templates = count;
  for (uint32_t t = 0; t < count; ++t) {
    const ee_endpoint_matching& m = templates[t];
    if (m.endpoint_count != ends) throw std::invalid_argument("template endpoint count != 2k");
    uint32_t used = 0;
    for (uint32_t p = 0; p < ends; ++p) {
      const uint32_t q = m.mate[p];
      if (q >= ends || q == p || m.mate[q] != p) {
        throw std::invalid_argument("template is not a perfect matching of the 2k ports");
      }
      if (p > q) continue;
      if (pair_index[p][q] < 0) {
        pair_index[p][q] = static_cast<int>(pair_count);
        pairs[pair_count++] = {static_cast<uint8_t>(p), static_cast<uint8_t>(q)};
      }
      plan.tpl[t] |= static_cast<uint64_t>(pair_index[p][q]) << (8 * used++);
    }
  }
  plan.pairs = pair_count;
  for (uint32_t i = 0; i < pair_count; ++i) {
    plan.pair_pq[i] = static_cast<uint16_t>(pairs[i].first | (pairs[i].second << 8));
  }
  return plan;
}

// ValidateKOptCostTask (src/cpu/kopt_search.cpp:784-807).
void ValidateTasks(int32_t n, uint32_t k, const ee_kopt_task* tasks, size_t count) {
  for (size_t r = 0; r < count; ++r) {
    const ee_kopt_task& t = tasks[r];
    if (t.deleted_cost < 0) throw std::invalid_argument("k-opt task has a negative deleted_cost");
    uint64_t deleted[kMaxK];
    for (uint32_t e = 0; e < k; ++e) {
      const int32_t a = t.port_nodes[2 * e], b = t.port_nodes[2 * e + 1];
      if (a < 0 || b < 0 || a >= n || b >= n || a == b) {
        throw std::invalid_argument("k-opt task has an invalid or repeated deleted edge");
      }
      const uint64_t key = (static_cast<uint64_t>(std::min(a, b)) << 32) |
                           static_cast<uint32_t>(std::max(a, b));
      for (uint32_t f = 0; f < e; ++f) {
        if (deleted[f] == key) {
          throw std::invalid_argument("k-opt task has an invalid or repeated deleted edge");
        }
      }
      deleted[e] = key;
    }
  }
}


}  // namespace ee

using ee::StatusFromException;

extern "C" {

const char* ee_last_error(void) { return ee::LastError(); }

const char* ee_version(void) { return "ee_hip 0.1.0 (gfx950)"; }


ee_status ee_kopt_build_templates(uint32_t k, ee_endpoint_matching* templates, uint32_t capacity,
                                  uint32_t* count, uint64_t* generator_hash) {
  try {
    uint64_t hash = 0;
    const auto table = ee::CanonicalTemplates(k, &hash);
    if (count) *count = static_cast<uint32_t>(table.size());
    if (generator_hash) *generator_hash = hash;
    if (templates) {
      if (capacity < table.size()) throw std::invalid_argument("template capacity too small");
      std::memcpy(templates, table.data(), table.size() * sizeof(ee_endpoint_matching));
    }
    return EE_OK;
  } catch (...) {
    return StatusFromException();
  }
}


}  // extern "C"
