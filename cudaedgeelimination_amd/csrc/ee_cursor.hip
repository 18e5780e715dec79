// Deletion-set producer on the device: the task rows the reference's k-opt
// cost cursor builds on the host, one per combination of k selectable tour
// positions, are built by a kernel from the rank of each combination, and
// fed straight into the plain k-opt evaluation.  No task row crosses PCIe.
//
// Reference: KOptSearchCursor::PrepareNextBlock / PrepareNextK
// (src/cpu/kopt_search.cpp:2123-2156, 2258-2271), AdvanceCombination
// (kopt_search.cpp:736-752), BuildKOptCostTask (kopt_search.cpp:561-584).
#include <algorithm>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <vector>

#include "ee_runtime.h"

namespace ee {
namespace {

// Selectable positions beyond this make C(n, 4) overflow the 64-bit
// binomials below (n (n-1) / 2 * (n-2) (n-3) / 2 must stay below 2^64).
constexpr uint32_t kMaxSelectable = 65536;
constexpr uint32_t kCursorBlock = 256;

// C(n, j) for j in 1..4 and n < 2^16, exact in 64 bits.
__host__ __device__ inline uint64_t Binom(uint64_t n, uint32_t j) {
  if (n < j) return 0;
  switch (j) {
    case 1:
      return n;
    case 2:
      return n * (n - 1) / 2;
    case 3:
      return n * (n - 1) / 2 * (n - 2) / 3;
    default:
      return (n * (n - 1) / 2) * ((n - 2) * (n - 3) / 2) / 6;
  }
}

// Sets the reference cursor enumerates for one k: AdvanceCombination never
// moves index 0, so a set is selectable[0] plus k - 1 of the other
// `labels` = |selectable| - 1 positions, in lexicographic order of their
// indices.  PrepareNextK skips k >= |selectable|.
uint64_t SetCount(uint32_t selectable_count, uint32_t k) {
  if (k < 3 || k > 5 || k >= selectable_count) return 0;
  return Binom(selectable_count - 1, k - 1);
}

// One thread per set.  Set `first + i` is unranked in the combinatorial
// number system: the index of slot s is the largest v whose preceding
// choices (indices prev+1 .. v-1, each followed by C(labels - u, rem)
// completions) number at most the remaining rank; by the hockey-stick
// identity those number C(labels - prev, rem + 1) - C(labels - v + 1, rem + 1).
// The positions are sorted (tour order) and turned into ports and the summed
// deleted cost as BuildKOptCostTask does.
__global__ __launch_bounds__(kCursorBlock) void KoptCursorKernel(
    const int32_t* __restrict__ tour, const int64_t* __restrict__ edge_costs,
    const uint64_t* __restrict__ selectable, uint32_t tour_length, uint32_t labels, uint32_t k,
    uint64_t first, uint32_t count, ee_kopt_task* __restrict__ tasks,
    uint32_t* __restrict__ positions, uint32_t* __restrict__ status) {
  const uint32_t i = blockIdx.x * kCursorBlock + threadIdx.x;
  if (i >= count) return;
  uint64_t rank = first + i;
  uint64_t pos[kMaxK];
  pos[0] = selectable[0];
  uint32_t prev = 0;
  for (uint32_t s = 1; s < k; ++s) {
    const uint32_t rem = k - 1 - s;
    const uint64_t total = Binom(labels - prev, rem + 1);
    uint32_t lo = prev + 1, hi = labels - rem;
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo + 1) / 2;
      if (total - Binom(labels - mid + 1, rem + 1) <= rank) {
        lo = mid;
      } else {
        hi = mid - 1;
      }
    }
    rank -= total - Binom(labels - lo + 1, rem + 1);
    prev = lo;
    pos[s] = selectable[lo];
  }
  for (uint32_t s = 1; s < k; ++s) {  // insertion sort, k <= 5
    const uint64_t v = pos[s];
    uint32_t t = s;
    while (t > 0 && pos[t - 1] > v) {
      pos[t] = pos[t - 1];
      --t;
    }
    pos[t] = v;
  }
  ee_kopt_task task{};
  uint64_t sum = 0;
  bool overflow = false;
  for (uint32_t e = 0; e < k; ++e) {
    const uint64_t p = pos[e];
    const uint64_t next = p + 1 == tour_length ? 0 : p + 1;
    task.port_nodes[2 * e] = tour[p];
    task.port_nodes[2 * e + 1] = tour[next];
    overflow |= __builtin_add_overflow(sum, static_cast<uint64_t>(edge_costs[p]), &sum);
    if (positions != nullptr) positions[static_cast<uint64_t>(i) * k + e] = static_cast<uint32_t>(p);
  }
  overflow |= sum > static_cast<uint64_t>(std::numeric_limits<int64_t>::max());
  if (overflow) atomicOr(status, 1u);
  task.deleted_cost = overflow ? 0 : static_cast<int64_t>(sum);
  tasks[i] = task;
}

void CheckTourView(const ee_tour_view* tour, int32_t n) {
  if (tour == nullptr || tour->tour == nullptr || tour->edge_costs == nullptr ||
      (tour->selectable == nullptr && tour->selectable_count > 0)) {
    throw std::invalid_argument("k-opt cursor: NULL tour view");
  }
  if (tour->tour_length < 3) throw std::invalid_argument("k-opt cursor: a tour needs three nodes");
  if (tour->selectable_count > kMaxSelectable) {
    throw std::invalid_argument("k-opt cursor: more than 65536 selectable positions");
  }
  if (n <= 0) throw std::invalid_argument("k-opt cursor: the snapshot has no nodes");
  // distinct nodes of the snapshot, so that every set's deleted edges are k
  // distinct non-loop edges (the rows pass ValidateKOptCostTask)
  std::vector<uint8_t> seen(static_cast<size_t>(n), 0);
  for (uint32_t p = 0; p < tour->tour_length; ++p) {
    const int32_t v = tour->tour[p];
    if (v < 0 || v >= n || seen[static_cast<size_t>(v)]) {
      throw std::invalid_argument("k-opt cursor: the tour must visit distinct snapshot nodes");
    }
    seen[static_cast<size_t>(v)] = 1;
  }
  std::vector<uint64_t> sorted(tour->selectable, tour->selectable + tour->selectable_count);
  std::sort(sorted.begin(), sorted.end());
  if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end()) {
    throw std::invalid_argument("k-opt cursor: repeated selectable position");
  }
  // BuildKOptCostTask's checks (kopt_search.cpp:568-573), made for every
  // selectable position up front
  for (const uint64_t p : sorted) {
    if (p >= tour->tour_length || tour->edge_costs[p] < 0) {
      throw std::runtime_error("k-opt cursor: deletion position or cached edge cost is invalid");
    }
  }
}

// Keeps the device copy of the tour view; returns whether it was resident.
bool PrepareTour(Context& ctx, const ee_tour_view& t) {
  KoptCursorCache& c = ctx.cursor;
  const size_t L = t.tour_length, S = t.selectable_count;
  const bool hit = c.valid && c.tour.size() == L && c.selectable.size() == S &&
                   std::memcmp(c.tour.data(), t.tour, L * sizeof(int32_t)) == 0 &&
                   std::memcmp(c.edge_costs.data(), t.edge_costs, L * sizeof(int64_t)) == 0 &&
                   (S == 0 || std::memcmp(c.selectable.data(), t.selectable, S * sizeof(uint64_t)) == 0);
  if (hit) return true;
  c.valid = false;
  if (c.d_tour.count() < L) c.d_tour.Allocate(Grow(c.d_tour.count(), L), ctx.device);
  if (c.d_edge_costs.count() < L) c.d_edge_costs.Allocate(Grow(c.d_edge_costs.count(), L), ctx.device);
  if (c.d_selectable.count() < std::max<size_t>(S, 1)) {
    c.d_selectable.Allocate(Grow(c.d_selectable.count(), std::max<size_t>(S, 1)), ctx.device);
  }
  if (c.d_status.count() < 1) {
    // the flag is cleared here and after a reported overflow, so a launch
    // finds it clear without a clearing launch of its own
    c.d_status.Allocate(1, ctx.device);
    CheckHip(hipMemsetAsync(c.d_status.get(), 0, sizeof(uint32_t), ctx.stream), "hipMemsetAsync status");
  }
  c.d_tour.Upload(t.tour, L, ctx.stream);
  c.d_edge_costs.Upload(t.edge_costs, L, ctx.stream);
  c.d_selectable.Upload(t.selectable, S, ctx.stream);
  ctx.Sync();
  c.tour.assign(t.tour, t.tour + L);
  c.edge_costs.assign(t.edge_costs, t.edge_costs + L);
  c.selectable.assign(t.selectable, t.selectable + S);
  c.valid = true;
  return false;
}

void LaunchCursor(Context& ctx, const ee_tour_view& t, uint32_t k, uint64_t first, uint32_t count,
                  ee_kopt_task* d_tasks, bool want_positions) {
  KoptCursorCache& c = ctx.cursor;
  if (want_positions && c.d_positions.count() < static_cast<size_t>(count) * k) {
    c.d_positions.Allocate(Grow(c.d_positions.count(), static_cast<size_t>(count) * k), ctx.device);
  }
  const uint32_t blocks = (count + kCursorBlock - 1) / kCursorBlock;
  hipLaunchKernelGGL(KoptCursorKernel, dim3(blocks), dim3(kCursorBlock), 0, ctx.stream,
                     c.d_tour.get(), c.d_edge_costs.get(), c.d_selectable.get(), t.tour_length,
                     t.selectable_count - 1, k, first, count, d_tasks,
                     want_positions ? c.d_positions.get() : nullptr, c.d_status.get());
  CheckHip(hipGetLastError(), "KoptCursorKernel launch");
}

void FinishCursor(Context& ctx, uint32_t k, uint32_t count, ee_kopt_task* tasks,
                  uint32_t* positions) {
  KoptCursorCache& c = ctx.cursor;
  uint32_t status = 0;
  c.d_status.Download(&status, 1, ctx.stream);
  if (tasks != nullptr) ctx.kopt.d_tasks.Download(tasks, count, ctx.stream);
  if (positions != nullptr) c.d_positions.Download(positions, static_cast<size_t>(count) * k, ctx.stream);
  ctx.Sync();
  if (status & 1u) {
    CheckHip(hipMemsetAsync(c.d_status.get(), 0, sizeof(uint32_t), ctx.stream), "hipMemsetAsync status");
    ctx.Sync();
    throw std::runtime_error("k-opt cursor: deleted cost sum overflows int64_t");
  }
}

void CursorEval(Context& ctx, const int64_t* x, const int64_t* y, int32_t n, uint8_t type,
                const ee_tour_view* tour, uint32_t k, const ee_endpoint_matching* templates,
                uint32_t template_count, uint64_t generator_hash, uint64_t first, uint32_t count,
                ee_kopt_task* tasks, uint32_t* positions, int64_t* costs, uint32_t* masks,
                uint32_t* produced, ee_kopt_usage* usage) {
  ctx.Activate();
  if (produced) *produced = 0;
  if (usage) *usage = ee_kopt_usage{};
  if (costs != nullptr && masks != nullptr) {
    throw std::invalid_argument("k-opt cursor: costs and masks are exclusive");
  }
  if (k < 3 || k > 5) throw std::invalid_argument("k-opt cursor: k must lie in [3, 5]");
  CheckTourView(tour, n);
  const uint64_t total = SetCount(tour->selectable_count, k);
  const uint32_t rows = first >= total ? 0 : static_cast<uint32_t>(std::min<uint64_t>(count, total - first));
  if (rows == 0) return;
  PrepareTour(ctx, *tour);
  const bool want_positions = positions != nullptr;
  try {
    if (costs == nullptr && masks == nullptr) {  // producer only
      KoptCache& kc = ctx.kopt;
      if (kc.d_tasks.count() < rows) kc.d_tasks.Allocate(Grow(kc.d_tasks.count(), rows), ctx.device);
      LaunchCursor(ctx, *tour, k, first, rows, kc.d_tasks.get(), want_positions);
    } else {
      const KoptTaskSource produce = [&](ee_kopt_task* d_tasks) {
        LaunchCursor(ctx, *tour, k, first, rows, d_tasks, want_positions);
      };
      EvalKoptCostsFrom(ctx, x, y, n, type, k, templates, template_count, generator_hash, rows,
                        produce, false, costs, masks, usage);
    }
  } catch (...) {
    // a failed evaluation may leave the overflow flag set; clear it (best
    // effort) so that the next call does not report this one's overflow
    static_cast<void>(hipMemsetAsync(ctx.cursor.d_status.get(), 0, sizeof(uint32_t), ctx.stream));
    static_cast<void>(hipStreamSynchronize(ctx.stream));
    throw;
  }
  FinishCursor(ctx, k, rows, tasks, positions);
  if (produced) *produced = rows;
}

template <typename F> ee_status Guarded(F&& body) {
  try {
    body();
    return EE_OK;
  } catch (...) {
    return StatusFromException();
  }
}

}  // namespace

uint64_t CursorSetCount(uint32_t selectable_count, uint32_t k) { return SetCount(selectable_count, k); }

void CheckCursorTour(const ee_tour_view* tour, int32_t n) { CheckTourView(tour, n); }

void LaunchCursorRows(const int32_t* d_tour, const int64_t* d_edge_costs, const uint64_t* d_selectable,
                      uint32_t tour_length, uint32_t selectable_count, uint32_t k, uint64_t first,
                      uint32_t count, ee_kopt_task* d_tasks, uint32_t* d_positions, uint32_t* d_status,
                      hipStream_t stream) {
  if (count == 0) return;
  const uint32_t blocks = (count + kCursorBlock - 1) / kCursorBlock;
  hipLaunchKernelGGL(KoptCursorKernel, dim3(blocks), dim3(kCursorBlock), 0, stream, d_tour, d_edge_costs,
                     d_selectable, tour_length, selectable_count - 1, k, first, count, d_tasks, d_positions,
                     d_status);
  CheckHip(hipGetLastError(), "KoptCursorKernel launch");
}

}  // namespace ee

extern "C" {

ee_status ee_kopt_cursor_count(const ee_tour_view* tour, uint32_t k, uint64_t* count) {
  return ee::Guarded([&] {
    if (tour == nullptr || count == nullptr) throw std::invalid_argument("k-opt cursor: NULL argument");
    if (tour->selectable_count > ee::kMaxSelectable) {
      throw std::invalid_argument("k-opt cursor: more than 65536 selectable positions");
    }
    *count = ee::SetCount(tour->selectable_count, k);
  });
}

ee_status ee_kopt_cursor_eval(ee_ctx* ctx, const int64_t* x, const int64_t* y, int32_t dimension,
                              uint8_t distance_type, const ee_tour_view* tour, uint32_t k,
                              const ee_endpoint_matching* templates, uint32_t template_count,
                              uint64_t generator_hash, uint64_t first, uint32_t count,
                              ee_kopt_task* tasks, uint32_t* positions, int64_t* costs,
                              uint32_t* masks, uint32_t* produced, ee_kopt_usage* usage) {
  return ee::Guarded([&] {
    if (ctx == nullptr) throw std::invalid_argument("NULL context");
    ee::CursorEval(*reinterpret_cast<ee::Context*>(ctx), x, y, dimension, distance_type, tour, k,
                   templates, template_count, generator_hash, first, count, tasks, positions, costs,
                   masks, produced, usage);
  });
}

}  // extern "C"
