// Hamilton-Tutte reply enumerations on the resident CSR: the Hamilton
// replies (pairs of a center's neighbours that survive the target edge's
// 2-opt and cost tests) and the end replies (a path end's active edges but
// the path-internal one), in the reference's deterministic order.
//
// Reference: CountHtRepliesKernel / WriteHtRepliesKernel /
// CountEndRepliesKernel / WriteEndRepliesKernel and their host drivers
// (src/cuda/ht_replies.cu:406-520, 585-752); CPU semantics
// AppendHtHamiltonRepliesUnchecked / AppendHtEndRepliesUnchecked
// (src/cpu/hamilton_tutte.cpp:230-277) with EdgesSurviveTwoOpt (:79-87).
//
// One wavefront per center.  Its lanes first compute, once per neighbour,
// the center edge's exact distance and its 2-opt verdict into LDS (the
// reference recomputes both per pair); then each lane takes one neighbour
// pair (pairs numbered in the reference's first-major order), and a ballot
// plus a popcount prefix place the surviving pairs in order.  The count and
// write passes run the same wave body.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "ee_exact.h"
#include "ee_runtime.h"

namespace ee {
namespace {

constexpr uint32_t kHtBlock = 256;
constexpr uint32_t kHtWaves = kHtBlock / 64;
constexpr uint32_t kHtLdsNeighbors = 64;  // per wave; larger rows recompute per pair

struct HtNeighbor {
  int64_t center_cost;
  int32_t node;
  int32_t survives;
};

// kNear: both coordinate extents of the snapshot are below 2^25, so the
// differences fit int32 and the FP64 root with one exact remainder applies
// (ExactDistanceNear); otherwise the 64-bit integer path.
template <bool kNear>
__device__ inline int64_t Dist(const longlong2* xy, int32_t a, int32_t b, int type) {
  const longlong2 p = xy[a], q = xy[b];
  if constexpr (kNear) {
    return ExactDistanceNear(static_cast<int32_t>(p.x - q.x), static_cast<int32_t>(p.y - q.y), type);
  } else {
    return ExactDistance(p.x, p.y, q.x, q.y, type);
  }
}

// EdgesSurviveTwoOpt for the edge (a, b) against the target (tu, tv): at
// least one 2-opt orientation is not a strict improvement.  Distances of a
// distance-safe snapshot are below 2^32, so the 64-bit sums are exact.
template <bool kNear>
__device__ inline bool Survives(const longlong2* xy, int32_t tu, int32_t tv, int64_t target_cost,
                                int32_t a, int32_t b, int64_t ab, int type) {
  const int32_t lo = min(a, b), hi = max(a, b);
  const int64_t original = target_cost + ab;
  const int64_t first = Dist<kNear>(xy, tu, hi, type) + Dist<kNear>(xy, tv, lo, type);
  const int64_t second = Dist<kNear>(xy, tu, lo, type) + Dist<kNear>(xy, tv, hi, type);
  return first >= original || second >= original;
}

// First index of row i of the pairs (i, j > i) of d items, first-major.
__device__ inline uint64_t PairRowStart(uint64_t i, uint64_t d) { return i * (2 * d - i - 1) / 2; }

// Centers per scan chunk: one workgroup sums each chunk's reply counts, and
// the scan pass gives each chunk to one workgroup.
constexpr uint32_t kScanThreads = 256;
constexpr uint32_t kScanPerThread = 4;
constexpr uint32_t kScanChunk = kScanThreads * kScanPerThread;

template <bool kWrite, bool kNear>
__global__ __launch_bounds__(kHtBlock) void HtHamiltonKernel(
    const longlong2* __restrict__ xy, const int32_t* __restrict__ row_offsets,
    const int32_t* __restrict__ neighbors, int type, int32_t tu, int32_t tv,
    const int32_t* __restrict__ centers, uint64_t center_count, uint64_t* __restrict__ counts,
    const uint64_t* __restrict__ offsets, int32_t* __restrict__ replies) {
  __shared__ HtNeighbor cache[kHtWaves][kHtLdsNeighbors];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  const uint64_t item = static_cast<uint64_t>(blockIdx.x) * kHtWaves + wave;
  if (item >= center_count) return;  // whole wavefronts leave together
  const int32_t c = centers[item];
  const int32_t begin = row_offsets[c];
  const uint64_t d = static_cast<uint64_t>(row_offsets[c + 1] - begin);
  const int64_t target_cost = Dist<kNear>(xy, tu, tv, type);
  const int64_t center_target = Dist<kNear>(xy, tu, c, type) + Dist<kNear>(xy, tv, c, type);
  const bool cached = d <= kHtLdsNeighbors;
  HtNeighbor* mine = cache[wave];
  if (cached && lane < d) {
    const int32_t nb = neighbors[begin + lane];
    const int64_t cost = Dist<kNear>(xy, c, nb, type);
    mine[lane] = {cost, nb, Survives<kNear>(xy, tu, tv, target_cost, c, nb, cost, type) ? 1 : 0};
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const uint64_t pairs = d < 2 ? 0 : d * (d - 1) / 2;
  uint64_t out = kWrite ? offsets[item] : 0;
  const uint64_t out_end = kWrite ? offsets[item + 1] : 0;
  uint64_t found = 0;
  for (uint64_t base = 0; base < pairs; base += 64) {
    const uint64_t p = base + lane;
    bool keep = false;
    if (p < pairs) {
      // row i of pair p: the root estimate, then exact corrections
      const double b = static_cast<double>(2 * d - 1);
      uint64_t i = static_cast<uint64_t>(fmax(0.0, floor((b - sqrt(b * b - 8.0 * static_cast<double>(p))) / 2)));
      if (i > d - 2) i = d - 2;
      while (i > 0 && PairRowStart(i, d) > p) --i;
      while (i + 1 <= d - 2 && PairRowStart(i + 1, d) <= p) ++i;
      const uint64_t j = p - PairRowStart(i, d) + i + 1;
      HtNeighbor f, s;
      if (cached) {
        f = mine[i];
        s = mine[j];
      } else {
        f.node = neighbors[begin + i];
        s.node = neighbors[begin + j];
        f.center_cost = Dist<kNear>(xy, c, f.node, type);
        s.center_cost = Dist<kNear>(xy, c, s.node, type);
        f.survives = Survives<kNear>(xy, tu, tv, target_cost, c, f.node, f.center_cost, type);
        s.survives = Survives<kNear>(xy, tu, tv, target_cost, c, s.node, s.center_cost, type);
      }
      const bool is_target = min(f.node, s.node) == tu && max(f.node, s.node) == tv;
      if (!is_target && f.survives && s.survives) {
        const int64_t original = target_cost + f.center_cost + s.center_cost;
        const int64_t replacement = Dist<kNear>(xy, f.node, s.node, type) + center_target;
        keep = original <= replacement;
      }
      if (kWrite && keep) {
        const uint64_t ballot_below = __ballot(keep) & ((uint64_t{1} << lane) - 1);
        const uint64_t at = out + __popcll(ballot_below);
        if (at < out_end) {  // the count pass found exactly out_end - offsets[item]
          int32_t* r = replies + 3 * at;
          r[0] = c;
          r[1] = f.node;
          r[2] = s.node;
        }
      }
    }
    const uint64_t ballot = __ballot(keep);
    found += __popcll(ballot);
    if (kWrite) out += __popcll(ballot);
  }
  if (!kWrite && lane == 0) counts[item] = found;
}

// chunk_sums[c] = counts of chunk c (kScanChunk centers), one workgroup per
// chunk: plain stores, no atomics shared by the chunk's centers.
__global__ __launch_bounds__(kScanThreads) void HtChunkSumKernel(const uint64_t* __restrict__ counts, uint64_t n,
                                                                  unsigned long long* __restrict__ chunk_sums) {
  __shared__ unsigned long long wave_sums[kScanThreads / 64];
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  const uint64_t first = static_cast<uint64_t>(blockIdx.x) * kScanChunk + static_cast<uint64_t>(t) * kScanPerThread;
  unsigned long long own = 0;
#pragma unroll
  for (uint32_t e = 0; e < kScanPerThread; ++e) own += first + e < n ? counts[first + e] : 0;
  for (uint32_t d = 32; d >= 1; d >>= 1) own += __shfl_xor(own, d, 64);
  if (lane == 0) wave_sums[wave] = own;
  __syncthreads();
  if (t == 0) {
    unsigned long long total = 0;
    for (uint32_t w = 0; w < kScanThreads / 64; ++w) total += wave_sums[w];
    chunk_sums[blockIdx.x] = total;
  }
}

// offsets[0] = 0, offsets[i + 1] = counts[0] + ... + counts[i], one
// workgroup per chunk of kScanChunk counts: the workgroup first sums the
// totals of the chunks before its own (chunk_sums, from HtChunkSumKernel),
// then scans its chunk, each thread holding 4 consecutive counts
// (wavefront shuffles, then the 4 wavefront totals in LDS).
__global__ __launch_bounds__(kScanThreads) void HtScanKernel(const uint64_t* __restrict__ counts,
                                                              uint64_t n,
                                                              const unsigned long long* __restrict__ chunk_sums,
                                                              uint64_t* __restrict__ offsets) {
  __shared__ unsigned long long partial[kScanThreads / 64];
  __shared__ unsigned long long wave_sums[kScanThreads / 64];
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  const uint64_t chunk = blockIdx.x;
  // prefix of the chunks before this one
  unsigned long long before = 0;
  for (uint64_t c = t; c < chunk; c += kScanThreads) before += chunk_sums[c];
  for (uint32_t d = 32; d >= 1; d >>= 1) before += __shfl_xor(before, d, 64);
  if (lane == 0) partial[wave] = before;
  const uint64_t first = chunk * kScanChunk + static_cast<uint64_t>(t) * kScanPerThread;
  unsigned long long v[kScanPerThread];
  unsigned long long own = 0;
#pragma unroll
  for (uint32_t e = 0; e < kScanPerThread; ++e) {
    v[e] = first + e < n ? counts[first + e] : 0;
    own += v[e];
  }
  unsigned long long inc = own;  // inclusive scan of the threads' sums in the wavefront
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const unsigned long long o = __shfl_up(inc, d, 64);
    if (lane >= d) inc += o;
  }
  if (lane == 63) wave_sums[wave] = inc;
  __syncthreads();
  unsigned long long run = inc - own;
  for (uint32_t w = 0; w < kScanThreads / 64; ++w) {
    run += partial[w];
    if (w < wave) run += wave_sums[w];
  }
  if (chunk == 0 && t == 0) offsets[0] = 0;
#pragma unroll
  for (uint32_t e = 0; e < kScanPerThread; ++e) {
    run += v[e];
    if (first + e < n) offsets[first + e + 1] = run;
  }
}

// End replies: one thread per task writes the endpoint's active edges but
// the internal one, canonical (u < v), from its precomputed offset.
__global__ __launch_bounds__(kHtBlock) void HtEndKernel(const int32_t* __restrict__ row_offsets,
                                                         const int32_t* __restrict__ neighbors,
                                                         const int32_t* __restrict__ tasks,
                                                         uint64_t task_count,
                                                         const uint64_t* __restrict__ offsets,
                                                         int32_t* __restrict__ replies) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kHtBlock + threadIdx.x;
  if (i >= task_count) return;
  const int32_t e = tasks[2 * i], internal = tasks[2 * i + 1];
  uint64_t out = offsets[i];
  const uint64_t end = offsets[i + 1];
  for (int32_t a = row_offsets[e]; a < row_offsets[e + 1] && out < end; ++a) {
    const int32_t nb = neighbors[a];
    if (nb == internal) continue;
    replies[2 * out] = min(e, nb);
    replies[2 * out + 1] = max(e, nb);
    ++out;
  }
}

bool RowHas(const ee_graph_view& g, int32_t a, int32_t b) {
  const int32_t* lo = g.neighbors + g.row_offsets[a];
  const int32_t* hi = g.neighbors + g.row_offsets[a + 1];
  return std::binary_search(lo, hi, b);
}

// ValidateGraphShape (src/cuda/ht_replies.cu:521-553): distance type and
// integer, distance-safe coordinates; n >= 4; CSR offsets in bounds and
// ascending; rows strictly ascending, in range, without self loops, every
// entry mirrored in its neighbour's row; and the CSR holding exactly the
// active edges.  Every kernel access depends on it.
void ValidateHtGraph(const ee_graph_view& g) {
  if (g.dimension < 4) throw std::invalid_argument("HT reply snapshot needs at least four nodes");
  if (!DistanceSafe(g.x, g.y, g.dimension)) {
    throw std::invalid_argument("HT reply snapshot needs distance-safe integer coordinates");
  }
  const int32_t n = g.dimension;
  if (g.row_offsets[0] != 0) throw std::invalid_argument("CSR row_offsets[0] must be 0");
  for (int32_t v = 0; v < n; ++v) {
    if (g.row_offsets[v + 1] < g.row_offsets[v]) {
      throw std::invalid_argument("CSR row_offsets must be non-decreasing");
    }
  }
  int64_t active = 0;
  for (int32_t e = 0; e < g.edge_count; ++e) active += g.edge_active[e] != 0;
  if (static_cast<int64_t>(g.row_offsets[n]) != 2 * active) {
    throw std::invalid_argument("HT reply CSR must hold exactly the active edges");
  }
  for (int32_t v = 0; v < n; ++v) {
    int32_t previous = -1;
    for (int32_t a = g.row_offsets[v]; a < g.row_offsets[v + 1]; ++a) {
      const int32_t nb = g.neighbors[a];
      if (nb <= previous || nb < 0 || nb >= n || nb == v) {
        throw std::invalid_argument("HT reply CSR row is not strictly ascending and in range");
      }
      previous = nb;
    }
  }
  for (int32_t v = 0; v < n; ++v) {
    for (int32_t a = g.row_offsets[v]; a < g.row_offsets[v + 1]; ++a) {
      if (!RowHas(g, g.neighbors[a], v)) throw std::invalid_argument("HT reply CSR is not symmetric");
    }
  }
}

// Keeps the device copy of the snapshot.  A snapshot equal to the resident
// one (coordinates, CSR and active flags) passed the HT checks when it was
// uploaded, so only a new one is checked (ValidateHtGraph) and uploaded.
bool PrepareHtGraph(Context& ctx, const ee_graph_view& g) {
  HtCache& c = ctx.ht;
  const size_t n = static_cast<size_t>(g.dimension);
  const size_t m = static_cast<size_t>(g.edge_count);
  const bool hit = c.valid && c.dimension == g.dimension && c.distance_type == g.distance_type &&
                   c.active.size() == m && c.row_offsets.size() == n + 1 &&
                   std::memcmp(c.row_offsets.data(), g.row_offsets, (n + 1) * 4) == 0 &&
                   std::memcmp(c.x.data(), g.x, n * 8) == 0 &&
                   std::memcmp(c.y.data(), g.y, n * 8) == 0 &&
                   (c.neighbors.empty() ||
                    std::memcmp(c.neighbors.data(), g.neighbors, c.neighbors.size() * 4) == 0) &&
                   (m == 0 || std::memcmp(c.active.data(), g.edge_active, m) == 0);
  if (hit) return true;
  ValidateHtGraph(g);
  const size_t a = static_cast<size_t>(g.row_offsets[g.dimension]);
  c.valid = false;
  std::vector<longlong2> xy(n);
  for (size_t i = 0; i < n; ++i) xy[i] = make_longlong2(g.x[i], g.y[i]);
  c.d_xy.Allocate(n, ctx.device);
  c.d_row_offsets.Allocate(n + 1, ctx.device);
  c.d_neighbors.Allocate(std::max<size_t>(a, 1), ctx.device);
  c.d_xy.Upload(xy.data(), n, ctx.stream);
  c.d_row_offsets.Upload(g.row_offsets, n + 1, ctx.stream);
  c.d_neighbors.Upload(g.neighbors, a, ctx.stream);
  ctx.Sync();
  c.x.assign(g.x, g.x + n);
  c.y.assign(g.y, g.y + n);
  c.row_offsets.assign(g.row_offsets, g.row_offsets + n + 1);
  c.neighbors.assign(g.neighbors, g.neighbors + a);
  c.active.assign(g.edge_active, g.edge_active + m);
  c.near = NearExtent(g.x, g.y, g.dimension);
  c.dimension = g.dimension;
  c.distance_type = g.distance_type;
  c.valid = true;
  return false;
}

template <typename T> bool Reserve(DeviceArray<T>& a, size_t count, int device) {
  if (a.count() >= std::max<size_t>(count, 1)) return true;
  a.Allocate(Grow(a.count(), std::max<size_t>(count, 1)), device);
  return false;
}

uint64_t HtResidentBytes(const HtCache& c) {
  return c.d_xy.bytes() + c.d_row_offsets.bytes() + c.d_neighbors.bytes() + c.d_items.bytes() +
         c.d_counts.bytes() + c.d_offsets.bytes() + c.d_chunk_sums.bytes() + c.d_replies.bytes();
}

void HamiltonReplies(Context& ctx, const ee_graph_view* graph, int32_t target_u, int32_t target_v,
                     const int32_t* centers, size_t count, uint64_t* offsets,
                     ee_ht_neighbor_pair* replies, uint64_t capacity, uint64_t* total,
                     ee_ht_usage* usage) {
  ee_ht_usage u{};
  ctx.Activate();
  CheckGraphView(graph);
  const ee_graph_view& g = *graph;
  if (offsets == nullptr || total == nullptr || (count > 0 && centers == nullptr) ||
      (capacity > 0 && replies == nullptr)) {
    throw std::invalid_argument("HT Hamilton replies: NULL argument");
  }
  auto t0 = std::chrono::steady_clock::now();
  u.graph_hit = PrepareHtGraph(ctx, g) ? 1 : 0;  // checks a new snapshot
  const int32_t tu = std::min(target_u, target_v), tv = std::max(target_u, target_v);
  if (tu < 0 || tv >= g.dimension || tu == tv || !RowHas(g, tu, tv)) {
    throw std::invalid_argument("HT Hamilton replies: the target is not an active edge");
  }
  for (size_t i = 0; i < count; ++i) {
    const int32_t c = centers[i];
    if (c < 0 || c >= g.dimension || c == tu || c == tv) {
      throw std::invalid_argument("HT Hamilton replies: invalid center");
    }
  }
  offsets[0] = 0;
  *total = 0;
  HtCache& c = ctx.ht;
  c.last_kind = 0;
  c.last_total = 0;
  if (count == 0) {
    if (usage) *usage = u;
    return;
  }
  const uint64_t chunks = (count + kScanChunk - 1) / kScanChunk;
  bool ws = Reserve(c.d_items, count, ctx.device);
  ws = Reserve(c.d_counts, count, ctx.device) && ws;
  ws = Reserve(c.d_offsets, count + 1, ctx.device) && ws;
  ws = Reserve(c.d_chunk_sums, chunks, ctx.device) && ws;
  c.d_items.Upload(centers, count, ctx.stream);
  ctx.Sync();
  u.h2d_ms = MsSince(t0);

  // count pass and offset scan, then the exact reply count (the reference
  // sizes its reply buffer the same way, ht_replies.cu:626-637)
  t0 = std::chrono::steady_clock::now();
  const uint64_t blocks = (count + kHtWaves - 1) / kHtWaves;
  const auto count_kernel = c.near ? HtHamiltonKernel<false, true> : HtHamiltonKernel<false, false>;
  const auto write_kernel = c.near ? HtHamiltonKernel<true, true> : HtHamiltonKernel<true, false>;
  hipLaunchKernelGGL(count_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(kHtBlock), 0,
                     ctx.stream, c.d_xy.get(), c.d_row_offsets.get(), c.d_neighbors.get(),
                     static_cast<int>(g.distance_type), tu, tv, c.d_items.get(),
                     static_cast<uint64_t>(count), c.d_counts.get(), nullptr, nullptr);
  CheckHip(hipGetLastError(), "HtHamiltonKernel count launch");
  hipLaunchKernelGGL(HtChunkSumKernel, dim3(static_cast<uint32_t>(chunks)), dim3(kScanThreads), 0, ctx.stream,
                     c.d_counts.get(), static_cast<uint64_t>(count), c.d_chunk_sums.get());
  CheckHip(hipGetLastError(), "HtChunkSumKernel launch");
  hipLaunchKernelGGL(HtScanKernel, dim3(static_cast<uint32_t>(chunks)), dim3(kScanThreads), 0,
                     ctx.stream, c.d_counts.get(), static_cast<uint64_t>(count),
                     c.d_chunk_sums.get(), c.d_offsets.get());
  CheckHip(hipGetLastError(), "HtScanKernel launch");
  c.d_offsets.Download(offsets, count + 1, ctx.stream);
  ctx.Sync();
  *total = offsets[count];
  uint64_t bound = 0;  // every neighbour pair of every center
  for (size_t i = 0; i < count; ++i) {
    const uint64_t d = static_cast<uint64_t>(g.row_offsets[centers[i] + 1] - g.row_offsets[centers[i]]);
    bound += d < 2 ? 0 : d * (d - 1) / 2;
  }
  if (*total > bound) throw std::logic_error("HT Hamilton replies exceed the pair bound");
  if (*total > UINT64_MAX / 3) throw std::overflow_error("HT Hamilton reply count overflows");
  ws = Reserve(c.d_replies, 3 * *total, ctx.device) && ws;
  u.workspace_hit = ws ? 1 : 0;
  hipLaunchKernelGGL(write_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(kHtBlock), 0,
                     ctx.stream, c.d_xy.get(), c.d_row_offsets.get(), c.d_neighbors.get(),
                     static_cast<int>(g.distance_type), tu, tv, c.d_items.get(),
                     static_cast<uint64_t>(count), nullptr, c.d_offsets.get(), c.d_replies.get());
  CheckHip(hipGetLastError(), "HtHamiltonKernel write launch");
  ctx.Sync();
  u.kernel_ms = MsSince(t0);
  c.last_kind = 1;
  c.last_total = *total;

  t0 = std::chrono::steady_clock::now();
  const uint64_t copy = std::min(*total, capacity);
  if (copy > 0) DownloadToCaller(ctx, replies, c.d_replies.get(), copy * sizeof(ee_ht_neighbor_pair));
  u.d2h_ms = MsSince(t0);
  u.resident_bytes = HtResidentBytes(c);
  if (usage) *usage = u;
}

// Replies [first, first + count) of the context's last reply call.
void CopyReplies(Context& ctx, uint64_t first, uint64_t count, void* out) {
  HtCache& c = ctx.ht;
  if (c.last_kind == 0) throw std::logic_error("HT copy replies: no reply call has completed");
  if (first > c.last_total || count > c.last_total - first) {
    throw std::invalid_argument("HT copy replies: range beyond the last call's replies");
  }
  if (count == 0) return;
  if (out == nullptr) throw std::invalid_argument("HT copy replies: NULL buffer");
  ctx.Activate();
  const size_t record = c.last_kind == 1 ? sizeof(ee_ht_neighbor_pair) : sizeof(ee_node_edge);
  DownloadToCaller(ctx, out, reinterpret_cast<const char*>(c.d_replies.get()) + first * record, count * record);
}

void EndReplies(Context& ctx, const ee_graph_view* graph, const ee_ht_end_task* tasks, size_t count,
                uint64_t* offsets, ee_node_edge* replies, uint64_t capacity, uint64_t* total,
                ee_ht_usage* usage) {
  ee_ht_usage u{};
  ctx.Activate();
  CheckGraphView(graph);
  const ee_graph_view& g = *graph;
  if (offsets == nullptr || total == nullptr || (count > 0 && tasks == nullptr) ||
      (capacity > 0 && replies == nullptr)) {
    throw std::invalid_argument("HT end replies: NULL argument");
  }
  auto t0 = std::chrono::steady_clock::now();
  u.graph_hit = PrepareHtGraph(ctx, g) ? 1 : 0;  // checks a new snapshot
  offsets[0] = 0;
  for (size_t i = 0; i < count; ++i) {
    const int32_t e = tasks[i].endpoint, in = tasks[i].internal_neighbor;
    if (e < 0 || e >= g.dimension || in < 0 || in >= g.dimension || e == in || !RowHas(g, e, in)) {
      throw std::invalid_argument("HT end replies: a task must name an active path-internal edge");
    }
    // the internal edge is in the row exactly once, so the rest are replies
    offsets[i + 1] = offsets[i] + static_cast<uint64_t>(g.row_offsets[e + 1] - g.row_offsets[e] - 1);
  }
  *total = offsets[count];
  HtCache& c = ctx.ht;
  c.last_kind = 0;
  c.last_total = 0;
  if (count == 0) {
    if (usage) *usage = u;
    return;
  }
  bool ws = Reserve(c.d_items, 2 * count, ctx.device);
  ws = Reserve(c.d_offsets, count + 1, ctx.device) && ws;
  ws = Reserve(c.d_replies, 2 * *total, ctx.device) && ws;
  u.workspace_hit = ws ? 1 : 0;
  static_assert(sizeof(ee_ht_end_task) == 8, "task layout");
  c.d_items.Upload(reinterpret_cast<const int32_t*>(tasks), 2 * count, ctx.stream);
  c.d_offsets.Upload(offsets, count + 1, ctx.stream);
  ctx.Sync();
  u.h2d_ms = MsSince(t0);
  t0 = std::chrono::steady_clock::now();
  const uint64_t blocks = (count + kHtBlock - 1) / kHtBlock;
  hipLaunchKernelGGL(HtEndKernel, dim3(static_cast<uint32_t>(blocks)), dim3(kHtBlock), 0, ctx.stream,
                     c.d_row_offsets.get(), c.d_neighbors.get(), c.d_items.get(),
                     static_cast<uint64_t>(count), c.d_offsets.get(), c.d_replies.get());
  CheckHip(hipGetLastError(), "HtEndKernel launch");
  ctx.Sync();
  u.kernel_ms = MsSince(t0);
  c.last_kind = 2;
  c.last_total = *total;
  t0 = std::chrono::steady_clock::now();
  const uint64_t copy = std::min(*total, capacity);
  if (copy > 0) DownloadToCaller(ctx, replies, c.d_replies.get(), copy * sizeof(ee_node_edge));
  u.d2h_ms = MsSince(t0);
  u.resident_bytes = HtResidentBytes(c);
  if (usage) *usage = u;
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
}  // namespace ee

extern "C" {

ee_status ee_ht_hamilton_replies(ee_ctx* ctx, const ee_graph_view* graph, int32_t target_u,
                                 int32_t target_v, const int32_t* centers, size_t center_count,
                                 uint64_t* offsets, ee_ht_neighbor_pair* replies, uint64_t capacity,
                                 uint64_t* total, ee_ht_usage* usage) {
  return ee::Guarded([&] {
    if (ctx == nullptr) throw std::invalid_argument("NULL context");
    ee::HamiltonReplies(*reinterpret_cast<ee::Context*>(ctx), graph, target_u, target_v, centers,
                        center_count, offsets, replies, capacity, total, usage);
  });
}

ee_status ee_ht_end_replies(ee_ctx* ctx, const ee_graph_view* graph, const ee_ht_end_task* tasks,
                            size_t task_count, uint64_t* offsets, ee_node_edge* replies,
                            uint64_t capacity, uint64_t* total, ee_ht_usage* usage) {
  return ee::Guarded([&] {
    if (ctx == nullptr) throw std::invalid_argument("NULL context");
    ee::EndReplies(*reinterpret_cast<ee::Context*>(ctx), graph, tasks, task_count, offsets, replies,
                   capacity, total, usage);
  });
}

ee_status ee_ht_copy_replies(ee_ctx* ctx, uint64_t first, uint64_t count, void* replies) {
  return ee::Guarded([&] {
    if (ctx == nullptr) throw std::invalid_argument("NULL context");
    ee::CopyReplies(*reinterpret_cast<ee::Context*>(ctx), first, count, replies);
  });
}

ee_status ee_ht_clear_cache(ee_ctx* ctx) {
  return ee::Guarded([&] {
    if (ctx == nullptr) throw std::invalid_argument("NULL context");
    ee::Context& c = *reinterpret_cast<ee::Context*>(ctx);
    c.Activate();
    c.ht = ee::HtCache{};
  });
}

}  // extern "C"
