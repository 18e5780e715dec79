// Exact 3/4/5-opt leaf cost matrix on gfx950.
//
// Reference kernel: KOptTemplateCostsKernel (src/cuda/kopt_cost.cu:349-409),
// one thread per (row, template) cell recomputing k exact distances per cell.
//
// Here a workgroup owns a tile of task rows of one segment (one instance, one
// k).  Three LDS phases per tile:
//   A. stage the 2k port node ids and their coordinates (one gather per port);
//   B. for every distinct port pair the templates use, compute once per row
//      its canonical node pair and its exact distance, or INT64_MAX when it is
//      a self loop or re-adds one of the row's deleted edges;
//   C. every thread walks cells of the tile in row-major order: a cell sums
//      the k pair costs of its template and is INT64_MAX if a pair is invalid
//      or two pairs join the same node pair.  Consecutive threads write
//      consecutive int64 cells, so the store stream is fully coalesced.
// The cell rules are the reference's (kopt_cost.cu:370-408); the sum of at most
// five exact distances of a distance-safe graph is below 2^35, so the
// reference's overflow branch cannot fire and is not evaluated per cell.
#include "ee_exact.h"
#include "ee_internal.h"

namespace ee {
namespace {

constexpr int64_t kInvalid = INT64_MAX;

struct PairCell {
  int64_t cost;
  uint64_t key;  // (u << 32) | v of the canonical added edge
};

// q = c / d for c < 2^24 via a float reciprocal plus one correction step.
__device__ __forceinline__ uint32_t DivSmall(uint32_t c, uint32_t d, float inv) {
  uint32_t q = static_cast<uint32_t>(static_cast<float>(c) * inv);
  if (q * d > c) --q;
  if ((q + 1) * d <= c) ++q;
  return q;
}

__device__ __forceinline__ uint64_t EdgeKey(int32_t a, int32_t b) {
  const uint32_t lo = static_cast<uint32_t>(a < b ? a : b);
  const uint32_t hi = static_cast<uint32_t>(a < b ? b : a);
  return (static_cast<uint64_t>(lo) << 32) | hi;
}

// LDS layout of one tile (all offsets 16-byte aligned).
struct TileLayout {
  uint32_t xy_off, pairs_off, nodes_off, plan_off, bytes;
};

__host__ __device__ inline uint32_t Align16(uint32_t v) { return (v + 15u) & ~15u; }

__host__ __device__ inline TileLayout Layout(uint32_t k, uint32_t rows, uint32_t pairs,
                                             uint32_t templates) {
  TileLayout l;
  const uint32_t ports = 2 * k;
  l.xy_off = 0;
  l.pairs_off = Align16(l.xy_off + rows * ports * 16u);
  l.nodes_off = Align16(l.pairs_off + rows * pairs * 16u);
  l.plan_off = Align16(l.nodes_off + rows * ports * 4u);
  // plan: pair_p[45], pair_q[45], then templates padded to 8 bytes each
  l.bytes = Align16(l.plan_off + 2u * kMaxPairs + templates * 8u);
  return l;
}

template <int K>
__device__ __forceinline__ void LoadPlan(const KoptPlan& plan, uint8_t* s_plan) {
  for (uint32_t i = threadIdx.x; i < kMaxPairs; i += blockDim.x) {
    s_plan[i] = plan.pair_p[i];
    s_plan[kMaxPairs + i] = plan.pair_q[i];
  }
  uint8_t* s_tpl = s_plan + 2 * kMaxPairs;
  for (uint32_t i = threadIdx.x; i < plan.templates * 8u; i += blockDim.x) {
    const uint32_t t = i >> 3, j = i & 7u;
    s_tpl[i] = j < static_cast<uint32_t>(K) ? plan.tpl[t * K + j] : 0;
  }
}

template <int K, bool kNonTemporal>
__device__ void KoptTile(const KoptLaunch& a, const KoptSegment& seg, uint64_t tile, char* smem,
                         uint32_t P, uint32_t T) {
  constexpr uint32_t kPorts = 2 * K;
  const TileLayout lay = Layout(K, seg.rows_per_tile, P, T);
  longlong2* s_xy = reinterpret_cast<longlong2*>(smem + lay.xy_off);
  PairCell* s_pairs = reinterpret_cast<PairCell*>(smem + lay.pairs_off);
  int32_t* s_nodes = reinterpret_cast<int32_t*>(smem + lay.nodes_off);
  const uint8_t* s_pp = reinterpret_cast<const uint8_t*>(smem + lay.plan_off);
  const uint8_t* s_pq = s_pp + kMaxPairs;
  const uint64_t* s_tpl = reinterpret_cast<const uint64_t*>(s_pp + 2 * kMaxPairs);

  const uint64_t row0 = (tile - seg.tile_begin) * seg.rows_per_tile;
  const uint64_t left = seg.task_count - row0;
  const uint32_t rows = left < seg.rows_per_tile ? static_cast<uint32_t>(left) : seg.rows_per_tile;
  const KoptInstance inst = a.instances[seg.instance];
  const longlong2* xy = a.xy + inst.coord_base;
  const int32_t* ports =
      reinterpret_cast<const int32_t*>(a.tasks + seg.task_offset + row0);  // 12 ints per task

  // A. port nodes and coordinates
  for (uint32_t s = threadIdx.x; s < rows * kPorts; s += blockDim.x) {
    const uint32_t r = s / kPorts, p = s - r * kPorts;
    const int32_t node = ports[r * 12 + p];
    s_nodes[s] = node;
    s_xy[s] = xy[node];
  }
  __syncthreads();

  // B. distinct port-pair costs of every row
  const float invP = 1.0f / static_cast<float>(P);
  const int type = static_cast<int>(inst.distance_type);
  for (uint32_t s = threadIdx.x; s < rows * P; s += blockDim.x) {
    const uint32_t r = DivSmall(s, P, invP), pi = s - r * P;
    const uint32_t p = s_pp[pi], q = s_pq[pi];
    const int32_t* nodes = s_nodes + r * kPorts;
    const int32_t na = nodes[p], nb = nodes[q];
    const uint64_t key = EdgeKey(na, nb);
    bool bad = na == nb;
#pragma unroll
    for (int e = 0; e < K; ++e) bad |= key == EdgeKey(nodes[2 * e], nodes[2 * e + 1]);
    int64_t cost = kInvalid;
    if (!bad) {
      const longlong2 pa = s_xy[r * kPorts + p], pb = s_xy[r * kPorts + q];
      cost = ExactDistance(pa.x, pa.y, pb.x, pb.y, type);
    }
    s_pairs[s] = PairCell{cost, key};
  }
  __syncthreads();

  // C. cells, row-major, coalesced stores
  int64_t* out = a.costs + seg.cost_offset + row0 * T;
  const float invT = 1.0f / static_cast<float>(T);
  for (uint32_t c = threadIdx.x; c < rows * T; c += blockDim.x) {
    const uint32_t r = DivSmall(c, T, invT), t = c - r * T;
    const uint64_t tp = s_tpl[t];
    const PairCell* row = s_pairs + r * P;
    uint64_t keys[K];
    uint64_t total = 0;
    bool bad = false;
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const PairCell pc = row[(tp >> (8 * j)) & 0xffu];
      keys[j] = pc.key;
      bad |= pc.cost == kInvalid;
      total += static_cast<uint64_t>(pc.cost);
    }
#pragma unroll
    for (int i = 0; i < K; ++i) {
#pragma unroll
      for (int j = i + 1; j < K; ++j) bad |= keys[i] == keys[j];
    }
    const int64_t value = bad ? kInvalid : static_cast<int64_t>(total);
    if (kNonTemporal) {
      __builtin_nontemporal_store(value, out + c);
    } else {
      out[c] = value;
    }
  }
  __syncthreads();
}

template <bool kNonTemporal>
__global__ void __launch_bounds__(256) KoptKernel(KoptLaunch a) {
  extern __shared__ __attribute__((aligned(16))) char smem[];
  uint32_t loaded_k = 0;
  for (uint64_t tile = blockIdx.x; tile < a.tile_count; tile += gridDim.x) {
    const KoptSegment seg = a.segments[a.tile_segment ? a.tile_segment[tile] : 0u];
    const KoptPlan& plan = a.plans[seg.k - 3];
    const uint32_t P = plan.pairs, T = plan.templates;
    if (seg.k != loaded_k) {
      uint8_t* s_plan = reinterpret_cast<uint8_t*>(
          smem + Layout(seg.k, seg.rows_per_tile, P, T).plan_off);
      // every tile of a segment shares one layout; a new k starts a new plan
      switch (seg.k) {
        case 3: LoadPlan<3>(plan, s_plan); break;
        case 4: LoadPlan<4>(plan, s_plan); break;
        default: LoadPlan<5>(plan, s_plan); break;
      }
      loaded_k = seg.k;
      __syncthreads();
    }
    switch (seg.k) {
      case 3: KoptTile<3, kNonTemporal>(a, seg, tile, smem, P, T); break;
      case 4: KoptTile<4, kNonTemporal>(a, seg, tile, smem, P, T); break;
      default: KoptTile<5, kNonTemporal>(a, seg, tile, smem, P, T); break;
    }
  }
}

}  // namespace

size_t KoptSharedBytes(uint32_t k, uint32_t rows_per_tile, uint32_t pairs, uint32_t templates) {
  return Layout(k, rows_per_tile, pairs, templates).bytes;
}

void LaunchKopt(const KoptLaunch& args, const KoptTuning& tuning, size_t shared_bytes,
                hipStream_t stream) {
  if (args.tile_count == 0) return;
  if (shared_bytes > 65536) throw std::logic_error("k-opt tile needs more than 64 KiB of LDS");
  if (tuning.block == 0 || tuning.block > 256 || tuning.block % 64 != 0) {
    throw std::logic_error("k-opt block size must be a multiple of 64 up to 256");
  }
  uint64_t grid = tuning.grid == 0 ? args.tile_count : tuning.grid;
  if (grid > args.tile_count) grid = args.tile_count;
  if (grid > 0x7fffffffull) grid = 0x7fffffffull;
  const dim3 g(static_cast<uint32_t>(grid)), b(tuning.block);
  if (tuning.non_temporal) {
    hipLaunchKernelGGL(KoptKernel<true>, g, b, shared_bytes, stream, args);
  } else {
    hipLaunchKernelGGL(KoptKernel<false>, g, b, shared_bytes, stream, args);
  }
  CheckHip(hipGetLastError(), "k-opt kernel launch");
}

}  // namespace ee
