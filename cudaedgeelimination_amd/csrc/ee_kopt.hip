// Exact 3/4/5-opt leaf cost matrix on gfx950.
//
// Reference kernel: KOptTemplateCostsKernel (src/cuda/kopt_cost.cu:349-409),
// one thread per (row, template) cell recomputing k exact distances per cell.
//
// Here every 64-lane wavefront owns a tile of task rows of one segment (one
// instance, one k) and a private slice of LDS; the reconnect plans of the
// launch are loaded into LDS once per workgroup.  Per tile, with wave-level
// synchronisation only:
//   A. stage the 2k port node ids and coordinates (one gather per port);
//   A2. give every port the index of the first port holding the same node
//       (its class) and key every deleted edge by its class pair;
//   B. for every distinct port pair the templates use, compute once per row
//      the class-pair key of the added edge and its exact distance; a pair
//      that is a self loop or re-adds a deleted edge gets the sentinel 2^40;
//   C. lanes walk the tile's cells in row-major order: a cell sums the k pair
//      records of its template; a sentinel in the sum or two equal keys
//      (the same added edge twice) make it INT64_MAX.  Consecutive lanes write
//      consecutive int64 cells.
// The cell rules are the reference's (kopt_cost.cu:370-408).  Exact distances
// of a distance-safe graph are at most 2^32, so a pair record (40-bit cost or
// sentinel, 8-bit key) fits in 64 bits, the sum of five costs stays below
// 2^35 < 2^40, and the reference's overflow branch cannot fire.
#include <algorithm>

#include "ee_exact.h"
#include "ee_internal.h"

namespace ee {
namespace {

constexpr int kWave = 64;
constexpr int64_t kInvalid = INT64_MAX;
constexpr uint64_t kSentinel = uint64_t{1} << 40;
constexpr uint64_t kCostMask = (uint64_t{1} << 48) - 1;

__device__ __forceinline__ void WaveSync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// q = c / d for c < 2^24 via a float reciprocal plus one correction step.
__device__ __forceinline__ uint32_t DivSmall(uint32_t c, uint32_t d) {
  uint32_t q = static_cast<uint32_t>(static_cast<float>(c) * (1.0f / static_cast<float>(d)));
  if (q * d > c) --q;
  if ((q + 1) * d <= c) ++q;
  return q;
}

// Loads the plans of every k present into the block-shared LDS prefix.
__device__ void LoadPlans(const KoptLaunch& a, char* smem) {
  for (int i = 0; i < 3; ++i) {
    const KoptKLayout& L = a.layout[i];
    if (!L.present) continue;
    const KoptPlan& plan = a.plans[i];
    uint8_t* pp = reinterpret_cast<uint8_t*>(smem + L.pp_off);
    uint8_t* pq = reinterpret_cast<uint8_t*>(smem + L.pq_off);
    uint64_t* tpl = reinterpret_cast<uint64_t*>(smem + L.tpl_off);
    for (uint32_t j = threadIdx.x; j < L.P; j += blockDim.x) {
      pp[j] = plan.pair_p[j];
      pq[j] = plan.pair_q[j];
    }
    const uint32_t k = i + 3;
    for (uint32_t t = threadIdx.x; t < L.T; t += blockDim.x) {
      uint64_t packed = 0;
      for (uint32_t j = 0; j < k; ++j) packed |= static_cast<uint64_t>(plan.tpl[t * k + j]) << (8 * j);
      tpl[t] = packed;
    }
  }
}

template <int K, bool kNonTemporal>
__device__ void WaveTile(const KoptLaunch& a, const KoptKLayout& L, const KoptSegment& seg,
                         uint64_t tile, const char* smem, char* wave, int lane) {
  constexpr uint32_t kPorts = 2 * K;
  const uint8_t* s_pp = reinterpret_cast<const uint8_t*>(smem + L.pp_off);
  const uint8_t* s_pq = reinterpret_cast<const uint8_t*>(smem + L.pq_off);
  const uint64_t* s_tpl = reinterpret_cast<const uint64_t*>(smem + L.tpl_off);
  longlong2* w_xy = reinterpret_cast<longlong2*>(wave + L.xy_off);
  int32_t* w_node = reinterpret_cast<int32_t*>(wave + L.node_off);
  uint32_t* w_cls = reinterpret_cast<uint32_t*>(wave + L.cls_off);
  uint32_t* w_del = reinterpret_cast<uint32_t*>(wave + L.del_off);
  uint64_t* w_pair = reinterpret_cast<uint64_t*>(wave + L.pair_off);
  const uint32_t P = L.P, T = L.T;

  const uint64_t row0 = (tile - seg.tile_begin) * seg.rows_per_tile;
  const uint64_t left = seg.task_count - row0;
  const uint32_t rows = left < seg.rows_per_tile ? static_cast<uint32_t>(left) : seg.rows_per_tile;
  const KoptInstance inst = a.instances[seg.instance];
  const longlong2* xy = a.xy + inst.coord_base;
  const int32_t* ports = reinterpret_cast<const int32_t*>(a.tasks + seg.task_offset + row0);

  // A. port nodes and coordinates (a task row is 12 int32 words)
  for (uint32_t s = lane; s < rows * kPorts; s += kWave) {
    const uint32_t r = s / kPorts, p = s - r * kPorts;
    const int32_t node = ports[r * 12 + p];
    w_node[s] = node;
    w_xy[s] = xy[node];
  }
  WaveSync();

  // A2. port classes and deleted-edge keys
  for (uint32_t s = lane; s < rows * kPorts; s += kWave) {
    const uint32_t r = s / kPorts, p = s - r * kPorts;
    const int32_t* nd = w_node + r * kPorts;
    uint32_t c = p;
    for (uint32_t j = 0; j < p; ++j) {
      if (nd[j] == nd[p]) {
        c = j;
        break;
      }
    }
    w_cls[s] = c;
    if (p & 1u) {
      uint32_t c0 = p - 1;
      for (uint32_t j = 0; j + 1 < p; ++j) {
        if (nd[j] == nd[p - 1]) {
          c0 = j;
          break;
        }
      }
      w_del[r * K + (p >> 1)] = c0 < c ? c0 * 16 + c : c * 16 + c0;
    }
  }
  WaveSync();

  // B. pair records of every row
  const int type = static_cast<int>(inst.distance_type);
  for (uint32_t s = lane; s < rows * P; s += kWave) {
    const uint32_t r = DivSmall(s, P), pi = s - r * P;
    const uint32_t p = s_pp[pi], q = s_pq[pi];
    const uint32_t cp = w_cls[r * kPorts + p], cq = w_cls[r * kPorts + q];
    const uint32_t key = cp < cq ? cp * 16 + cq : cq * 16 + cp;
    bool bad = cp == cq;
#pragma unroll
    for (int e = 0; e < K; ++e) bad |= key == w_del[r * K + e];
    uint64_t rec = kSentinel;
    if (!bad) {
      const longlong2 pa = w_xy[r * kPorts + p], pb = w_xy[r * kPorts + q];
      rec = static_cast<uint64_t>(ExactDistance(pa.x, pa.y, pb.x, pb.y, type));
    }
    w_pair[s] = rec | (static_cast<uint64_t>(key) << 48);
  }
  WaveSync();

  // C. cells, row-major, coalesced stores
  int64_t* out = a.costs + seg.cost_offset + row0 * T;
  const uint32_t cells = rows * T;
  uint32_t r = DivSmall(static_cast<uint32_t>(lane), T);
  uint32_t t = static_cast<uint32_t>(lane) - r * T;
  for (uint32_t c = lane; c < cells; c += kWave) {
    const uint64_t tp = s_tpl[t];
    const uint64_t* row = w_pair + r * P;
    uint64_t sum = 0;
    uint32_t key[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const uint64_t rec = row[(tp >> (8 * j)) & 0xffu];
      sum += rec & kCostMask;
      key[j] = static_cast<uint32_t>(rec >> 48);
    }
    bool bad = sum >= kSentinel;
#pragma unroll
    for (int i = 0; i < K; ++i) {
#pragma unroll
      for (int j = i + 1; j < K; ++j) bad |= key[i] == key[j];
    }
    const int64_t value = bad ? kInvalid : static_cast<int64_t>(sum);
    if (kNonTemporal) {
      __builtin_nontemporal_store(value, out + c);
    } else {
      out[c] = value;
    }
    t += L.step_t;
    r += L.step_r;
    if (t >= T) {
      t -= T;
      ++r;
    }
  }
  WaveSync();
}

template <bool kNonTemporal>
__global__ void __launch_bounds__(256) KoptKernel(KoptLaunch a) {
  extern __shared__ __attribute__((aligned(16))) char smem[];
  LoadPlans(a, smem);
  __syncthreads();
  const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
  const uint32_t wib = threadIdx.x / kWave;
  const uint32_t waves_per_block = blockDim.x / kWave;
  char* wave = smem + a.wave_base + wib * a.wave_bytes;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * waves_per_block;
  for (uint64_t tile = __builtin_amdgcn_readfirstlane(blockIdx.x * waves_per_block + wib);
       tile < a.tile_count; tile += stride) {
    const KoptSegment seg = a.segments[a.tile_segment ? a.tile_segment[tile] : 0u];
    switch (seg.k) {
      case 3: WaveTile<3, kNonTemporal>(a, a.layout[0], seg, tile, smem, wave, lane); break;
      case 4: WaveTile<4, kNonTemporal>(a, a.layout[1], seg, tile, smem, wave, lane); break;
      default: WaveTile<5, kNonTemporal>(a, a.layout[2], seg, tile, smem, wave, lane); break;
    }
  }
}

uint32_t Align16(uint32_t v) { return (v + 15u) & ~15u; }

}  // namespace

size_t KoptWaveBytes(uint32_t k, uint32_t rows, uint32_t pairs) {
  const uint32_t ports = 2 * k;
  return Align16(rows * ports * 16) + Align16(rows * ports * 4) + Align16(rows * ports * 4) +
         Align16(rows * k * 4) + Align16(rows * pairs * 8);
}

size_t KoptPrepareLayout(KoptLaunch* a, const KoptPlan plans[3], const bool present[3],
                         const KoptTuning& tuning) {
  if (tuning.block == 0 || tuning.block > 256 || tuning.block % kWave != 0) {
    throw std::invalid_argument("k-opt block size must be a multiple of 64 up to 256");
  }
  uint32_t off = 0;
  uint32_t wave_bytes = 0;
  for (int i = 0; i < 3; ++i) {
    KoptKLayout& L = a->layout[i];
    L = KoptKLayout{};
    if (!present[i]) continue;
    const KoptPlan& plan = plans[i];
    const uint32_t k = i + 3, rows = tuning.rows_per_tile[i];
    if (rows == 0) throw std::invalid_argument("rows per tile must be positive");
    if (static_cast<uint64_t>(rows) * std::max(plan.pairs, plan.templates) >= (1u << 24)) {
      throw std::invalid_argument("k-opt tile too large");
    }
    L.present = 1;
    L.P = plan.pairs;
    L.T = plan.templates;
    L.rows = rows;
    L.step_r = kWave / L.T;
    L.step_t = kWave - L.step_r * L.T;
    L.pp_off = off;
    L.pq_off = off + kMaxPairs;
    L.tpl_off = Align16(off + 2 * kMaxPairs);
    off = Align16(L.tpl_off + L.T * 8);
    const uint32_t ports = 2 * k;
    L.xy_off = 0;
    L.node_off = Align16(rows * ports * 16);
    L.cls_off = L.node_off + Align16(rows * ports * 4);
    L.del_off = L.cls_off + Align16(rows * ports * 4);
    L.pair_off = L.del_off + Align16(rows * k * 4);
    wave_bytes = std::max<uint32_t>(wave_bytes, static_cast<uint32_t>(KoptWaveBytes(k, rows, L.P)));
  }
  a->wave_base = off;
  a->wave_bytes = wave_bytes;
  return static_cast<size_t>(off) + static_cast<size_t>(wave_bytes) * (tuning.block / kWave);
}

void LaunchKopt(const KoptLaunch& args, const KoptTuning& tuning, size_t shared_bytes,
                hipStream_t stream) {
  if (args.tile_count == 0) return;
  if (shared_bytes > 65536) throw std::invalid_argument("k-opt launch needs more than 64 KiB of LDS");
  const uint32_t waves_per_block = tuning.block / kWave;
  const uint64_t needed = (args.tile_count + waves_per_block - 1) / waves_per_block;
  uint64_t grid = tuning.grid == 0 ? needed : tuning.grid;
  if (grid > needed) grid = needed;
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
