// Exact 3/4/5-opt leaf cost matrix on gfx950.
//
// Reference kernel: KOptTemplateCostsKernel (src/cuda/kopt_cost.cu:349-409),
// one thread per (row, template) cell recomputing k exact distances per cell.
//
// Here every 64-lane wavefront owns a tile of task rows of one segment (one
// instance, one k) and a private slice of LDS; each workgroup runs the tiles
// of one k and loads that k's reconnect plan into LDS once.  Per tile, with
// wave-level synchronisation only:
//   A. stage the 2k port node ids and coordinates (one gather per port);
//   A2. give every port the bit mask of the ports holding the same node, and
//       mark per row the first port of every node held by two or more ports;
//   B. for every distinct port pair the templates use, build once per row a
//      64-bit pair record (below), stored pair-major so that one pair's
//      records of consecutive rows are adjacent;
//   C. lanes keep one template each and walk a run of rows: a cell is the sum
//      of its k pair records plus one mask test, written as int64.
//
// Pair record: bits 0..34 hold the exact distance of the added edge, bits
// 61..63 count sentinels (a self loop, or an added edge equal to a deleted
// one), and bits 35..54 hold ten 2-bit counters, one per unordered pair of
// nodes that each hold two or more ports of the row.  A perfect matching adds
// the same edge twice only between two such nodes, so a cell is invalid
// exactly when its summed record has a sentinel or a counter >= 2; these are
// the reference's cell rules (kopt_cost.cu:370-408).  Field widths:
//  - exact distances of a distance-safe graph are at most 2^32, so five of
//    them stay below 2^35 (the reference's overflow branch cannot fire);
//  - tasks are validated to have k distinct deleted edges without self loops,
//    so a node holding c ports has c distinct neighbours among the other
//    ports; two nodes holding four ports each would need at least 14 ports,
//    hence a counter never exceeds 3 for k <= 5, and at most five nodes hold
//    two or more of the 2k <= 10 ports (ten counters).
#include <algorithm>
#include <type_traits>

#include "ee_exact.h"
#include "ee_internal.h"
#include "ee_jv_body.h"

namespace ee {
namespace {

constexpr int kWave = 64;
constexpr int64_t kInvalid = INT64_MAX;
constexpr int kCounterBase = 35;
constexpr uint64_t kCostMask = (uint64_t{1} << kCounterBase) - 1;
constexpr uint64_t kSentinel = uint64_t{1} << 61;
constexpr uint64_t kBadMask = (uint64_t{0xAAAAA} << kCounterBase) | (uint64_t{7} << 61);
// Bytes of one port's record in the port area: coordinates (the low words on
// near tiles), then the two info words of phase A2.
constexpr uint32_t kNearPortBytes = 16;
constexpr uint32_t kFarPortBytes = 32;

// The launch's buffers, taken from the kernel arguments.
struct KoptPointers {
  const longlong2* xy;
  const KoptTile* tiles;
  const ee_kopt_task* tasks;
  const KoptPlan* plans;
  int64_t* costs;
  uint32_t* masks;
  // measurement only (null otherwise): wavefront w of the launch writes the
  // constant-rate device clock at its start to stamps[2w] and at its end to
  // stamps[2w + 1]
  unsigned long long* stamps;
};

__device__ __forceinline__ void WaveSync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A value every lane holds alike, moved to scalar registers.
__device__ __forceinline__ uint64_t Uniform64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v >> 32));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

// q = c / d for c < 2^24 via a float reciprocal plus one correction step.
__device__ __forceinline__ uint32_t DivSmall(uint32_t c, uint32_t d) {
  uint32_t q = static_cast<uint32_t>(static_cast<float>(c) * (1.0f / static_cast<float>(d)));
  if (q * d > c) --q;
  if ((q + 1) * d <= c) ++q;
  return q;
}

// Writes one cell at byte offset `at` of the tile's output.
template <bool kNonTemporal>
__device__ __forceinline__ void StoreCell(int64_t* out, uint32_t at, uint64_t sum) {
  const uint32_t hi = static_cast<uint32_t>(sum >> 32);
  const bool bad = (hi & static_cast<uint32_t>(kBadMask >> 32)) != 0;  // kBadMask has no low bits
  const int64_t value = bad ? kInvalid : static_cast<int64_t>(sum & kCostMask);
  int64_t* dst = reinterpret_cast<int64_t*>(reinterpret_cast<char*>(out) + at);
  if (kNonTemporal) {
    __builtin_nontemporal_store(value, dst);
  } else {
    *dst = value;
  }
}

// Phase A: the tile's port node ids (into the pair record area, which is
// free until phase B) and their coordinates, two gathers in flight per lane.
// kBounded: node ids at or above node_count read node 0's coordinates (task
// rows that another agent wrote, whose checks this kernel cannot see).
template <int K, bool kNear, bool kMask, bool kBounded = false>
__device__ __forceinline__ void StagePorts(const KoptLaunch& a, const KoptPointers& bufs, const KoptKLayout& L,
                                           const KoptTile& d, char* wave, int lane, uint32_t node_count = 0) {
  constexpr uint32_t kPorts = 2 * K;
  constexpr uint32_t kBatch = 2;
  char* w_port = wave + L.xy_off;
  uint32_t* w_row = reinterpret_cast<uint32_t*>(wave + L.row_off);
  int32_t* w_node = reinterpret_cast<int32_t*>(wave + L.pair_off);
  const uint32_t total = d.rows * kPorts;
  const longlong2* xy = bufs.xy + d.coord_base;
  const int32_t* ports = reinterpret_cast<const int32_t*>(bufs.tasks + d.task_index);
  for (uint32_t base = 0; base < total; base += kBatch * kWave) {
    // lanes past the end re-read port 0 of the tile, so every load is in range
    int32_t node[kBatch];
#pragma unroll
    for (uint32_t i = 0; i < kBatch; ++i) {
      const uint32_t s = base + i * kWave + static_cast<uint32_t>(lane);
      const uint32_t r = s / kPorts, p = s - r * kPorts;
      node[i] = ports[s < total ? r * 12 + p : 0u];  // a task row is 12 int32 words
      if (kBounded && static_cast<uint32_t>(node[i]) >= node_count) node[i] = 0;
    }
    int64_t cx[kBatch], cy[kBatch];
#pragma unroll
    for (uint32_t i = 0; i < kBatch; ++i) {
      const longlong2 c = xy[node[i]];
      cx[i] = c.x;
      cy[i] = c.y;
    }
#pragma unroll
    for (uint32_t i = 0; i < kBatch; ++i) {
      const uint32_t s = base + i * kWave + static_cast<uint32_t>(lane);
      if (s < total) {
        const uint32_t r = s / kPorts, p = s - r * kPorts;
        w_node[s] = node[i];
        if (kNear) {
          // coordinate extents below 2^25: differences of the low words are exact
          *reinterpret_cast<uint2*>(w_port + s * kNearPortBytes) =
              uint2{static_cast<uint32_t>(cx[i]), static_cast<uint32_t>(cy[i])};
        } else {
          *reinterpret_cast<longlong2*>(w_port + s * kFarPortBytes) = longlong2{cx[i], cy[i]};
        }
        if (p == 0) {
          w_row[r] = 0;
          // mask output: the row's deleted_cost (bytes 40..47 of the task row)
          if (kMask) {
            reinterpret_cast<int64_t*>(wave + L.dcost_off)[r] =
                reinterpret_cast<const int64_t*>(ports + r * 12)[5];
          }
        }
      }
    }
  }
  WaveSync();
}

// Duplicate-edge counters: the high word of a pair record holds, for every
// two nodes of ranks lo < hi < 5 (the row's nodes held by two or more ports,
// ranked by first port), a 2-bit counter at bits lo * (7 - lo) + 2 * hi + 1
// and one up, so a template that adds the same node pair twice carries into
// the upper bit.  RankMaskOf(a) has the low counter bit of every pair of
// nodes that includes rank a: for ranks a != b, RankMaskOf(a) &
// RankMaskOf(b) is exactly the counter bit of pair (a, b).
constexpr uint32_t CounterBit(uint32_t lo, uint32_t hi) { return 1u << (lo * (7u - lo) + 2u * hi + 1u); }
constexpr uint32_t RankMaskOf(uint32_t a) {
  uint32_t m = 0;
  for (uint32_t x = 0; x < 5; ++x) {
    if (x != a) m |= x < a ? CounterBit(x, a) : CounterBit(a, x);
  }
  return m;
}
static_assert((RankMaskOf(1) & RankMaskOf(3)) == CounterBit(1, 3), "one common bit");
static_assert((RankMaskOf(0) & RankMaskOf(4)) == CounterBit(0, 4), "one common bit");

__device__ __forceinline__ uint32_t RankMask(uint32_t rank) {
  return rank == 0 ? RankMaskOf(0)
       : rank == 1 ? RankMaskOf(1)
       : rank == 2 ? RankMaskOf(2)
       : rank == 3 ? RankMaskOf(3)
                   : RankMaskOf(4);
}

// Phase B: pair records, pair-major (record of pair pi and row r at
// pi * stride + r), from the port records of p and q:
//  - sentinel: q's node is p's node, or the two nodes are the end nodes of
//    one deleted edge (info word 0: the same-node mask of p in bits 0..9, the
//    mask together with the deleted-edge partners of its ports in bits
//    16..25);
//  - duplicate-edge counter bit: info word 1 of p AND that of q (RankMask of
//    the node's rank, 0 for a node held by one port).
// Every lane computes the distance, sentinel or not, so the loop body does
// not diverge; the next iteration's pair ports are read at the top of the
// current one.  kType is the distance type (0 EUC_2D, 1 CEIL_2D).
template <int K, bool kNear, int kType>
__device__ __forceinline__ void PairRecords(const KoptKLayout& L, uint32_t rows, const uint16_t* s_pq,
                                            const char* wave, uint64_t* w_pair, int lane) {
  constexpr uint32_t kPorts = 2 * K;
  constexpr uint32_t kPortBytes = kNear ? kNearPortBytes : kFarPortBytes;
  const char* w_port = wave + L.xy_off;
  const uint32_t P = L.P, stride = L.pair_stride;
  uint32_t r = DivSmall(static_cast<uint32_t>(lane), P);
  uint32_t pi = static_cast<uint32_t>(lane) - r * P;
  uint32_t pq = s_pq[pi];
  while (r < rows) {
    const uint32_t p = pq & 0xffu, q = pq >> 8;
    const uint32_t row = r, slot = pi;
    r += L.pair_step_r;
    pi += L.pair_step_p;
    if (pi >= P) {
      pi -= P;
      ++r;
    }
    pq = s_pq[pi];  // pi < P always, so this read is in range even past the last row
    const uint32_t base = __umul24(row, kPorts);
    const char* ra = w_port + (base + p) * kPortBytes;
    const char* rb = w_port + (base + q) * kPortBytes;
    uint32_t lo_word, hi_word;
    uint2 ia, ib;
    if (kNear) {
      const uint4 pa = *reinterpret_cast<const uint4*>(ra);
      const uint4 pb = *reinterpret_cast<const uint4*>(rb);
      lo_word = static_cast<uint32_t>(ExactDistanceNear(static_cast<int32_t>(pa.x - pb.x),
                                                        static_cast<int32_t>(pa.y - pb.y), kType));
      hi_word = 0;  // near distances are below 2^27
      ia = uint2{pa.z, pa.w};
      ib = uint2{pb.z, pb.w};
    } else {
      const longlong2 pa = *reinterpret_cast<const longlong2*>(ra);
      const longlong2 pb = *reinterpret_cast<const longlong2*>(rb);
      const uint64_t dist = static_cast<uint64_t>(ExactDistance(pa.x, pa.y, pb.x, pb.y, kType));
      lo_word = static_cast<uint32_t>(dist);
      hi_word = static_cast<uint32_t>(dist >> 32);
      ia = *reinterpret_cast<const uint2*>(ra + 16);
      ib = *reinterpret_cast<const uint2*>(rb + 16);
    }
    const bool sentinel = ((ia.x >> 16) & ib.x) != 0;
    const uint64_t rec = (static_cast<uint64_t>(hi_word | (ia.y & ib.y)) << 32) | lo_word;
    w_pair[__umul24(slot, stride) + row] = sentinel ? kSentinel : rec;
  }
}

// The cells of one lane's template: n steps of G rows (rows g, g + G, ...),
// plus row g + n * G when `extra_row`.  `off` holds the byte offsets of the
// template's k records for row g; `at` the byte offset of its first cell.
// Near tiles sum the two 32-bit halves of the records separately: their
// distances are below 2^26, so the low halves sum to the cost without a carry
// and the high halves to the counters and sentinels.  kG is G when it is a
// compile-time constant, else 0.
//
// Mask mode (m.words > 0): no cells are stored; a lane whose cell is below
// the row's deleted cost sets its template's bit in the row's mask words,
// staged in LDS (zeroed first) and copied to the output once per tile with
// full-width stores.  Improving cells are rare, so most steps write nothing.
struct MaskOut {
  uint32_t* stage;       // LDS: mask of tile row 0
  uint32_t words;        // words per row; 0 = cost output
  const int64_t* dcost;  // LDS: deleted cost of each tile row
  uint32_t t;            // the lane's template
  uint32_t g;            // lane group (its rows are g, g + G, ...)
};

__device__ __forceinline__ void WriteMaskRow(const MaskOut& m, uint32_t row, bool improving) {
  if (improving) atomicOr(m.stage + row * m.words + m.t / 32, 1u << (m.t % 32));
}

template <int K, bool kNear, bool kNonTemporal, uint32_t kG, bool kMask>
__device__ __forceinline__ void CellSteps(const char* pairs, const uint32_t (&off)[K], int64_t* out,
                                          uint32_t at, uint32_t step_bytes, uint32_t g_rows, uint32_t n,
                                          bool extra_row, const MaskOut& m) {
  const uint32_t G = kG != 0 ? kG : g_rows;
  auto steps = [&](auto unroll, uint32_t i) {
    constexpr uint32_t U = decltype(unroll)::value;
    if (kMask && kNear) {
      // near coordinates: as for the cells below, the cost is the low word
      uint32_t lo[U], hi[U];
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) lo[u] = hi[u] = 0;
#pragma unroll
      for (int j = 0; j < K; ++j) {
        const uint2* rec = reinterpret_cast<const uint2*>(pairs + off[j] + i * G * 8);
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
          const uint2 v = rec[u * G];
          lo[u] += v.x;
          hi[u] += v.y;
        }
      }
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) {
        const uint32_t row = m.g + (i + u) * G;
        const bool bad = (hi[u] & static_cast<uint32_t>(kBadMask >> 32)) != 0;
        WriteMaskRow(m, row, !bad && static_cast<int64_t>(lo[u]) < m.dcost[row]);
      }
    } else if (kMask) {
      uint64_t sum[U];
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) sum[u] = 0;
#pragma unroll
      for (int j = 0; j < K; ++j) {
        const uint64_t* rec = reinterpret_cast<const uint64_t*>(pairs + off[j] + i * G * 8);
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) sum[u] += rec[u * G];
      }
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) {
        const uint32_t row = m.g + (i + u) * G;
        const uint32_t hi = static_cast<uint32_t>(sum[u] >> 32);
        const bool bad = (hi & static_cast<uint32_t>(kBadMask >> 32)) != 0;
        const int64_t value = bad ? kInvalid : static_cast<int64_t>(sum[u] & kCostMask);
        WriteMaskRow(m, row, value < m.dcost[row]);
      }
    } else if (kNear) {
      uint32_t lo[U], hi[U];
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) lo[u] = hi[u] = 0;
#pragma unroll
      for (int j = 0; j < K; ++j) {
        const uint2* rec = reinterpret_cast<const uint2*>(pairs + off[j] + i * G * 8);
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
          const uint2 v = rec[u * G];
          lo[u] += v.x;
          hi[u] += v.y;
        }
      }
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) {
        const bool bad = (hi[u] & static_cast<uint32_t>(kBadMask >> 32)) != 0;
        const int64_t value = bad ? kInvalid : static_cast<int64_t>(lo[u]);
        int64_t* dst = reinterpret_cast<int64_t*>(reinterpret_cast<char*>(out) + at);
        if (kNonTemporal) {
          __builtin_nontemporal_store(value, dst);
        } else {
          *dst = value;
        }
        at += step_bytes;
      }
    } else {
      uint64_t sum[U];
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) sum[u] = 0;
#pragma unroll
      for (int j = 0; j < K; ++j) {
        const uint64_t* rec = reinterpret_cast<const uint64_t*>(pairs + off[j] + i * G * 8);
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) sum[u] += rec[u * G];
      }
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) {
        StoreCell<kNonTemporal>(out, at, sum[u]);
        at += step_bytes;
      }
    }
  };
  uint32_t i = 0;
  for (; i + 2 <= n; i += 2) steps(std::integral_constant<uint32_t, 2>{}, i);
  if (i < n) steps(std::integral_constant<uint32_t, 1>{}, i);
  if (extra_row) steps(std::integral_constant<uint32_t, 1>{}, n);  // row g + n * G < rows
}

// Phases A2, B and C of one staged tile.
template <int K, bool kNear, bool kNonTemporal, bool kMask>
__device__ void TileCells(const KoptLaunch& a, const KoptPointers& bufs, const KoptKLayout& L, const KoptTile& d,
                          const char* smem, char* wave, int lane) {
  constexpr uint32_t kPorts = 2 * K;
  const uint16_t* s_pq = reinterpret_cast<const uint16_t*>(smem + L.pq_off);
  const uint64_t* s_tpl = reinterpret_cast<const uint64_t*>(smem + L.tpl_off);
  uint32_t* w_row = reinterpret_cast<uint32_t*>(wave + L.row_off);
  uint64_t* w_pair = reinterpret_cast<uint64_t*>(wave + L.pair_off);
  const int32_t* w_node = reinterpret_cast<const int32_t*>(wave + L.pair_off);
  const uint32_t T = L.T, stride = L.pair_stride;
  const uint32_t rows = d.rows;
  // a port's info words in its port record; the first pass parks its
  // same-node mask in word 0, where only its own lane reads it back
  char* const w_port = wave + L.xy_off;
  constexpr uint32_t kPortBytes = kNear ? kNearPortBytes : kFarPortBytes;
  constexpr uint32_t kInfoOffset = kNear ? 8 : 16;

  // A2. same-node masks; per row, the first port of every node held by two
  // or more ports
  for (uint32_t s = lane; s < rows * kPorts; s += kWave) {
    const uint32_t r = s / kPorts, p = s - r * kPorts;
    const int32_t* nd = w_node + r * kPorts;
    const int32_t me = nd[p];
    uint32_t eq = 0;
#pragma unroll
    for (uint32_t j = 0; j < kPorts; ++j) eq |= static_cast<uint32_t>(nd[j] == me) << j;
    *reinterpret_cast<uint32_t*>(w_port + s * kPortBytes + kInfoOffset) = eq;
    if (eq != (1u << p) && __builtin_ctz(eq) == p) atomicOr(&w_row[r], 1u << p);
  }
  WaveSync();
  // port info words: word 0 the same-node mask eq in bits 0..9 and, in bits
  // 16..25, eq together with the deleted-edge partners of its ports (port
  // 2e <-> 2e + 1); word 1 RankMask of the rank of the port's node among
  // the row's nodes held by two or more ports (by first port), or 0 for a
  // node held by this port alone
  for (uint32_t s = lane; s < rows * kPorts; s += kWave) {
    const uint32_t r = s / kPorts;
    const uint32_t eq = *reinterpret_cast<const uint32_t*>(w_port + s * kPortBytes + kInfoOffset);
    const uint32_t below = w_row[r] & ((eq & (0u - eq)) - 1u);
    const uint32_t mask =
        (eq & (eq - 1u)) != 0 ? RankMask(static_cast<uint32_t>(__builtin_popcount(below))) : 0u;
    const uint32_t partners = eq | ((eq & 0x155u) << 1) | ((eq >> 1) & 0x155u);
    *reinterpret_cast<uint2*>(w_port + s * kPortBytes + kInfoOffset) = uint2{eq | (partners << 16), mask};
  }
  WaveSync();

  // B. pair records
  if (d.distance_type == 0) {
    PairRecords<K, kNear, 0>(L, rows, s_pq, wave, w_pair, lane);
  } else {
    PairRecords<K, kNear, 1>(L, rows, s_pq, wave, w_pair, lane);
  }
  WaveSync();

  // C. cells: each lane keeps template t.  With T <= 64 the wave holds G =
  // 64 / T rows at a time (lane group g takes rows g, g + G, g + 2G, ...), so
  // one store writes G consecutive rows, G * T contiguous cells; with T > 64
  // it walks every row once per pass of 64 templates.
  const uint64_t cost_index = Uniform64(d.cost_index);
  int64_t* const out = bufs.costs + cost_index;
  const char* const pairs = reinterpret_cast<const char*>(w_pair);
  const uint32_t G = L.group;
  const uint32_t n = __builtin_amdgcn_readfirstlane(rows / G);  // rows every group has
  const uint32_t extra = rows - n * G;                          // groups g < extra have one more
  const uint32_t step_bytes = G * T * 8;
  // mask mode: the tile's mask words are staged in the port records, which
  // phase C no longer reads (the layout sizes them for rows x mask words)
  uint32_t* const w_mask = reinterpret_cast<uint32_t*>(wave + L.xy_off);
  const uint32_t mask_count = kMask ? rows * a.mask_words : 0u;
  if (kMask) {
    for (uint32_t q = lane; q < mask_count; q += kWave) w_mask[q] = 0u;
    WaveSync();
  }
  for (uint32_t pass = 0; pass < L.passes; ++pass) {
    uint32_t g = 0, t = pass * kWave + static_cast<uint32_t>(lane);
    if (T <= static_cast<uint32_t>(kWave)) {
      g = DivSmall(static_cast<uint32_t>(lane), T);
      t = static_cast<uint32_t>(lane) - g * T;
    }
    const bool active = g < G && t < T;
    const uint64_t tp = s_tpl[active ? t : 0u];
    uint32_t off[K];  // byte offset of record (pair j of the template, row g)
#pragma unroll
    for (int j = 0; j < K; ++j) off[j] = (static_cast<uint32_t>((tp >> (8 * j)) & 0xffu) * stride + g) * 8;
    // byte offset of the lane's next cell; a tile's cells stay below 2^24
    const uint32_t at = (g * T + t) * 8;
    if (!active) continue;
    // the standard templates give T = 4 / 25 / 208, so G = 16 / 2 / 1: with G
    // a constant the rows of a step are immediate offsets of one address
    constexpr uint32_t kStdG = K == 3 ? 16u : (K == 4 ? 2u : 1u);
    if (kMask) {
      MaskOut m;
      m.stage = w_mask;
      m.words = a.mask_words;
      m.dcost = reinterpret_cast<const int64_t*>(wave + L.dcost_off);
      m.t = t;
      m.g = g;
      if (G == kStdG) {
        CellSteps<K, kNear, kNonTemporal, kStdG, true>(pairs, off, out, at, step_bytes, G, n, g < extra, m);
      } else {
        CellSteps<K, kNear, kNonTemporal, 0, true>(pairs, off, out, at, step_bytes, G, n, g < extra, m);
      }
    } else if (G == kStdG) {
      CellSteps<K, kNear, kNonTemporal, kStdG, false>(pairs, off, out, at, step_bytes, G, n, g < extra, MaskOut{});
    } else {
      CellSteps<K, kNear, kNonTemporal, 0, false>(pairs, off, out, at, step_bytes, G, n, g < extra, MaskOut{});
    }
  }
  WaveSync();
  if (kMask) {
    // the tile's rows are consecutive task rows: one contiguous block
    uint32_t* const dst = bufs.masks + Uniform64(d.task_index) * a.mask_words;
    for (uint32_t q = lane; q < mask_count; q += kWave) dst[q] = w_mask[q];
    WaveSync();
  }
}

// Descriptor of `tile` (a tile of this k): from the inline segment table
// when the launch has one, else from the descriptor array.
template <int K>
__device__ __forceinline__ KoptTile TileAt(const KoptLaunch& a, const KoptPointers& bufs, const KoptKLayout& L,
                                           uint64_t tile) {
  if (a.inline_count == 0) return bufs.tiles[tile];
  constexpr int ki = K - 3;
  uint32_t s = a.k_inline_begin[ki];
  const uint32_t e = a.k_inline_begin[ki + 1];
  while (s + 1 < e && tile >= a.inline_seg[s + 1].tile_begin) ++s;
  const KoptInlineSegment& g = a.inline_seg[s];
  const uint32_t row0 = static_cast<uint32_t>(tile - g.tile_begin) * L.rows;
  KoptTile d;
  d.task_index = g.task_offset + row0;
  d.cost_index = g.cost_offset + static_cast<uint64_t>(row0) * L.T;
  d.coord_base = g.coord_base;
  d.rows = g.task_count - row0 < L.rows ? g.task_count - row0 : L.rows;
  d.distance_type = g.distance_type;
  d.near = g.near;
  return d;
}

// The tiles of one k handed to this workgroup, one per wavefront at a time.
// The plan is fetched into registers first, so its latency overlaps the
// first tile's gathers; it reaches LDS before the only block barrier.
template <int K, bool kNonTemporal, bool kMask>
__device__ void TileLoop(const KoptLaunch& a, const KoptPointers& bufs, char* smem) {
  constexpr int ki = K - 3;
  // k = 5 tiles hold most of the cells and end the launch: their wavefronts
  // take issue priority over the k = 3 and k = 4 wavefronts sharing a SIMD,
  // which fill the gaps (the single-stream launch, retuned, 10.28 -> 9.75 us)
  if constexpr (K == 5) __builtin_amdgcn_s_setprio(3);
  constexpr uint32_t kPlanRegs = 4;
  const KoptKLayout& L = a.layout[ki];
  const KoptPlan& plan = bufs.plans[ki];
  const uint32_t tid = threadIdx.x, bd = blockDim.x;
  uint64_t tv[kPlanRegs];
#pragma unroll
  for (uint32_t i = 0; i < kPlanRegs; ++i) {
    const uint32_t t = tid + i * bd;
    tv[i] = t < L.T ? plan.tpl[t] : 0;
  }
  const uint16_t pqv = tid < L.P ? plan.pair_pq[tid] : 0;

  const int lane = static_cast<int>(tid) & (kWave - 1);
  const uint32_t wib = tid / kWave;
  const uint32_t waves_per_block = bd / kWave;
  char* wave = smem + a.wave_base + wib * a.wave_bytes;
  const uint32_t wave_id = __builtin_amdgcn_readfirstlane(blockIdx.x * waves_per_block + wib);
  if (bufs.stamps != nullptr && lane == 0) bufs.stamps[2 * static_cast<uint64_t>(wave_id)] = __builtin_amdgcn_s_memrealtime();
  const uint32_t block = blockIdx.x - a.k_block_first[ki];
  const uint32_t blocks = a.k_block_count[ki];
  const uint64_t stride = static_cast<uint64_t>(blocks) * waves_per_block;
  const uint64_t end = a.k_tile_begin[ki] + a.k_tile_count[ki];
  uint64_t tile = a.k_tile_begin[ki] + __builtin_amdgcn_readfirstlane(block * waves_per_block + wib);
  KoptTile d{};
  if (tile < end) {
    d = TileAt<K>(a, bufs, L, tile);
    if (d.near != 0) {
      StagePorts<K, true, kMask>(a, bufs, L, d, wave, lane);
    } else {
      StagePorts<K, false, kMask>(a, bufs, L, d, wave, lane);
    }
  }

  uint16_t* s_pq = reinterpret_cast<uint16_t*>(smem + L.pq_off);
  uint64_t* s_tpl = reinterpret_cast<uint64_t*>(smem + L.tpl_off);
#pragma unroll
  for (uint32_t i = 0; i < kPlanRegs; ++i) {
    const uint32_t t = tid + i * bd;
    if (t < L.T) s_tpl[t] = tv[i];
  }
  for (uint32_t t = tid + kPlanRegs * bd; t < L.T; t += bd) s_tpl[t] = plan.tpl[t];
  if (tid < L.P) s_pq[tid] = pqv;
  __syncthreads();

  for (bool first = true; tile < end; tile += stride, first = false) {
    if (!first) d = TileAt<K>(a, bufs, L, tile);
    if (d.near != 0) {
      if (!first) StagePorts<K, true, kMask>(a, bufs, L, d, wave, lane);
      TileCells<K, true, kNonTemporal, kMask>(a, bufs, L, d, smem, wave, lane);
    } else {
      if (!first) StagePorts<K, false, kMask>(a, bufs, L, d, wave, lane);
      TileCells<K, false, kNonTemporal, kMask>(a, bufs, L, d, smem, wave, lane);
    }
  }
  if (bufs.stamps != nullptr && lane == 0) bufs.stamps[2 * static_cast<uint64_t>(wave_id) + 1] = __builtin_amdgcn_s_memrealtime();
}

template <bool kNonTemporal, bool kMask>
__device__ __forceinline__ void KoptBody(const KoptLaunch& a, const KoptPointers& bufs, char* smem) {
  if (blockIdx.x < a.k_block_first[1]) {
    TileLoop<3, kNonTemporal, kMask>(a, bufs, smem);
  } else if (blockIdx.x < a.k_block_first[2]) {
    TileLoop<4, kNonTemporal, kMask>(a, bufs, smem);
  } else {
    TileLoop<5, kNonTemporal, kMask>(a, bufs, smem);
  }
}

// The launch arguments by value, in the kernel argument segment.
template <bool kNonTemporal, bool kMask>
__global__ void __launch_bounds__(256) KoptKernel(KoptLaunch a) {
  extern __shared__ __attribute__((aligned(16))) char smem[];
  const KoptPointers bufs{a.xy, a.tiles, a.tasks, a.plans, a.costs, a.masks, nullptr};
  KoptBody<kNonTemporal, kMask>(a, bufs, smem);
}

// The launch arguments in ordinary device memory, for callers that launch
// the same arguments many times.  The argument segment is rewritten by the
// host for every launch and read uncached, and every wavefront's first
// memory accesses depend on a chain of reads from it; a resident copy is
// served from the caches after the first wavefront.  The buffer pointers
// still come as arguments (one batch of independent reads), so the compiler
// knows them as global memory.
template <bool kNonTemporal, bool kMask>
__global__ void __launch_bounds__(256) KoptKernelResident(const KoptLaunch* __restrict__ ap, KoptPointers bufs) {
  extern __shared__ __attribute__((aligned(16))) char smem[];
  KoptBody<kNonTemporal, kMask>(*ap, bufs, smem);
}

// ------------------------------------------------------------ work queue
// Control words are read with relaxed atomic loads: those are never served
// from the scalar cache, and at agent or system scope they miss the
// non-coherent caches, so a poll sees the other side's latest store.
__device__ __forceinline__ uint64_t PollAgent(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t PollAgent(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t PollSystem(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ uint32_t PollSystem(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ void StoreAgent(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void StoreAgent(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Entry `entry` from the device mirror, every field uniform.
__device__ __forceinline__ KoptQueueEntry MirrorEntry(const KoptQueueArgs& q, uint64_t entry) {
  const uint64_t* src = reinterpret_cast<const uint64_t*>(q.mirror + entry % q.ring_size);
  uint64_t w[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) w[i] = Uniform64(PollAgent(src + i));
  KoptQueueEntry e;
  __builtin_memcpy(&e, w, sizeof(e));
  return e;
}

// The entry's fields stay inside the arenas and the coordinate pool, so a
// tile of it reads and writes in bounds whatever its task rows hold.
__device__ __forceinline__ bool EntryInBounds(const KoptLaunch& a, const KoptQueueArgs& q, const KoptQueueEntry& e) {
  if (e.k == 0) {
    // a JV recheck: its tiles' rows (marks, and the slots, 12 per row) and
    // its witnesses (two per cell) inside the arenas
    const uint64_t tiles = (static_cast<uint64_t>(e.task_count) + kQueueJvTileEdges - 1) / kQueueJvTileEdges;
    return q.jv_slot_count != 0 && e.task_count != 0 && e.tiles == tiles && tiles <= kQueueMaxTiles &&
           e.task_offset <= q.task_capacity && tiles <= q.task_capacity - e.task_offset &&
           e.cost_offset <= q.cost_capacity &&
           (static_cast<uint64_t>(e.task_count) + 1) / 2 <= q.cost_capacity - e.cost_offset;
  }
  if (e.k < 3 || e.k > 5 || e.task_count == 0 || e.node_count == 0) return false;
  const KoptKLayout& L = a.layout[e.k - 3];
  return e.task_offset <= q.task_capacity && e.task_count <= q.task_capacity - e.task_offset &&
         e.cost_offset <= q.cost_capacity &&
         static_cast<uint64_t>(e.task_count) * L.T <= q.cost_capacity - e.cost_offset &&
         e.coord_base <= q.coord_count && e.node_count <= q.coord_count - e.coord_base &&
         e.tiles == (e.task_count + L.rows - 1) / L.rows && e.tiles <= kQueueMaxTiles;
}

// Refuses an entry: its flag says so and the queue goes out of service.
__device__ __forceinline__ void Refuse(const KoptQueueArgs& q, uint64_t entry, int lane) {
  if (lane == 0) {
    StoreAgent(&q.dev->refused, 1u);
    __hip_atomic_store(&q.flags[entry % q.ring_size], (entry + 1) | kQueueRefused, __ATOMIC_RELEASE,
                       __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// Copies entries [seen, tail) of the host ring to the mirror, eight at a
// time (lane = 8 x entry + word), writes their tiles' tickets and publishes
// both.  An entry whose sequence number or tile count is off is refused
// here; the tile waves check the rest of it.
__device__ void PublishEntries(const KoptQueueArgs& q, int lane, uint64_t seen, uint64_t tail) {
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");  // the host's entries and task rows
  uint64_t written = Uniform64(PollAgent(&q.dev->written));
  const uint32_t word = static_cast<uint32_t>(lane) & 7u, which = static_cast<uint32_t>(lane) >> 3;
  for (uint64_t base = seen; base < tail; base += 8) {
    const uint32_t count = tail - base < 8 ? static_cast<uint32_t>(tail - base) : 8u;
    uint64_t v = 0;
    if (which < count) {
      const uint64_t slot = (base + which) % q.ring_size;
      v = PollSystem(reinterpret_cast<const uint64_t*>(q.ring + slot) + word);
      StoreAgent(reinterpret_cast<uint64_t*>(q.mirror + slot) + word, v);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // entries before their tickets
    for (uint32_t j = 0; j < count; ++j) {
      const uint64_t entry = base + j;
      const uint64_t seq = static_cast<uint64_t>(__shfl(static_cast<long long>(v), static_cast<int>(8 * j)));
      // word 5: task_count in the low half, tiles in the high half
      const uint64_t w5 = static_cast<uint64_t>(__shfl(static_cast<long long>(v), static_cast<int>(8 * j + 5)));
      const uint32_t tiles = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(w5 >> 32));
      if (Uniform64(seq) != entry || tiles == 0 || tiles > kQueueMaxTiles || tiles > q.ticket_capacity) {
        Refuse(q, entry, lane);
        continue;
      }
      const uint64_t head = (entry + 1) << 24;
      for (uint32_t t = static_cast<uint32_t>(lane); t < tiles; t += kWave) {
        StoreAgent(&q.tickets[(written + t) % q.ticket_capacity], head | t);
      }
      written += tiles;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  if (lane == 0) {
    StoreAgent(&q.dev->written, written);
    StoreAgent(&q.dev->published, tail);
  }
}

// Workgroup 0: rewinds the ticket index past the tickets an ended launch
// handed out but never filled, opens the launch, then copies entries the
// host publishes to the mirror and writes their tickets, until the host sets
// stop, no entry arrives for idle_ticks, or the launch has lived life_ticks.
__device__ void QueuePoller(const KoptQueueArgs& q, int lane, uint64_t start) {
  // every wave of earlier launches has ended (one stream): no ticket of theirs is outstanding
  if (lane == 0) {
    StoreAgent(&q.dev->taken, PollAgent(&q.dev->written));
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    StoreAgent(&q.dev->open_epoch, q.epoch);
  }
  uint64_t seen = Uniform64(PollAgent(&q.dev->published));
  uint64_t last = start;
  for (;;) {
    const uint64_t tail = Uniform64(PollSystem(&q.host->tail));
    if (tail != seen) {
      if (tail < seen || tail - seen > q.ring_size) {  // never published so by the host
        if (lane == 0) StoreAgent(&q.dev->refused, 1u);
        break;
      }
      PublishEntries(q, lane, seen, tail);
      seen = tail;
      last = __builtin_amdgcn_s_memrealtime();
    }
    const uint64_t now = __builtin_amdgcn_s_memrealtime();
    const bool expired = now - start > q.life_ticks;
    const bool idle = now - last > q.idle_ticks;
    const uint32_t stop = idle || expired ? 0u : __builtin_amdgcn_readfirstlane(PollSystem(&q.host->stop));
    if (stop != 0 || expired || idle) {
      // Announce the end, then look at the tail once more.  The host stores
      // the tail before it reads the announcement, so an entry it published
      // without seeing the announcement is seen here; one it published later
      // goes to the launch it then starts.
      if (lane == 0) {
        __hip_atomic_store(&q.host->ended_epoch, static_cast<uint64_t>(q.epoch), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_SYSTEM);
      }
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "");
      const uint64_t again = Uniform64(PollSystem(&q.host->tail));
      if (again != seen && again > seen && again - seen <= q.ring_size) {
        if (idle && !expired && stop == 0) {
          // new work ends the idle spell: withdraw the announcement
          if (lane == 0) {
            __hip_atomic_store(&q.host->ended_epoch, uint64_t{0}, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
          }
          continue;
        }
        PublishEntries(q, lane, seen, again);  // the last entries of this launch
      }
      break;
    }
    if (tail == seen) __builtin_amdgcn_s_sleep(1);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  if (lane == 0) StoreAgent(&q.dev->exit_epoch, q.epoch);
}

template <int K>
__device__ __forceinline__ void QueueTile(const KoptLaunch& a, const KoptPointers& bufs, const KoptQueueEntry& e,
                                          uint32_t tile, const char* smem, char* wave, int lane) {
  const KoptKLayout& L = a.layout[K - 3];
  const uint32_t row0 = tile * L.rows;
  KoptTile d;
  d.task_index = e.task_offset + row0;
  d.cost_index = e.cost_offset + static_cast<uint64_t>(row0) * L.T;
  d.coord_base = e.coord_base;
  d.rows = e.task_count - row0 < L.rows ? e.task_count - row0 : L.rows;
  d.distance_type = e.distance_type;
  d.near = e.near;
  if (d.near != 0) {
    StagePorts<K, true, false, true>(a, bufs, L, d, wave, lane, e.node_count);
    TileCells<K, true, false, false>(a, bufs, L, d, smem, wave, lane);
  } else {
    StagePorts<K, false, false, true>(a, bufs, L, d, wave, lane, e.node_count);
    TileCells<K, false, false, false>(a, bufs, L, d, smem, wave, lane);
  }
}

// One JV recheck tile: up to kQueueJvTileEdges edge slots, copied from the
// task arena into the workgroup's staging row (a slot outside the union reads
// slot 0: rows written by another agent), tested in pairs with JvKernel's
// bodies into the row's witness half, then stored into the cost arena with
// system-scope stores.  Those write through with byte enables: tiles of one
// request, on different XCDs, share the cost arena's cache lines.
__device__ void QueueJvTile(const KoptLaunch& a, const KoptQueueArgs& q, const KoptQueueEntry& e, uint32_t tile,
                            jvbody::JvShared& jsh, int lane) {
  const uint32_t first = tile * kQueueJvTileEdges;
  const uint32_t n = e.task_count - first < kQueueJvTileEdges ? e.task_count - first : kQueueJvTileEdges;
  int32_t* const stage = q.jv_stage + static_cast<uint64_t>(blockIdx.x) * (2 * kQueueJvTileEdges);
  const int32_t* const src = reinterpret_cast<const int32_t*>(a.tasks + e.task_offset) + first;
  if (static_cast<uint32_t>(lane) < n) {
    const int32_t v = src[lane];
    stage[lane] = static_cast<uint64_t>(static_cast<uint32_t>(v)) < q.jv_slot_count ? v : 0;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  JvLaunch g = q.jv;
  g.edge_list = stage;
  g.edge_count = static_cast<int32_t>(n);
  g.witnesses = stage + kQueueJvTileEdges;
  for (int32_t pair = 0; pair < static_cast<int32_t>(n + 1) / 2; ++pair) {
    const uint64_t rest = __ballot(jvbody::PairOfEdges(g, jsh, 0, lane, pair));
    for (int h = 0; h < 2; ++h) {  // uniform
      if (((rest >> (jvbody::kHalf * h)) & 1ull) != 0) {
        const int32_t item = 2 * pair + h;
        const int32_t result = jvbody::OneEdge(g, jsh, 0, lane, item);
        if (lane == 0) g.witnesses[item] = result;
      }
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  int32_t* const out = reinterpret_cast<int32_t*>(a.costs + e.cost_offset) + first;
  if (static_cast<uint32_t>(lane) < n) {
    __hip_atomic_store(out + lane, g.witnesses[lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// One wavefront per workgroup.  Workgroup 0 polls the host ring; the others
// take a ticket index, wait for the ticket and run its tile with the batched
// kernel's tile code.
__global__ void __launch_bounds__(kWave) KoptQueueKernel(KoptLaunch a, KoptQueueArgs q) {
  extern __shared__ __attribute__((aligned(16))) char smem[];
  __shared__ jvbody::JvShared jsh;  // JV tiles use wavefront slot 0
  const int lane = static_cast<int>(threadIdx.x);
  const uint64_t start = __builtin_amdgcn_s_memrealtime();
  if (blockIdx.x == 0) {
    QueuePoller(q, lane, start);
    return;
  }
  for (int i = 0; i < 3; ++i) {
    const KoptKLayout& L = a.layout[i];
    uint16_t* s_pq = reinterpret_cast<uint16_t*>(smem + L.pq_off);
    uint64_t* s_tpl = reinterpret_cast<uint64_t*>(smem + L.tpl_off);
    for (uint32_t t = lane; t < L.T; t += kWave) s_tpl[t] = a.plans[i].tpl[t];
    for (uint32_t t = lane; t < L.P; t += kWave) s_pq[t] = a.plans[i].pair_pq[t];
  }
  __syncthreads();
  const KoptPointers bufs{a.xy, nullptr, a.tasks, a.plans, a.costs, nullptr, nullptr};
  char* const wave = smem + a.wave_base;
  // a poller that never ran ends nothing: every wait also ends on the launch's own clock
  const uint64_t limit = 2 * q.life_ticks + q.idle_ticks;
  // wait until this launch's poller has rewound the ticket index
  while (__builtin_amdgcn_readfirstlane(PollAgent(&q.dev->open_epoch)) != q.epoch) {
    if (__builtin_amdgcn_s_memrealtime() - start > limit) return;
    __builtin_amdgcn_s_sleep(4);
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  for (;;) {
    uint64_t index = 0;
    if (lane == 0) index = __hip_atomic_fetch_add(&q.dev->taken, uint64_t{1}, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    index = Uniform64(index);
    uint64_t* const slot = q.tickets + index % q.ticket_capacity;
    uint64_t ticket = 0;
    for (uint32_t polls = 0;; ++polls) {
      ticket = Uniform64(PollAgent(slot));
      if (ticket != 0) break;
      // every 16th look: has the poller ended (then no ticket comes), or the clock run out?
      if ((polls & 15u) == 15u) {
        const bool ended = __builtin_amdgcn_readfirstlane(PollAgent(&q.dev->exit_epoch)) == q.epoch;
        if (ended || __builtin_amdgcn_s_memrealtime() - start > limit) {
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
          ticket = Uniform64(PollAgent(slot));
          break;
        }
      }
      // nap briefly at first, longer once the wait has lasted
      if (polls < 512) {
        __builtin_amdgcn_s_sleep(2);
      } else {
        __builtin_amdgcn_s_sleep(16);
      }
    }
    if (ticket == 0) return;
    if (lane == 0) StoreAgent(slot, uint64_t{0});
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // the mirror's entry
    const uint64_t entry = (ticket >> 24) - 1;
    const uint32_t tile = static_cast<uint32_t>(ticket & kQueueMaxTiles);
    const KoptQueueEntry e = MirrorEntry(q, entry);
    if (e.seq != entry || !EntryInBounds(a, q, e) || tile >= e.tiles) {  // never so from the poller
      Refuse(q, entry, lane);
      continue;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");  // the host's task rows
    if (e.k == 0) {
      QueueJvTile(a, q, e, tile, jsh, lane);
    } else if (e.k == 3) {
      QueueTile<3>(a, bufs, e, tile, smem, wave, lane);
    } else if (e.k == 4) {
      QueueTile<4>(a, bufs, e, tile, smem, wave, lane);
    } else {
      QueueTile<5>(a, bufs, e, tile, smem, wave, lane);
    }
    // the tile's cells, then its mark: the same wave's stores reach host
    // memory in order
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
    const uint64_t slot_of_entry = entry % q.ring_size;
    uint64_t before = 0;
    if (lane == 0) {
      __hip_atomic_store(&q.marks[e.task_offset + tile], entry + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      before = __hip_atomic_fetch_add(&q.done[slot_of_entry], uint64_t{1}, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
    }
    before = Uniform64(before);
    // the last tile's wave raises the flag; the host still waits for every mark
    if (before + 1 == e.done_target && lane == 0) {
      __hip_atomic_store(&q.flags[slot_of_entry], entry + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
}

uint32_t Align16(uint32_t v) { return (v + 15u) & ~15u; }

}  // namespace

size_t KoptQueueLayout(KoptLaunch* a, const KoptPlan plans[3], const KoptTuning& tuning) {
  KoptTuning t = tuning;
  t.block = kWave;
  const bool present[3] = {true, true, true};
  a->masks = nullptr;
  a->mask_words = 0;
  KoptPrepareLayout(a, plans, present, t, false);
  // one wavefront holds the plans of all three k
  uint32_t off = 0;
  for (int i = 0; i < 3; ++i) {
    KoptKLayout& L = a->layout[i];
    L.pq_off = off;
    L.tpl_off = off + Align16(2 * kMaxPairs);
    off = Align16(L.tpl_off + L.T * 8);
  }
  a->wave_base = off;
  return static_cast<size_t>(off) + a->wave_bytes;
}

void LaunchKoptQueue(const KoptLaunch& a, const KoptQueueArgs& q, uint32_t workers, size_t shared_bytes,
                     hipStream_t stream) {
  if (workers == 0 || q.ring_size == 0 || q.epoch == 0) throw std::invalid_argument("bad k-opt queue launch");
  if (shared_bytes > 65536) throw std::invalid_argument("k-opt queue launch needs more than 64 KiB of LDS");
  hipLaunchKernelGGL(KoptQueueKernel, dim3(workers + 1), dim3(kWave), shared_bytes, stream, a, q);
  CheckHip(hipGetLastError(), "k-opt queue kernel launch");
}

size_t KoptPrepareLayout(KoptLaunch* a, const KoptPlan plans[3], const bool present[3],
                         const KoptTuning& tuning, bool near_only) {
  if (tuning.block == 0 || tuning.block > 256 || tuning.block % kWave != 0) {
    throw std::invalid_argument("k-opt block size must be a multiple of 64 up to 256");
  }
  uint32_t plan_bytes = 0;  // every workgroup holds the plan of its one k
  uint32_t wave_bytes = 0;
  for (int i = 0; i < 3; ++i) {
    KoptKLayout& L = a->layout[i];
    L = KoptKLayout{};
    if (!present[i]) continue;
    const KoptPlan& plan = plans[i];
    const uint32_t k = i + 3, rows = tuning.rows_per_tile[i];
    if (rows == 0 || rows > static_cast<uint32_t>(kWave)) {
      throw std::invalid_argument("rows per tile must be in [1, 64]");
    }
    if (plan.templates == 0 || plan.pairs == 0) throw std::invalid_argument("empty k-opt plan");
    L.present = 1;
    L.P = plan.pairs;
    L.T = plan.templates;
    L.rows = rows;
    if (L.T <= static_cast<uint32_t>(kWave)) {
      L.group = kWave / L.T;
      L.passes = 1;
    } else {
      L.group = 1;
      L.passes = (L.T + kWave - 1) / kWave;
    }
    // records per pair: at least rows, and odd, so that 32 lanes reading one
    // 8-byte record each of 32 different pairs hit 64 distinct LDS banks
    L.pair_stride = rows | 1u;
    L.pair_step_r = kWave / L.P;
    L.pair_step_p = kWave - L.pair_step_r * L.P;
    L.pq_off = 0;
    L.tpl_off = Align16(2 * kMaxPairs);
    plan_bytes = std::max(plan_bytes, Align16(L.tpl_off + L.T * 8));
    const uint32_t ports = 2 * k;
    // every block of a launch reserves the LDS of its largest k, so the
    // per-wave arrays are kept to what each phase needs: the same-node masks
    // of A2 live in the port records, the deleted costs only in mask mode
    L.xy_off = 0;
    // in mask mode the port records also hold the tile's mask words in phase C
    const uint32_t port_bytes = rows * ports * (near_only ? kNearPortBytes : kFarPortBytes);
    const uint32_t stage_bytes = a->masks != nullptr ? rows * a->mask_words * 4 : 0u;
    L.row_off = Align16(std::max(port_bytes, stage_bytes));
    L.dcost_off = L.row_off + Align16(rows * 4);
    L.pair_off = L.dcost_off + (a->masks != nullptr ? Align16(rows * 8) : 0u);
    // the port node ids of phase A share the pair record area
    const uint32_t pair_bytes = std::max(L.P * L.pair_stride * 8, rows * ports * 4);
    wave_bytes = std::max<uint32_t>(wave_bytes, L.pair_off + Align16(pair_bytes));
  }
  a->wave_base = plan_bytes;
  a->wave_bytes = wave_bytes;
  return static_cast<size_t>(plan_bytes) + static_cast<size_t>(wave_bytes) * (tuning.block / kWave);
}

uint64_t KoptCountTiles(const KoptSegment* segments, size_t count, const KoptTuning& tuning) {
  uint64_t tiles = 0;
  for (size_t i = 0; i < count; ++i) {
    const uint32_t rows = tuning.rows_per_tile[segments[i].k - 3];
    tiles += (segments[i].task_count + rows - 1) / rows;
  }
  return tiles;
}

void KoptAssignTiles(KoptSegment* segments, size_t count, const KoptInstance* instances,
                     const KoptPlan plans[3], const KoptTuning& tuning, KoptTile* tiles,
                     KoptLaunch* a) {
  uint64_t tile = 0;
  bool inline_ok = count <= kInlineSegments;
  uint32_t n_inline = 0;
  for (int i = 0; i < 3; ++i) {
    a->k_tile_begin[i] = tile;
    a->k_inline_begin[i] = n_inline;
    for (size_t s = 0; s < count; ++s) {
      KoptSegment& seg = segments[s];
      if (seg.k != static_cast<uint32_t>(3 + i)) continue;
      seg.rows_per_tile = tuning.rows_per_tile[i];
      seg.tile_begin = tile;
      const uint64_t n_tiles = (seg.task_count + seg.rows_per_tile - 1) / seg.rows_per_tile;
      const KoptInstance& inst = instances[seg.instance];
      if (inline_ok && tile <= UINT32_MAX && seg.task_count < (uint64_t{1} << 31) && inst.coord_base <= UINT32_MAX) {
        KoptInlineSegment& g = a->inline_seg[n_inline++];
        g.task_offset = seg.task_offset;
        g.cost_offset = seg.cost_offset;
        g.tile_begin = static_cast<uint32_t>(tile);
        g.task_count = static_cast<uint32_t>(seg.task_count);
        g.coord_base = static_cast<uint32_t>(inst.coord_base);
        g.distance_type = inst.distance_type;
        g.near = inst.near;
      } else {
        inline_ok = false;
      }
      if (tiles != nullptr) {
        const uint64_t T = plans[i].templates;
        for (uint64_t j = 0; j < n_tiles; ++j) {
          const uint64_t row0 = j * seg.rows_per_tile;
          KoptTile& d = tiles[tile + j];
          d.task_index = seg.task_offset + row0;
          d.cost_index = seg.cost_offset + row0 * T;
          d.coord_base = inst.coord_base;
          d.rows = static_cast<uint32_t>(std::min<uint64_t>(seg.rows_per_tile, seg.task_count - row0));
          d.distance_type = inst.distance_type;
          d.near = inst.near;
        }
      }
      tile += n_tiles;
    }
    a->k_tile_count[i] = tile - a->k_tile_begin[i];
  }
  a->k_inline_begin[3] = n_inline;
  a->tile_count = tile;
  a->inline_count = inline_ok && tile <= UINT32_MAX ? n_inline : 0;
}

uint32_t KoptFinalize(KoptLaunch* out, const KoptTuning& tuning, size_t shared_bytes) {
  const KoptLaunch& args = *out;
  if (args.tile_count == 0) return 0;
  if (shared_bytes > 65536) throw std::invalid_argument("k-opt launch needs more than 64 KiB of LDS");
  uint64_t tiles = 0;
  for (int i = 0; i < 3; ++i) tiles += args.k_tile_count[i];
  if (tiles != args.tile_count) throw std::logic_error("k-opt tile ranges do not cover the launch");
  const uint32_t waves_per_block = tuning.block / kWave;
  // workgroups per k: one tile per wavefront, or a share of tuning.grid in
  // proportion to the cells of each k
  uint64_t need[3], blocks[3];
  uint64_t total_need = 0;
  double work[3], total_work = 0.0;
  for (int i = 0; i < 3; ++i) {
    need[i] = (args.k_tile_count[i] + waves_per_block - 1) / waves_per_block;
    total_need += need[i];
    work[i] = args.k_tile_count[i] == 0
                  ? 0.0
                  : static_cast<double>(args.k_tile_count[i]) * args.layout[i].rows *
                        (args.layout[i].T + args.layout[i].P);
    total_work += work[i];
  }
  for (int i = 0; i < 3; ++i) {
    blocks[i] = need[i];
    if (tuning.grid != 0 && tuning.grid < total_need && need[i] > 0) {
      const double share = static_cast<double>(tuning.grid) * work[i] / total_work;
      blocks[i] = std::min<uint64_t>(need[i], std::max<uint64_t>(1, static_cast<uint64_t>(share + 0.5)));
    }
  }
  if (blocks[0] + blocks[1] + blocks[2] > 0x7fffffffull) {
    throw std::invalid_argument("k-opt launch needs too many workgroups");
  }
  if (args.masks != nullptr && args.mask_words == 0) {
    throw std::invalid_argument("mask output needs mask_words > 0");
  }
  uint32_t first = 0;
  for (int i = 0; i < 3; ++i) {
    out->k_block_first[i] = first;
    out->k_block_count[i] = static_cast<uint32_t>(blocks[i]);
    first += out->k_block_count[i];
  }
  return first;
}

void LaunchKopt(const KoptLaunch& args, const KoptTuning& tuning, size_t shared_bytes,
                hipStream_t stream) {
  KoptLaunch a = args;
  const uint32_t grid = KoptFinalize(&a, tuning, shared_bytes);
  if (grid == 0) return;
  const dim3 g(grid), b(tuning.block);
  if (a.masks != nullptr) {
    hipLaunchKernelGGL((KoptKernel<false, true>), g, b, shared_bytes, stream, a);
  } else if (tuning.non_temporal) {
    hipLaunchKernelGGL((KoptKernel<true, false>), g, b, shared_bytes, stream, a);
  } else {
    hipLaunchKernelGGL((KoptKernel<false, false>), g, b, shared_bytes, stream, a);
  }
  CheckHip(hipGetLastError(), "k-opt kernel launch");
}

void LaunchKoptResident(const KoptLaunch& finalized, const KoptLaunch* device_args, uint32_t grid,
                        const KoptTuning& tuning, size_t shared_bytes, hipStream_t stream,
                        unsigned long long* stamps) {
  if (grid == 0) return;
  if (device_args == nullptr) throw std::invalid_argument("resident k-opt arguments are NULL");
  const dim3 gd(grid), b(tuning.block);
  const KoptPointers ptrs{finalized.xy, finalized.tiles, finalized.tasks, finalized.plans, finalized.costs,
                          finalized.masks, stamps};
  if (finalized.masks != nullptr) {
    hipLaunchKernelGGL((KoptKernelResident<false, true>), gd, b, shared_bytes, stream, device_args, ptrs);
  } else if (tuning.non_temporal) {
    hipLaunchKernelGGL((KoptKernelResident<true, false>), gd, b, shared_bytes, stream, device_args, ptrs);
  } else {
    hipLaunchKernelGGL((KoptKernelResident<false, false>), gd, b, shared_bytes, stream, device_args, ptrs);
  }
  CheckHip(hipGetLastError(), "k-opt kernel launch");
}

}  // namespace ee
