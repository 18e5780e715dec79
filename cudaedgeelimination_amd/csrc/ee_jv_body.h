// The JV edge test of one wavefront, shared by JvKernel (ee_jv.hip) and the
// device work queue's kernel (ee_kopt.hip): constants, the per-wave LDS
// layout, and the one-edge and pair-of-edges bodies.  How the test works is
// described at the top of ee_jv.hip.
#pragma once

#include "ee_exact.h"
#include "ee_internal.h"

namespace ee {
namespace jvbody {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kCandidates = 10;  // jv_elimination.cpp:28, jv_cuda.cu:21

__device__ __forceinline__ void WaveSync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ bool KeyLess(uint64_t sa, int32_t na, uint64_t sb, int32_t nb) {
  return sa < sb || (sa == sb && na < nb);
}

__device__ __forceinline__ int32_t ReadLane(int32_t v, int lane) {
  return __builtin_amdgcn_readlane(v, lane);
}

__device__ __forceinline__ uint64_t ReadLane64(uint64_t v, int lane) {
  const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v), lane));
  const uint32_t hi =
      static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(v >> 32), lane));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Number of lanes j < lanes with bit j of `mask` set and key_j < (s, n).
__device__ __forceinline__ int CountLessInWave(uint64_t mask, int lanes, uint64_t my_s,
                                               int32_t my_n, uint64_t s, int32_t n) {
  int count = 0;
  for (int j = 0; j < lanes; ++j) {
    if (((mask >> j) & 1ull) == 0) continue;  // uniform branch
    const uint64_t js = ReadLane64(my_s, j);
    const int32_t jn = ReadLane(my_n, j);
    count += KeyLess(js, jn, s, n) ? 1 : 0;
  }
  return count;
}

// The FP64 path when the coordinate extents are below 2^25 (same results).
__device__ __forceinline__ int64_t Dist(const longlong2 p, const longlong2 q, int type, bool near) {
  if (near) {
    const int32_t dx = static_cast<int32_t>(static_cast<uint32_t>(p.x) - static_cast<uint32_t>(q.x));
    const int32_t dy = static_cast<int32_t>(static_cast<uint32_t>(p.y) - static_cast<uint32_t>(q.y));
    return ExactDistanceNear(dx, dy, type);
  }
  return ExactDistance(p.x, p.y, q.x, q.y, type);
}

// Packed candidate keys: (score << 24) | node, ordered as (score, node) when
// the score is below 2^40 and the node below 2^24.
constexpr int kNodeBits = 24;
constexpr int kScoreBits = 40;

constexpr int kHalf = kWave / 2;
// pair path LDS rows, padded so that the two halves' broadcast reads fall in
// different banks (16-byte multiples for the vector reads)
constexpr int kHalfNbrStride = 36;  // int32 entries
constexpr int kHalfKeyStride = 34;  // uint64 entries

struct JvShared {
  // one-edge path: the running top ten (double-buffered for the merge)
  uint64_t score[kWavesPerBlock][2][kCandidates];
  int32_t node[kWavesPerBlock][2][kCandidates];
  // one-edge path, one-pass rows: N(a) (padded with -1) and every lane's key
  alignas(16) int32_t nbr[kWavesPerBlock][kWave];
  alignas(16) uint64_t key[kWavesPerBlock][kWave];
  // the candidates' row bounds and D(a, c), D(b, c), stored by the lanes that
  // scored them: index 0 on the one-edge path, the half on the pair path
  int32_t cr0[kWavesPerBlock][2][kCandidates];
  int32_t cdeg[kWavesPerBlock][2][kCandidates];
  uint64_t cac[kWavesPerBlock][2][kCandidates];
  uint64_t cbc[kWavesPerBlock][2][kCandidates];
  // pair path: each half's N(a) and keys
  alignas(16) int32_t pnbr[kWavesPerBlock][2][kHalfNbrStride];
  alignas(16) uint64_t pkey[kWavesPerBlock][2][kHalfKeyStride];
};

// What the witness test needs of the edge (a, b) under test.
struct EdgeUnderTest {
  int32_t a = 0, b = 0;
  longlong2 pa{}, pb{};
  uint64_t wab = 0;
  int type = 0;
  bool near = false;
};

// Is neighbour d (adjacency entry `slot` of candidate c's row) compatible:
// w_ab + w_cd <= D(a,c) + D(d,b) or w_ab + w_cd <= D(a,d) + D(b,c), d not in
// {a, b} (jv_elimination.cpp:86-117).
__device__ __forceinline__ bool Compatible(const JvLaunch& g, const EdgeUnderTest& t, int32_t slot,
                                           uint64_t cac, uint64_t cbc) {
  const int32_t d = g.neighbors[slot];
  const int32_t eid = g.csr_edge_ids[slot];
  if (d == t.a || d == t.b) return false;
  const longlong2 pd = g.xy[d];
  const uint64_t lhs = t.wab + static_cast<uint64_t>(g.edge_weight[eid]);
  return lhs <= cac + static_cast<uint64_t>(Dist(pd, t.pb, t.type, t.near)) ||
         lhs <= static_cast<uint64_t>(Dist(t.pa, pd, t.type, t.near)) + cbc;
}

__device__ __forceinline__ uint32_t HalfBits(uint64_t ballot, int h) {
  return static_cast<uint32_t>(ballot >> (kHalf * h));
}

// One edge slot (item, wave-uniform) on the whole wavefront; returns the
// witness or -1.
__device__ __forceinline__ int32_t OneEdge(const JvLaunch& g, JvShared& sh, int wib, int lane,
                                           int32_t item) {
  const int32_t e = g.edge_list != nullptr ? g.edge_list[item] : item;
  int32_t result = -1;
  EdgeUnderTest t;
  t.type = g.distance_type;
  t.near = g.near != 0;
  if (g.edge_mode != nullptr) {
    const uint32_t mode = g.edge_mode[e];
    t.type = static_cast<int>(mode & 1u);
    t.near = (mode & 2u) != 0;
  }
  const int2 uv = g.edge_uv[e];
  const int32_t a = uv.x, b = uv.y;
  t.a = a;
  t.b = b;
  const int32_t ra0 = g.row_offsets[a], ra1 = g.row_offsets[a + 1];
  const int32_t rb0 = g.row_offsets[b], rb1 = g.row_offsets[b + 1];
  const int32_t deg_a = ra1 - ra0, deg_b = rb1 - rb0;
  if (g.edge_active[e] != 0 && deg_a > 2 && deg_b > 2) {
    const int type = t.type;
    const bool near = t.near;
    t.pa = g.xy[a];
    t.pb = g.xy[b];
    t.wab = static_cast<uint64_t>(g.edge_weight[e]);
    const longlong2 pa = t.pa, pb = t.pb;

    // ---- 1+2: score N(a) then N(b), keep the ten smallest (score, node)
    const int32_t total = deg_a + deg_b;
    const bool one_pass = total <= kWave;
    int cnt = 0, cur = 0;
    bool ranked = false;
    if (one_pass) {
      // both rows in one pass: lane s takes entry s of N(a) then N(b).  The
      // b side skips x in N(a) by a broadcast scan of N(a) in LDS; the keys
      // are ranked by counting smaller keys, two per LDS read, all lanes
      // reading the same address (a broadcast).
      const int32_t s = lane;
      const bool side_b = s >= deg_a;
      int32_t raw = -1, eid = 0;
      if (s < total) {
        const int32_t slot = side_b ? rb0 + (s - deg_a) : ra0 + s;
        raw = g.neighbors[slot];
        eid = g.csr_edge_ids[slot];
      }
      sh.nbr[wib][lane] = s < deg_a ? raw : -1;
      WaveSync();
      bool skip = s >= total || raw == a || raw == b;
      if (side_b && !skip) {
        const int4* row = reinterpret_cast<const int4*>(sh.nbr[wib]);
        bool in_a = false;
        for (int j = 0; j < (deg_a + 7) / 8; ++j) {  // uniform; entries past deg_a hold -1
#pragma unroll
          for (int u = 0; u < 2; ++u) {
            const int4 q = row[2 * j + u];
            in_a |= q.x == raw || q.y == raw || q.z == raw || q.w == raw;
          }
        }
        skip = in_a;
      }
      uint64_t sc = ~0ull, d_ax = 0, d_bx = 0;
      int32_t x_r0 = 0, x_r1 = 0;
      if (!skip) {
        // what the witness test needs of x if it becomes a candidate,
        // fetched alongside the score's operands
        x_r0 = g.row_offsets[raw];
        x_r1 = g.row_offsets[raw + 1];
        const longlong2 px = g.xy[raw];
        d_ax = static_cast<uint64_t>(Dist(pa, px, type, near));
        d_bx = static_cast<uint64_t>(Dist(pb, px, type, near));
        sc = static_cast<uint64_t>(g.edge_weight[eid]) + (side_b ? d_ax : d_bx);
      }
      const bool valid = !skip;
      const uint64_t wide = __ballot(valid && ((sc >> kScoreBits) != 0 || (raw >> kNodeBits) != 0));
      if (wide == 0) {
        const uint64_t key = valid ? (sc << kNodeBits) | static_cast<uint64_t>(raw) : ~0ull;
        sh.key[wib][lane] = key;
        WaveSync();
        const ulonglong2* keys = reinterpret_cast<const ulonglong2*>(sh.key[wib]);
        int rank = 0;
        for (int j = 0; j < (total + 1) / 2; ++j) {  // uniform; lanes past total hold ~0
          const ulonglong2 q = keys[j];
          rank += (q.x < key ? 1 : 0) + (q.y < key ? 1 : 0);
        }
        if (valid && rank < kCandidates) {
          sh.score[wib][1][rank] = sc;
          sh.node[wib][1][rank] = raw;
          sh.cr0[wib][0][rank] = x_r0;
          sh.cdeg[wib][0][rank] = x_r1 - x_r0;
          sh.cac[wib][0][rank] = d_ax;
          sh.cbc[wib][0][rank] = d_bx;
        }
        WaveSync();
        const int npass = __popcll(__ballot(valid));
        cnt = npass < kCandidates ? npass : kCandidates;
        cur = 1;
        ranked = true;
      }
    }
    for (int32_t base = 0; !ranked && base < total; base += kWave) {
      const int32_t s = base + lane;
      bool valid = false;
      uint64_t sc = ~0ull;
      int32_t x = INT32_MAX;
      int32_t raw = -1;
      const bool side_b = s >= deg_a;
      int32_t slot = 0;
      if (s < total) {
        slot = side_b ? rb0 + (s - deg_a) : ra0 + s;
        raw = g.neighbors[slot];
      }
      bool skip = s >= total || raw == a || raw == b;
      if (one_pass) {
        // every lane joins the (uniform) scan; lanes 0..deg_a-1 hold N(a)
        bool in_a = false;
        for (int j = 0; j < deg_a; ++j) in_a |= ReadLane(raw, j) == raw;
        skip |= side_b && in_a;
      } else if (side_b && !skip) {
        int32_t lo = ra0, hi = ra1;
        while (lo < hi) {
          const int32_t mid = lo + (hi - lo) / 2;
          if (g.neighbors[mid] < raw) {
            lo = mid + 1;
          } else {
            hi = mid;
          }
        }
        skip = lo < ra1 && g.neighbors[lo] == raw;
      }
      if (!skip) {
        const longlong2 px = g.xy[raw];
        sc = static_cast<uint64_t>(g.edge_weight[g.csr_edge_ids[slot]]) +
             static_cast<uint64_t>(Dist(px, side_b ? pa : pb, type, near));
        x = raw;
        valid = true;
      }
      const uint64_t vmask = __ballot(valid);
      const int lanes = total - base < kWave ? total - base : kWave;
      const int npass = __popcll(vmask);
      // rank among this pass plus the running list
      int rank = CountLessInWave(vmask, lanes, sc, x, sc, x);
      for (int j = 0; j < cnt; ++j) {
        rank += KeyLess(sh.score[wib][cur][j], sh.node[wib][cur][j], sc, x) ? 1 : 0;
      }
      // new positions of the running list entries
      uint64_t ls = ~0ull;
      int32_t ln = INT32_MAX;
      int lrank = kCandidates;
      if (lane < cnt) {
        ls = sh.score[wib][cur][lane];
        ln = sh.node[wib][cur][lane];
      }
      if (cnt > 0) {  // uniform: the first pass has no running list
        const int c = CountLessInWave(vmask, lanes, sc, x, ls, ln);
        if (lane < cnt) lrank = lane + c;
      }
      const int nxt = cur ^ 1;
      if (valid && rank < kCandidates) {
        sh.score[wib][nxt][rank] = sc;
        sh.node[wib][nxt][rank] = x;
      }
      if (lane < cnt && lrank < kCandidates) {
        sh.score[wib][nxt][lrank] = ls;
        sh.node[wib][nxt][lrank] = ln;
      }
      WaveSync();
      cnt = cnt + npass < kCandidates ? cnt + npass : kCandidates;
      cur = nxt;
    }

    // ---- 3: witness test in candidate order.  Lane q < cnt holds
    // candidate q (node, row, D(a,c), D(b,c)); consecutive candidates whose
    // rows fit 64 lanes together are tested in one pass, lane j testing the
    // j-th neighbour of its candidate, and the first candidate whose lanes
    // have no compatible neighbour is the witness.
    int32_t c_node = -1, c_r0 = 0, c_deg = 0;
    uint64_t c_ac = 0, c_bc = 0;
    if (lane < cnt) {
      c_node = sh.node[wib][cur][lane];
      if (ranked) {
        c_ac = sh.cac[wib][0][lane];
        c_bc = sh.cbc[wib][0][lane];
        c_r0 = sh.cr0[wib][0][lane];
        c_deg = sh.cdeg[wib][0][lane];
      } else {
        const longlong2 pc = g.xy[c_node];
        c_ac = static_cast<uint64_t>(Dist(pa, pc, type, near));
        c_bc = static_cast<uint64_t>(Dist(pb, pc, type, near));
        c_r0 = g.row_offsets[c_node];
        c_deg = g.row_offsets[c_node + 1] - c_r0;
      }
    }
    // rows of at most 16 entries (the usual case): candidate q takes lanes
    // 16 (q % 4) .. 16 (q % 4) + 15 of pass q / 4, so one ballot gives the
    // verdicts of four candidates as 16-bit chunks, with no per-candidate
    // bookkeeping; longer rows take the packing below
    constexpr int kSlot = 16;
    const bool slots = __ballot(lane < cnt && c_deg > kSlot) == 0;
    for (int base_q = 0; slots && base_q < cnt; base_q += kWave / kSlot) {  // uniform
      const int q = base_q + lane / kSlot, k = lane % kSlot;
      const int src = q < cnt ? q : 0;
      const int32_t r0 = __shfl(c_r0, src);
      const int32_t dq = __shfl(c_deg, src);
      const uint64_t cac = (static_cast<uint64_t>(static_cast<uint32_t>(__shfl(static_cast<int>(c_ac >> 32), src))) << 32) |
                           static_cast<uint32_t>(__shfl(static_cast<int>(c_ac), src));
      const uint64_t cbc = (static_cast<uint64_t>(static_cast<uint32_t>(__shfl(static_cast<int>(c_bc >> 32), src))) << 32) |
                           static_cast<uint32_t>(__shfl(static_cast<int>(c_bc), src));
      const uint64_t hits = __ballot(q < cnt && k < dq && Compatible(g, t, r0 + k, cac, cbc));
      bool found = false;
#pragma unroll 1
      for (int u = 0; u < kWave / kSlot; ++u) {  // uniform, in candidate order
        if (!found && base_q + u < cnt && ((hits >> (kSlot * u)) & 0xffffull) == 0) {
          result = ReadLane(c_node, base_q + u);
          found = true;
        }
      }
      if (found) break;
    }
    for (int i = 0; !slots && i < cnt;) {  // uniform
      const int deg_i = ReadLane(c_deg, i);
      if (deg_i > kWave) {
        // a row longer than a wavefront: this candidate alone, pass by pass
        const int32_t r0 = ReadLane(c_r0, i), r1 = r0 + deg_i;
        const uint64_t cac = ReadLane64(c_ac, i), cbc = ReadLane64(c_bc, i);
        bool compatible = false;
        for (int32_t base = r0; base < r1 && !compatible; base += kWave) {
          const int32_t slot = base + lane;
          compatible = __ballot(slot < r1 && Compatible(g, t, slot, cac, cbc)) != 0;
        }
        if (!compatible) {
          result = ReadLane(c_node, i);
          break;
        }
        ++i;
        continue;
      }
      // candidates i .. j-1: lanes [start_q, start_q + deg_q) test candidate q
      int j = i, used = 0, my_q = -1, my_k = 0;
      while (j < cnt) {
        const int dj = ReadLane(c_deg, j);
        if (used + dj > kWave) break;
        if (lane >= used && lane < used + dj) {
          my_q = j;
          my_k = lane - used;
        }
        used += dj;
        ++j;
      }
      const int src = my_q < 0 ? 0 : my_q;
      const int32_t r0 = __shfl(c_r0, src);
      const uint64_t cac = (static_cast<uint64_t>(static_cast<uint32_t>(__shfl(static_cast<int>(c_ac >> 32), src))) << 32) |
                           static_cast<uint32_t>(__shfl(static_cast<int>(c_ac), src));
      const uint64_t cbc = (static_cast<uint64_t>(static_cast<uint32_t>(__shfl(static_cast<int>(c_bc >> 32), src))) << 32) |
                           static_cast<uint32_t>(__shfl(static_cast<int>(c_bc), src));
      const uint64_t hits = __ballot(my_q >= 0 && Compatible(g, t, r0 + my_k, cac, cbc));
      bool found = false;
      for (int q = i, start = 0; q < j; ++q) {  // uniform
        const int dq = ReadLane(c_deg, q);
        const uint64_t span = dq >= kWave ? ~0ull : ((1ull << dq) - 1ull);
        if ((hits & (span << start)) == 0) {
          result = ReadLane(c_node, q);
          found = true;
          break;
        }
        start += dq;
      }
      if (found) break;
      i = j;
    }
  }
  WaveSync();
  return result;
}

// Edge slots 2 pair and 2 pair + 1 on the two 32-lane halves.  Lane 0 of a
// half writes its edge's witness (or -1), unless the edge does not fit the
// half; then the function returns true on that half's lanes and writes
// nothing for it.
__device__ __forceinline__ bool PairOfEdges(const JvLaunch& g, JvShared& sh, int wib, int lane,
                                            int32_t pair) {
  const int h = lane / kHalf, hl = lane % kHalf;
  const int32_t item = 2 * pair + h;
  const bool present = item < g.edge_count;
  EdgeUnderTest t;
  t.type = g.distance_type;
  t.near = g.near != 0;
  int32_t e = 0, ra0 = 0, rb0 = 0, deg_a = 0, deg_b = 0;
  bool eligible = false;
  if (present) {
    e = g.edge_list != nullptr ? g.edge_list[item] : item;
    if (g.edge_mode != nullptr) {
      const uint32_t mode = g.edge_mode[e];
      t.type = static_cast<int>(mode & 1u);
      t.near = (mode & 2u) != 0;
    }
    const int2 uv = g.edge_uv[e];
    t.a = uv.x;
    t.b = uv.y;
    ra0 = g.row_offsets[t.a];
    rb0 = g.row_offsets[t.b];
    deg_a = g.row_offsets[t.a + 1] - ra0;
    deg_b = g.row_offsets[t.b + 1] - rb0;
    eligible = g.edge_active[e] != 0 && deg_a > 2 && deg_b > 2;
  }
  const int32_t total = deg_a + deg_b;
  bool fallback = eligible && total > kHalf;
  bool work = eligible && !fallback;  // uniform within the half from here on
  if (work) {
    t.pa = g.xy[t.a];
    t.pb = g.xy[t.b];
    t.wab = static_cast<uint64_t>(g.edge_weight[e]);
  }

  // ---- 1: lane hl scores entry hl of N(a) then N(b)
  const bool side_b = hl >= deg_a;
  int32_t raw = -1, eid = 0;
  if (work && hl < total) {
    const int32_t slot = side_b ? rb0 + (hl - deg_a) : ra0 + hl;
    raw = g.neighbors[slot];
    eid = g.csr_edge_ids[slot];
  }
  sh.pnbr[wib][h][hl] = hl < deg_a ? raw : -1;
  // both halves scan their N(a) for the same number of steps (deg_a <= 29
  // when the edge fits, so 4 steps of 8 entries stay in the half's row)
  const int32_t scan_a = work ? deg_a : 0;
  const int32_t scan_max = max(ReadLane(scan_a, 0), ReadLane(scan_a, kHalf));
  WaveSync();
  bool skip = !work || hl >= total || raw == t.a || raw == t.b;
  if (side_b && !skip) {
    const int4* row = reinterpret_cast<const int4*>(sh.pnbr[wib][h]);
    bool in_a = false;
    for (int j = 0; j < (scan_max + 7) / 8; ++j) {  // uniform; entries past deg_a hold -1
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int4 q = row[2 * j + u];
        in_a |= q.x == raw || q.y == raw || q.z == raw || q.w == raw;
      }
    }
    skip = in_a;
  }
  uint64_t sc = ~0ull, d_ax = 0, d_bx = 0;
  int32_t x_r0 = 0, x_deg = 0;
  if (!skip) {
    x_r0 = g.row_offsets[raw];
    x_deg = g.row_offsets[raw + 1] - x_r0;
    const longlong2 px = g.xy[raw];
    const uint64_t wx = static_cast<uint64_t>(g.edge_weight[eid]);
    d_ax = static_cast<uint64_t>(Dist(t.pa, px, t.type, t.near));
    d_bx = static_cast<uint64_t>(Dist(t.pb, px, t.type, t.near));
    sc = wx + (side_b ? d_ax : d_bx);
  }
  const bool scored = !skip;
  if (HalfBits(__ballot(scored && ((sc >> kScoreBits) != 0 || (raw >> kNodeBits) != 0)), h) != 0) {
    fallback = true;  // a key that does not pack: the one-edge path ranks it
    work = false;
  }

  // ---- 2: rank the half's packed keys by broadcast reads
  const bool valid = scored && work;
  const uint64_t key = valid ? (sc << kNodeBits) | static_cast<uint64_t>(raw) : ~0ull;
  sh.pkey[wib][h][hl] = key;
  const int32_t rank_n = work ? total : 0;
  const int32_t rank_max = max(ReadLane(rank_n, 0), ReadLane(rank_n, kHalf));
  WaveSync();
  const ulonglong2* keys = reinterpret_cast<const ulonglong2*>(sh.pkey[wib][h]);
  int rank = 0;
  for (int j = 0; j < (rank_max + 1) / 2; ++j) {  // uniform; lanes past total hold ~0
    const ulonglong2 q = keys[j];
    rank += (q.x < key ? 1 : 0) + (q.y < key ? 1 : 0);
  }
  if (valid && rank < kCandidates) {
    sh.node[wib][h][rank] = raw;
    sh.cr0[wib][h][rank] = x_r0;
    sh.cdeg[wib][h][rank] = x_deg;
    sh.cac[wib][h][rank] = d_ax;
    sh.cbc[wib][h][rank] = d_bx;
  }
  const int in_half = __popc(HalfBits(__ballot(valid), h));
  const int cnt = in_half < kCandidates ? in_half : kCandidates;
  WaveSync();

  // ---- 3: witness test.  Most candidates have a compatible neighbour among
  // the first three of their row (94% on k-NN graphs), so one probe pass
  // gives every candidate 3 lanes (candidate hl / 3, neighbours 0..2); then,
  // in rank order, the lowest candidate without a compatible neighbour yet
  // gets its whole remaining row (at most 29 entries, one per lane) per pass.
  // The first candidate whose row has no compatible neighbour is the
  // witness, as in the reference's scan in rank order with early exit.  A
  // candidate row longer than 32 sends the edge to the one-edge path.
  constexpr int kProbe = 3;
  const int32_t my_deg = hl < cnt ? sh.cdeg[wib][h][hl] : 0;
  if (HalfBits(__ballot(my_deg > kHalf), h) != 0) {
    fallback = true;
    work = false;
  }
  uint32_t hitmask = 0;  // per half: candidates with a compatible neighbour
  {
    const int q = (hl * 11) >> 5;  // hl / 3 for hl < 32
    const int k = hl - kProbe * q;
    bool hit = false;
    if (work && q < cnt) {
      const int32_t dq = sh.cdeg[wib][h][q];
      if (k < dq) hit = Compatible(g, t, sh.cr0[wib][h][q] + k, sh.cac[wib][h][q], sh.cbc[wib][h][q]);
    }
    const uint32_t lanes = HalfBits(__ballot(hit), h);
    // lane q < cnt gathers candidate q's three probe lanes
    hitmask = HalfBits(__ballot(hl < cnt && ((lanes >> (kProbe * hl)) & 7u) != 0), h);
  }
  int32_t result = -1;
  bool done = !work || cnt == 0;
  while (__ballot(!done) != 0) {  // uniform
    const uint32_t open = ~hitmask & ((1u << cnt) - 1u);
    int q = 0;
    int32_t dq = 0;
    if (!done) {
      if (open == 0) {
        done = true;  // every candidate has a compatible neighbour: no witness
      } else {
        q = __builtin_ctz(open);
        dq = sh.cdeg[wib][h][q];
        if (dq <= kProbe) {  // the probe saw its whole row
          result = sh.node[wib][h][q];
          done = true;
        }
      }
    }
    const bool test = !done && kProbe + hl < dq;
    bool hit = false;
    if (test) hit = Compatible(g, t, sh.cr0[wib][h][q] + kProbe + hl, sh.cac[wib][h][q], sh.cbc[wib][h][q]);
    const bool any = HalfBits(__ballot(hit), h) != 0;
    if (!done) {
      if (any) {
        hitmask |= 1u << q;
      } else {
        result = sh.node[wib][h][q];
        done = true;
      }
    }
  }
  if (hl == 0 && present && !fallback) g.witnesses[item] = result;
  WaveSync();
  return fallback;
}

}  // namespace jvbody
}  // namespace ee
