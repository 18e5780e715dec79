// Content hash of snapshots' edge parts on the device (see ee_hash.h).
#include "ee_hash.h"

#include <algorithm>
#include <cmath>

namespace ee {
namespace {

constexpr uint64_t kFnvPrime = 1099511628211ull;  // 2^40 + 0x1b3
constexpr uint32_t kPrimeLow = 0xb3u;             // the prime mod 256
constexpr uint64_t kBytesPerSlot = 17;            // u, v, weight, active
constexpr uint32_t kStates = 256;                 // low bytes
constexpr uint32_t kMaxGroup = 256;               // chunks per group, groups per segment
constexpr uint32_t kMinEdgesPerChunk = 8;
constexpr uint32_t kPermThreads = kStates / 2;    // two starting bytes per thread
constexpr uint32_t kStageSlots = 256;             // slots staged in LDS at a time
constexpr uint32_t kThreads = 256;

using Chunk = DeviceContentHash::Chunk;
using Group = DeviceContentHash::Group;
using Segment = DeviceContentHash::Segment;

__device__ __forceinline__ uint64_t PrimePower(uint64_t e) {
  uint64_t r = 1, b = kFnvPrime;
  while (e != 0) {
    if ((e & 1u) != 0) r *= b;
    b *= b;
    e >>= 1;
  }
  return r;
}

// Two low bytes at once, one in each 16-bit half: l <- ((l xor b) * 0xb3)
// mod 256.  Each half of x is below 2^8, so x < 2^24 for the 24-bit
// multiply, and a half's product (below 2^16) does not reach the other half.
__device__ __forceinline__ uint32_t LowStep2(uint32_t l2, uint32_t b) {
  const uint32_t x = l2 ^ (b * 0x10001u);
  return __umul24(x, kPrimeLow) & 0x00ff00ffu;
}

// 1. Per chunk (one workgroup): the low-byte permutation of the chunk's
// bytes; thread t follows starting bytes t and t + 128.  Slots are staged in
// LDS as five words (u, v, weight low, weight high, active).
__global__ __launch_bounds__(kPermThreads) void ChunkPermKernel(const Chunk* __restrict__ chunks,
                                                                const int2* __restrict__ uv,
                                                                const int64_t* __restrict__ weight,
                                                                const int32_t* __restrict__ active,
                                                                uint8_t* __restrict__ perm) {
  __shared__ uint32_t words[kStageSlots * 5];
  const uint64_t c = blockIdx.x;
  const Chunk d = chunks[c];
  uint32_t l2 = threadIdx.x | ((threadIdx.x + kPermThreads) << 16);
  for (int64_t base = d.first; base < d.end; base += kStageSlots) {
    const uint32_t ns = static_cast<uint32_t>(min(static_cast<int64_t>(kStageSlots), d.end - base));
    __syncthreads();  // the previous batch is consumed
    for (uint32_t i = threadIdx.x; i < ns; i += kPermThreads) {
      const int2 p = uv[base + i];
      const uint64_t w = static_cast<uint64_t>(weight[base + i]);
      uint32_t* s = words + 5 * i;
      s[0] = static_cast<uint32_t>(p.x - d.node_offset);
      s[1] = static_cast<uint32_t>(p.y - d.node_offset);
      s[2] = static_cast<uint32_t>(w);
      s[3] = static_cast<uint32_t>(w >> 32);
      s[4] = active[base + i] != 0 ? 1u : 0u;
    }
    __syncthreads();
    for (uint32_t i = 0; i < ns; ++i) {
      const uint32_t* s = words + 5 * i;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t wd = s[k];
#pragma unroll
        for (int j = 0; j < 4; ++j) l2 = LowStep2(l2, (wd >> (8 * j)) & 0xffu);
      }
      l2 = LowStep2(l2, s[4]);
    }
  }
  uint8_t* dst = perm + c * kStates;
  dst[threadIdx.x] = static_cast<uint8_t>(l2);
  dst[threadIdx.x + kPermThreads] = static_cast<uint8_t>(l2 >> 16);
}

// Tables [first, first + count) of 256 bytes into LDS (16-byte copies).
__device__ __forceinline__ void StageTables(uint8_t* t, const uint8_t* __restrict__ tables, uint64_t first,
                                            uint32_t count) {
  const uint4* src = reinterpret_cast<const uint4*>(tables + first * kStates);
  uint4* dst = reinterpret_cast<uint4*>(t);
  for (uint32_t q = threadIdx.x; q < count * (kStates / 16); q += kThreads) dst[q] = src[q];
}

// 2. Per group of chunks: the composed permutation (thread = starting byte).
__global__ __launch_bounds__(kThreads) void GroupPermKernel(const Group* __restrict__ groups,
                                                            const uint8_t* __restrict__ perm,
                                                            uint8_t* __restrict__ group_perm) {
  __shared__ __attribute__((aligned(16))) uint8_t t[kMaxGroup * kStates];
  const Group d = groups[blockIdx.x];
  StageTables(t, perm, d.first_chunk, d.count);
  __syncthreads();
  uint32_t l = threadIdx.x;
  for (uint32_t i = 0; i < d.count; ++i) l = t[i * kStates + l];
  group_perm[static_cast<uint64_t>(blockIdx.x) * kStates + threadIdx.x] = static_cast<uint8_t>(l);
}

// 3. Per segment: the starting low byte of each of its groups, one walk from
// the prefix's; its result is set to prefix * P^(all its edge bytes), which
// step 5 adds to.
__global__ __launch_bounds__(kThreads) void TopWalkKernel(const Segment* __restrict__ segments,
                                                          const uint8_t* __restrict__ group_perm,
                                                          const uint64_t* __restrict__ prefixes,
                                                          uint8_t* __restrict__ group_start,
                                                          uint64_t* __restrict__ out) {
  __shared__ __attribute__((aligned(16))) uint8_t t[kMaxGroup * kStates];
  const Segment d = segments[blockIdx.x];
  StageTables(t, group_perm, d.first_group, d.groups);
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint64_t prefix = prefixes[blockIdx.x];
    uint32_t l = static_cast<uint32_t>(prefix & 0xffu);
    for (uint32_t g = 0; g < d.groups; ++g) {
      group_start[d.first_group + g] = static_cast<uint8_t>(l);
      l = t[g * kStates + l];
    }
    out[blockIdx.x] = prefix * PrimePower(d.slots * kBytesPerSlot);
  }
}

// 4. The starting low byte of every chunk: one walk per group.
__global__ __launch_bounds__(kThreads) void ChunkWalkKernel(const Group* __restrict__ groups,
                                                            const uint8_t* __restrict__ perm,
                                                            const uint8_t* __restrict__ group_start,
                                                            uint8_t* __restrict__ chunk_start) {
  __shared__ __attribute__((aligned(16))) uint8_t t[kMaxGroup * kStates];
  const Group d = groups[blockIdx.x];
  StageTables(t, perm, d.first_chunk, d.count);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t l = group_start[blockIdx.x];
    for (uint32_t i = 0; i < d.count; ++i) {
      chunk_start[static_cast<uint64_t>(d.first_chunk) + i] = static_cast<uint8_t>(l);
      l = t[i * kStates + l];
    }
  }
}

// 5. Per chunk (one thread): G at the chunk's starting low byte by Horner,
// g <- (g + d) * P with d = (l xor b) - l, times P^(bytes after the chunk in
// its segment), added to the segment's result (mod 2^64): one atomic per
// wavefront when its chunks share a segment, else one per chunk.
__global__ __launch_bounds__(kThreads) void ChunkSumKernel(const Chunk* __restrict__ chunks, uint64_t count,
                                                           const uint8_t* __restrict__ chunk_start,
                                                           const int2* __restrict__ uv,
                                                           const int64_t* __restrict__ weight,
                                                           const int32_t* __restrict__ active,
                                                           uint64_t* __restrict__ out) {
  const uint64_t c = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  uint64_t part = 0;
  int32_t segment = -1;
  if (c < count) {
    const Chunk d = chunks[c];
    segment = d.segment;
    uint32_t l = chunk_start[c];
    uint64_t g = 0;
    auto step = [&](uint32_t b) {
      const uint32_t x = l ^ b;
      const int64_t diff = static_cast<int64_t>(static_cast<int32_t>(x) - static_cast<int32_t>(l));
      g = (g + static_cast<uint64_t>(diff)) * kFnvPrime;
      l = (x * kPrimeLow) & 0xffu;
    };
    auto word = [&](uint32_t wd) {
#pragma unroll
      for (int j = 0; j < 4; ++j) step((wd >> (8 * j)) & 0xffu);
    };
    for (int64_t e = d.first; e < d.end; ++e) {
      const int2 p = uv[e];
      const uint64_t w = static_cast<uint64_t>(weight[e]);
      word(static_cast<uint32_t>(p.x - d.node_offset));
      word(static_cast<uint32_t>(p.y - d.node_offset));
      word(static_cast<uint32_t>(w));
      word(static_cast<uint32_t>(w >> 32));
      step(active[e] != 0 ? 1u : 0u);
    }
    part = g * PrimePower(static_cast<uint64_t>(d.segment_end - d.end) * kBytesPerSlot);
  }
  // lane 0 holds the wavefront's first chunk whenever the wavefront has one
  const int32_t first = __shfl(segment, 0, 64);
  const bool shared = __ballot(segment != first) == 0;
  if (!shared) {
    if (segment >= 0) atomicAdd(reinterpret_cast<unsigned long long*>(out + segment), static_cast<unsigned long long>(part));
    return;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const uint32_t lo = static_cast<uint32_t>(__shfl_down(static_cast<int>(part), off, 64));
    const uint32_t hi = static_cast<uint32_t>(__shfl_down(static_cast<int>(part >> 32), off, 64));
    part += (static_cast<uint64_t>(hi) << 32) | lo;
  }
  if ((threadIdx.x & 63u) == 0 && first >= 0) {
    atomicAdd(reinterpret_cast<unsigned long long*>(out + first), static_cast<unsigned long long>(part));
  }
}

}  // namespace

void DeviceContentHash::Allocate(const std::vector<size_t>& first, int device, hipStream_t stream,
                                 const std::vector<int32_t>& node_offset) {
  if (first.size() < 2) throw std::invalid_argument("device content hash: no segments");
  if (!node_offset.empty() && node_offset.size() + 1 != first.size()) {
    throw std::invalid_argument("device content hash: one node offset per segment");
  }
  std::vector<Chunk> ch;
  std::vector<Group> gr;
  std::vector<Segment> sg(first.size() - 1);
  const uint64_t max_chunks = uint64_t{kMaxGroup} * kMaxGroup;
  for (size_t s = 0; s + 1 < first.size(); ++s) {
    if (first[s + 1] < first[s] || first[s + 1] >= (size_t{1} << 31)) {
      throw std::invalid_argument("device content hash: bad segment bounds");
    }
    const uint64_t m = first[s + 1] - first[s];
    sg[s].first_group = static_cast<uint32_t>(gr.size());
    sg[s].groups = 0;
    sg[s].slots = m;
    if (m == 0) continue;
    const uint64_t e = std::max<uint64_t>(kMinEdgesPerChunk, (m + max_chunks - 1) / max_chunks);
    const uint64_t nch = (m + e - 1) / e;
    uint64_t gs = static_cast<uint64_t>(std::sqrt(static_cast<double>(nch)));
    while (gs * gs < nch) ++gs;
    gs = std::max<uint64_t>(gs, 1);
    const uint64_t ngr = (nch + gs - 1) / gs;
    if (gs > kMaxGroup || ngr > kMaxGroup) throw std::logic_error("device content hash: group layout");
    const uint64_t c0 = ch.size();
    for (uint64_t i = 0; i < nch; ++i) {
      Chunk c{};
      c.first = static_cast<int64_t>(first[s] + i * e);
      c.end = static_cast<int64_t>(std::min<uint64_t>(first[s] + (i + 1) * e, first[s + 1]));
      c.segment_end = static_cast<int64_t>(first[s + 1]);
      c.segment = static_cast<int32_t>(s);
      c.node_offset = node_offset.empty() ? 0 : node_offset[s];
      ch.push_back(c);
    }
    for (uint64_t j = 0; j < ngr; ++j) {
      Group g{};
      g.first_chunk = static_cast<uint32_t>(c0 + j * gs);
      g.count = static_cast<uint32_t>(std::min<uint64_t>(gs, nch - j * gs));
      g.segment = static_cast<uint32_t>(s);
      gr.push_back(g);
    }
    sg[s].groups = static_cast<uint32_t>(ngr);
  }
  if (ch.size() >= (size_t{1} << 32)) throw std::invalid_argument("device content hash: too many chunks");
  segments_ = sg.size();
  chunks_ = ch.size();
  groups_ = gr.size();
  chunk_desc_.Allocate(std::max<size_t>(chunks_, 1), device);
  group_desc_.Allocate(std::max<size_t>(groups_, 1), device);
  segment_desc_.Allocate(segments_, device);
  prefixes_.Allocate(segments_, device);
  perm_.Allocate(std::max<size_t>(chunks_, 1) * kStates, device);
  group_perm_.Allocate(std::max<size_t>(groups_, 1) * kStates, device);
  chunk_start_.Allocate(std::max<size_t>(chunks_, 1), device);
  group_start_.Allocate(std::max<size_t>(groups_, 1), device);
  if (chunks_ > 0) chunk_desc_.Upload(ch.data(), chunks_, stream);
  if (groups_ > 0) group_desc_.Upload(gr.data(), groups_, stream);
  segment_desc_.Upload(sg.data(), segments_, stream);
  CheckHip(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

void DeviceContentHash::SetPrefixes(const uint64_t* prefixes, hipStream_t stream) {
  if (segments_ == 0) throw std::logic_error("device content hash not allocated");
  prefixes_.Upload(prefixes, segments_, stream);
  CheckHip(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

void DeviceContentHash::Enqueue(const int2* uv, const int64_t* weight, const int32_t* active, uint64_t* out,
                                hipStream_t stream) {
  if (segments_ == 0) throw std::logic_error("device content hash not allocated");
  const uint32_t segs = static_cast<uint32_t>(segments_);
  if (chunks_ > 0) {
    hipLaunchKernelGGL(ChunkPermKernel, dim3(static_cast<uint32_t>(chunks_)), dim3(kPermThreads), 0, stream,
                       chunk_desc_.get(), uv, weight, active, perm_.get());
    CheckHip(hipGetLastError(), "hash permutation launch");
    hipLaunchKernelGGL(GroupPermKernel, dim3(static_cast<uint32_t>(groups_)), dim3(kThreads), 0, stream,
                       group_desc_.get(), perm_.get(), group_perm_.get());
    CheckHip(hipGetLastError(), "hash group launch");
  }
  hipLaunchKernelGGL(TopWalkKernel, dim3(segs), dim3(kThreads), 0, stream, segment_desc_.get(), group_perm_.get(),
                     prefixes_.get(), group_start_.get(), out);
  CheckHip(hipGetLastError(), "hash walk launch");
  if (chunks_ > 0) {
    hipLaunchKernelGGL(ChunkWalkKernel, dim3(static_cast<uint32_t>(groups_)), dim3(kThreads), 0, stream,
                       group_desc_.get(), perm_.get(), group_start_.get(), chunk_start_.get());
    CheckHip(hipGetLastError(), "hash chunk walk launch");
    hipLaunchKernelGGL(ChunkSumKernel, dim3(static_cast<uint32_t>((chunks_ + kThreads - 1) / kThreads)),
                       dim3(kThreads), 0, stream, chunk_desc_.get(), static_cast<uint64_t>(chunks_),
                       chunk_start_.get(), uv, weight, active, out);
    CheckHip(hipGetLastError(), "hash sum launch");
  }
}

}  // namespace ee

extern "C" {

ee_status ee_graph_content_hash_hip(ee_ctx* handle, int32_t n, uint8_t distance_type, const int64_t* x,
                                    const int64_t* y, int32_t m, const int32_t* edge_u, const int32_t* edge_v,
                                    const int64_t* edge_weight, const uint8_t* edge_active, uint64_t* hash) {
  try {
    if (handle == nullptr || hash == nullptr) throw std::invalid_argument("NULL argument");
    if (n < 0 || m < 0 || (n > 0 && (x == nullptr || y == nullptr))) throw std::invalid_argument("bad snapshot");
    if (m > 0 && (edge_u == nullptr || edge_v == nullptr || edge_weight == nullptr)) {
      throw std::invalid_argument("bad edge arrays");
    }
    auto* ctx = reinterpret_cast<ee::Context*>(handle);
    ctx->Activate();
    const uint64_t prefix = ee::ContentHashPrefix(n, distance_type, x, y);
    if (m == 0) {
      *hash = prefix;
      return EE_OK;
    }
    const size_t count = static_cast<size_t>(m);
    std::vector<int2> uv(count);
    std::vector<int32_t> active(count);
    for (size_t e = 0; e < count; ++e) {
      uv[e] = make_int2(edge_u[e], edge_v[e]);
      active[e] = edge_active == nullptr || edge_active[e] != 0 ? 1 : 0;
    }
    ee::DeviceArray<int2> d_uv(count, ctx->device);
    ee::DeviceArray<int64_t> d_weight(count, ctx->device);
    ee::DeviceArray<int32_t> d_active(count, ctx->device);
    ee::DeviceArray<uint64_t> d_out(1, ctx->device);
    d_uv.Upload(uv.data(), count, ctx->stream);
    d_weight.Upload(edge_weight, count, ctx->stream);
    d_active.Upload(active.data(), count, ctx->stream);
    ee::DeviceContentHash h;
    h.Allocate({0, count}, ctx->device, ctx->stream);
    h.SetPrefixes(&prefix, ctx->stream);
    h.Enqueue(d_uv.get(), d_weight.get(), d_active.get(), d_out.get(), ctx->stream);
    d_out.Download(hash, 1, ctx->stream);
    ctx->Sync();
    return EE_OK;
  } catch (...) {
    return ee::StatusFromException();
  }
}

}  // extern "C"
