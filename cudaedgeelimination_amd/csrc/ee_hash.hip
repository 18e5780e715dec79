// Content hash of a snapshot's edge part on the device (see ee_hash.h).
#include "ee_hash.h"

#include <algorithm>
#include <cmath>

namespace ee {
namespace {

constexpr uint64_t kFnvPrime = 1099511628211ull;  // 2^40 + 0x1b3
constexpr uint32_t kPrimeLow = 0xb3u;             // the prime mod 256
constexpr uint64_t kBytesPerSlot = 17;            // u, v, weight, active
constexpr uint32_t kStates = 256;                 // low bytes
constexpr uint32_t kMaxGroup = 256;               // chunks per group, groups per walk
constexpr uint32_t kMinEdgesPerChunk = 8;
constexpr uint32_t kPermThreads = kStates / 2;    // two starting bytes per thread
constexpr uint32_t kStageSlots = 256;             // slots staged in LDS at a time
constexpr uint32_t kThreads = 256;

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
// LDS as five words (u, v, weight low, weight high, active).  Block 0 also
// clears the output, which step 5 adds into.
__global__ __launch_bounds__(kPermThreads) void ChunkPermKernel(const int2* __restrict__ uv,
                                                                const int64_t* __restrict__ weight,
                                                                const int32_t* __restrict__ active, int64_t m,
                                                                uint32_t edges_per_chunk, uint8_t* __restrict__ perm,
                                                                uint64_t* __restrict__ out) {
  __shared__ uint32_t words[kStageSlots * 5];
  const uint64_t c = blockIdx.x;
  const int64_t first = static_cast<int64_t>(c) * edges_per_chunk;
  const int64_t end = min(first + static_cast<int64_t>(edges_per_chunk), m);
  if (c == 0 && threadIdx.x == 0) *out = 0;
  uint32_t l2 = threadIdx.x | ((threadIdx.x + kPermThreads) << 16);
  for (int64_t base = first; base < end; base += kStageSlots) {
    const uint32_t ns = static_cast<uint32_t>(min(static_cast<int64_t>(kStageSlots), end - base));
    __syncthreads();  // the previous batch is consumed
    for (uint32_t i = threadIdx.x; i < ns; i += kPermThreads) {
      const int2 p = uv[base + i];
      const uint64_t w = static_cast<uint64_t>(weight[base + i]);
      uint32_t* s = words + 5 * i;
      s[0] = static_cast<uint32_t>(p.x);
      s[1] = static_cast<uint32_t>(p.y);
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

// The permutations of chunks [c0, c0 + count) into LDS (16-byte copies).
__device__ __forceinline__ void StagePerms(uint8_t* t, const uint8_t* __restrict__ perm, uint64_t c0,
                                           uint32_t count) {
  const uint4* src = reinterpret_cast<const uint4*>(perm + c0 * kStates);
  uint4* dst = reinterpret_cast<uint4*>(t);
  for (uint32_t q = threadIdx.x; q < count * (kStates / 16); q += kThreads) dst[q] = src[q];
}

// 2. Per group of chunks: the composed permutation (thread = starting byte).
__global__ __launch_bounds__(kThreads) void GroupPermKernel(const uint8_t* __restrict__ perm, uint64_t chunks,
                                                            uint32_t group_size, uint8_t* __restrict__ group_perm) {
  __shared__ __attribute__((aligned(16))) uint8_t t[kMaxGroup * kStates];
  const uint64_t c0 = static_cast<uint64_t>(blockIdx.x) * group_size;
  const uint32_t count = static_cast<uint32_t>(min(static_cast<uint64_t>(group_size), chunks - c0));
  StagePerms(t, perm, c0, count);
  __syncthreads();
  uint32_t l = threadIdx.x;
  for (uint32_t i = 0; i < count; ++i) l = t[i * kStates + l];
  group_perm[static_cast<uint64_t>(blockIdx.x) * kStates + threadIdx.x] = static_cast<uint8_t>(l);
}

// 3. The starting low byte of every group: one walk from the prefix's.
__global__ __launch_bounds__(kThreads) void TopWalkKernel(const uint8_t* __restrict__ group_perm, uint32_t groups,
                                                          uint32_t start, uint8_t* __restrict__ group_start) {
  __shared__ __attribute__((aligned(16))) uint8_t t[kMaxGroup * kStates];
  StagePerms(t, group_perm, 0, groups);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t l = start;
    for (uint32_t g = 0; g < groups; ++g) {
      group_start[g] = static_cast<uint8_t>(l);
      l = t[g * kStates + l];
    }
  }
}

// 4. The starting low byte of every chunk: one walk per group.
__global__ __launch_bounds__(kThreads) void ChunkWalkKernel(const uint8_t* __restrict__ perm, uint64_t chunks,
                                                            uint32_t group_size,
                                                            const uint8_t* __restrict__ group_start,
                                                            uint8_t* __restrict__ chunk_start) {
  __shared__ __attribute__((aligned(16))) uint8_t t[kMaxGroup * kStates];
  const uint64_t c0 = static_cast<uint64_t>(blockIdx.x) * group_size;
  const uint32_t count = static_cast<uint32_t>(min(static_cast<uint64_t>(group_size), chunks - c0));
  StagePerms(t, perm, c0, count);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t l = group_start[blockIdx.x];
    for (uint32_t i = 0; i < count; ++i) {
      chunk_start[c0 + i] = static_cast<uint8_t>(l);
      l = t[i * kStates + l];
    }
  }
}

// 5. Per chunk (one thread): G at the chunk's starting low byte by Horner,
// g <- (g + d) * P with d = (l xor b) - l, times P^(bytes after the chunk);
// plus prefix * P^(all edge bytes); summed into *out (mod 2^64).
__global__ __launch_bounds__(kThreads) void ChunkSumKernel(const int2* __restrict__ uv,
                                                           const int64_t* __restrict__ weight,
                                                           const int32_t* __restrict__ active, int64_t m,
                                                           uint32_t edges_per_chunk, uint64_t chunks,
                                                           const uint8_t* __restrict__ chunk_start, uint64_t prefix,
                                                           uint64_t* __restrict__ out) {
  const uint64_t c = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  uint64_t part = 0;
  if (c < chunks) {
    const int64_t first = static_cast<int64_t>(c) * edges_per_chunk;
    const int64_t end = min(first + static_cast<int64_t>(edges_per_chunk), m);
    uint32_t l = chunk_start[c];
    uint64_t g = 0;
    auto step = [&](uint32_t b) {
      const uint32_t x = l ^ b;
      const int64_t d = static_cast<int64_t>(static_cast<int32_t>(x) - static_cast<int32_t>(l));
      g = (g + static_cast<uint64_t>(d)) * kFnvPrime;
      l = (x * kPrimeLow) & 0xffu;
    };
    auto word = [&](uint32_t wd) {
#pragma unroll
      for (int j = 0; j < 4; ++j) step((wd >> (8 * j)) & 0xffu);
    };
    for (int64_t e = first; e < end; ++e) {
      const int2 p = uv[e];
      const uint64_t w = static_cast<uint64_t>(weight[e]);
      word(static_cast<uint32_t>(p.x));
      word(static_cast<uint32_t>(p.y));
      word(static_cast<uint32_t>(w));
      word(static_cast<uint32_t>(w >> 32));
      step(active[e] != 0 ? 1u : 0u);
    }
    part = g * PrimePower(static_cast<uint64_t>(m - end) * kBytesPerSlot);
  }
  if (c == 0) part += prefix * PrimePower(static_cast<uint64_t>(m) * kBytesPerSlot);
  // wave sum (lane 0 ends with it), one atomic per wavefront
  for (int off = 32; off > 0; off >>= 1) {
    const uint32_t lo = static_cast<uint32_t>(__shfl_down(static_cast<int>(part), off, 64));
    const uint32_t hi = static_cast<uint32_t>(__shfl_down(static_cast<int>(part >> 32), off, 64));
    part += (static_cast<uint64_t>(hi) << 32) | lo;
  }
  if ((threadIdx.x & 63u) == 0) {
    atomicAdd(reinterpret_cast<unsigned long long*>(out), static_cast<unsigned long long>(part));
  }
}

}  // namespace

void DeviceContentHash::Allocate(size_t m, int device) {
  if (m == 0 || m >= (size_t{1} << 31)) throw std::invalid_argument("device content hash: edge slots out of range");
  m_ = m;
  const uint64_t max_chunks = uint64_t{kMaxGroup} * kMaxGroup;
  edges_per_chunk_ = static_cast<uint32_t>(std::max<uint64_t>(kMinEdgesPerChunk, (m + max_chunks - 1) / max_chunks));
  chunks_ = (m + edges_per_chunk_ - 1) / edges_per_chunk_;
  uint64_t gs = static_cast<uint64_t>(std::sqrt(static_cast<double>(chunks_)));
  while (gs * gs < chunks_) ++gs;
  group_size_ = static_cast<uint32_t>(std::max<uint64_t>(gs, 1));
  groups_ = (chunks_ + group_size_ - 1) / group_size_;
  if (group_size_ > kMaxGroup || groups_ > kMaxGroup) throw std::logic_error("device content hash: group layout");
  perm_.Allocate(chunks_ * kStates, device);
  group_perm_.Allocate(groups_ * kStates, device);
  chunk_start_.Allocate(chunks_, device);
  group_start_.Allocate(groups_, device);
}

void DeviceContentHash::Enqueue(uint64_t prefix, const int2* uv, const int64_t* weight, const int32_t* active,
                                uint64_t* out, hipStream_t stream) {
  if (m_ == 0) throw std::logic_error("device content hash not allocated");
  const int64_t m = static_cast<int64_t>(m_);
  hipLaunchKernelGGL(ChunkPermKernel, dim3(static_cast<uint32_t>(chunks_)), dim3(kPermThreads), 0, stream, uv,
                     weight, active, m, edges_per_chunk_, perm_.get(), out);
  CheckHip(hipGetLastError(), "hash permutation launch");
  hipLaunchKernelGGL(GroupPermKernel, dim3(static_cast<uint32_t>(groups_)), dim3(kThreads), 0, stream, perm_.get(),
                     static_cast<uint64_t>(chunks_), group_size_, group_perm_.get());
  CheckHip(hipGetLastError(), "hash group launch");
  hipLaunchKernelGGL(TopWalkKernel, dim3(1), dim3(kThreads), 0, stream, group_perm_.get(),
                     static_cast<uint32_t>(groups_), static_cast<uint32_t>(prefix & 0xffu), group_start_.get());
  CheckHip(hipGetLastError(), "hash walk launch");
  hipLaunchKernelGGL(ChunkWalkKernel, dim3(static_cast<uint32_t>(groups_)), dim3(kThreads), 0, stream, perm_.get(),
                     static_cast<uint64_t>(chunks_), group_size_, group_start_.get(), chunk_start_.get());
  CheckHip(hipGetLastError(), "hash chunk walk launch");
  hipLaunchKernelGGL(ChunkSumKernel, dim3(static_cast<uint32_t>((chunks_ + kThreads - 1) / kThreads)),
                     dim3(kThreads), 0, stream, uv, weight, active, m, edges_per_chunk_,
                     static_cast<uint64_t>(chunks_), chunk_start_.get(), prefix, out);
  CheckHip(hipGetLastError(), "hash sum launch");
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
    h.Allocate(count, ctx->device);
    h.Enqueue(prefix, d_uv.get(), d_weight.get(), d_active.get(), d_out.get(), ctx->stream);
    d_out.Download(hash, 1, ctx->stream);
    ctx->Sync();
    return EE_OK;
  } catch (...) {
    return ee::StatusFromException();
  }
}

}  // extern "C"
