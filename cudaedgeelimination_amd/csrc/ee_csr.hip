// Commit on the device: clear the active flags of committed edge slots and
// compact the CSR (see ee_csr.h).  Four launches, each a pass over its
// arrays: deactivate, per-node live degree with per-chunk totals, offset
// scan, scatter of the live entries.
#include "ee_csr.h"

namespace ee {
namespace {

constexpr uint32_t kThreads = 256;
constexpr uint32_t kPerThread = 4;
constexpr uint32_t kChunk = kThreads * kPerThread;  // nodes per workgroup of the count and scan passes

__global__ __launch_bounds__(kThreads) void DeactivateKernel(const int32_t* __restrict__ commits,
                                                             int32_t capacity, int32_t* __restrict__ active,
                                                             int32_t m) {
  const int32_t count = min(commits[0], capacity);
  for (int32_t i = static_cast<int32_t>(blockIdx.x * blockDim.x + threadIdx.x); i < count;
       i += static_cast<int32_t>(gridDim.x * blockDim.x)) {
    const int32_t e = commits[1 + i];
    if (e >= 0 && e < m) active[e] = 0;
  }
}

// counts[v] = live entries of row v; chunk_sums[c] = their sum over chunk c.
__global__ __launch_bounds__(kThreads) void LiveDegreeKernel(const int32_t* __restrict__ row_offsets,
                                                             const int32_t* __restrict__ edge_ids,
                                                             const int32_t* __restrict__ active, int32_t n,
                                                             int32_t* __restrict__ counts,
                                                             unsigned long long* __restrict__ chunk_sums) {
  __shared__ unsigned long long wave_sums[kThreads / 64];
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  const int64_t first = static_cast<int64_t>(blockIdx.x) * kChunk + static_cast<int64_t>(t) * kPerThread;
  unsigned long long own = 0;
#pragma unroll
  for (uint32_t j = 0; j < kPerThread; ++j) {
    const int64_t v = first + j;
    if (v < n) {
      int32_t live = 0;
      for (int32_t a = row_offsets[v]; a < row_offsets[v + 1]; ++a) live += active[edge_ids[a]] != 0;
      counts[v] = live;
      own += static_cast<unsigned long long>(live);
    }
  }
  for (uint32_t d = 32; d >= 1; d >>= 1) own += __shfl_xor(own, d, 64);
  if (lane == 0) wave_sums[wave] = own;
  __syncthreads();
  if (t == 0) {
    unsigned long long s = 0;
    for (uint32_t w = 0; w < kThreads / 64; ++w) s += wave_sums[w];
    chunk_sums[blockIdx.x] = s;
  }
}

// new_ro[0] = 0, new_ro[v + 1] = counts[0] + ... + counts[v]: each
// workgroup sums the totals of the chunks before its own, then scans its
// chunk (4 consecutive counts per thread, wavefront shuffles, wavefront
// totals in LDS).
__global__ __launch_bounds__(kThreads) void OffsetScanKernel(const int32_t* __restrict__ counts, int32_t n,
                                                             const unsigned long long* __restrict__ chunk_sums,
                                                             int32_t* __restrict__ new_ro) {
  __shared__ unsigned long long partial[kThreads / 64];
  __shared__ unsigned long long wave_sums[kThreads / 64];
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  const uint64_t chunk = blockIdx.x;
  unsigned long long before = 0;
  for (uint64_t c = t; c < chunk; c += kThreads) before += chunk_sums[c];
  for (uint32_t d = 32; d >= 1; d >>= 1) before += __shfl_xor(before, d, 64);
  if (lane == 0) partial[wave] = before;
  const int64_t first = static_cast<int64_t>(chunk) * kChunk + static_cast<int64_t>(t) * kPerThread;
  unsigned long long v[kPerThread];
  unsigned long long own = 0;
#pragma unroll
  for (uint32_t j = 0; j < kPerThread; ++j) {
    v[j] = first + j < n ? static_cast<unsigned long long>(counts[first + j]) : 0ull;
    own += v[j];
  }
  unsigned long long inc = own;
  for (uint32_t d = 1; d < 64; d <<= 1) {
    const unsigned long long o = __shfl_up(inc, d, 64);
    if (lane >= d) inc += o;
  }
  if (lane == 63) wave_sums[wave] = inc;
  __syncthreads();
  unsigned long long run = inc - own;
  for (uint32_t w = 0; w < kThreads / 64; ++w) {
    run += partial[w];
    if (w < wave) run += wave_sums[w];
  }
  if (chunk == 0 && t == 0) new_ro[0] = 0;
#pragma unroll
  for (uint32_t j = 0; j < kPerThread; ++j) {
    run += v[j];
    if (first + j < n) new_ro[first + j + 1] = static_cast<int32_t>(run);
  }
}

// Row v's live entries, in row order, to [new_ro[v], new_ro[v + 1]).
__global__ __launch_bounds__(kThreads) void ScatterLiveKernel(const int32_t* __restrict__ row_offsets,
                                                              const int32_t* __restrict__ neighbors,
                                                              const int32_t* __restrict__ edge_ids,
                                                              const int32_t* __restrict__ active, int32_t n,
                                                              const int32_t* __restrict__ new_ro,
                                                              int32_t* __restrict__ new_nb,
                                                              int32_t* __restrict__ new_eid) {
  const int64_t v = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (v >= n) return;
  int32_t out = new_ro[v];
  const int32_t end = new_ro[v + 1];
  for (int32_t a = row_offsets[v]; a < row_offsets[v + 1] && out < end; ++a) {
    const int32_t e = edge_ids[a];
    if (active[e] != 0) {
      new_nb[out] = neighbors[a];
      new_eid[out] = e;
      ++out;
    }
  }
}

}  // namespace

void DeviceCsr::Allocate(size_t n, size_t m, size_t adjacency, size_t max_commits, int device) {
  if (n >= (size_t{1} << 31) || m >= (size_t{1} << 31) || adjacency >= (size_t{1} << 31) ||
      max_commits >= (size_t{1} << 31) - 1) {
    throw std::invalid_argument("device CSR exceeds int32 ids");
  }
  n_ = n;
  m_ = m;
  adjacency_ = std::max<size_t>(adjacency, 1);
  max_commits_ = max_commits;
  cur_ = 0;
  active_.Allocate(std::max<size_t>(m, 1), device);
  for (int b = 0; b < 2; ++b) {
    ro_[b].Allocate(n + 1, device);
    nb_[b].Allocate(adjacency_, device);
    eid_[b].Allocate(adjacency_, device);
  }
  counts_.Allocate(std::max<size_t>(n, 1), device);
  chunk_sums_.Allocate(std::max<size_t>((n + kChunk - 1) / kChunk, 1), device);
  commits_.Allocate(max_commits + 1, device);
  h_commits_.Reserve(max_commits + 1);
  h_commits_.get()[0] = 0;
}

void DeviceCsr::Upload(const int32_t* active, const int32_t* row_offsets, const int32_t* neighbors,
                       const int32_t* edge_ids, size_t adjacency, hipStream_t stream) {
  if (adjacency > adjacency_) throw std::logic_error("device CSR upload beyond its capacity");
  active_.Upload(active, m_, stream);
  ro_[cur_].Upload(row_offsets, n_ + 1, stream);
  nb_[cur_].Upload(neighbors, adjacency, stream);
  eid_[cur_].Upload(edge_ids, adjacency, stream);
}

void DeviceCsr::UploadCommits(hipStream_t stream) {
  const int32_t count = h_commits_.get()[0];
  if (count < 0 || static_cast<size_t>(count) > max_commits_) {
    throw std::logic_error("device CSR commit list beyond its capacity");
  }
  commits_.Upload(h_commits_.get(), static_cast<size_t>(count) + 1, stream);
}

void DeviceCsr::EnqueueCompact(hipStream_t stream) {
  const int nxt = cur_ ^ 1;
  const int32_t n = static_cast<int32_t>(n_), m = static_cast<int32_t>(m_);
  const uint32_t commit_blocks =
      static_cast<uint32_t>(std::min<size_t>((max_commits_ + kThreads - 1) / kThreads + 1, 1024));
  hipLaunchKernelGGL(DeactivateKernel, dim3(commit_blocks), dim3(kThreads), 0, stream, commits_.get(),
                     static_cast<int32_t>(max_commits_), active_.get(), m);
  CheckHip(hipGetLastError(), "deactivate launch");
  if (n_ == 0) {
    cur_ = nxt;
    return;
  }
  const uint32_t chunks = static_cast<uint32_t>((n_ + kChunk - 1) / kChunk);
  hipLaunchKernelGGL(LiveDegreeKernel, dim3(chunks), dim3(kThreads), 0, stream, ro_[cur_].get(),
                     eid_[cur_].get(), active_.get(), n, counts_.get(), chunk_sums_.get());
  CheckHip(hipGetLastError(), "live degree launch");
  hipLaunchKernelGGL(OffsetScanKernel, dim3(chunks), dim3(kThreads), 0, stream, counts_.get(), n,
                     chunk_sums_.get(), ro_[nxt].get());
  CheckHip(hipGetLastError(), "offset scan launch");
  const uint32_t blocks = static_cast<uint32_t>((n_ + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(ScatterLiveKernel, dim3(blocks), dim3(kThreads), 0, stream, ro_[cur_].get(),
                     nb_[cur_].get(), eid_[cur_].get(), active_.get(), n, ro_[nxt].get(), nb_[nxt].get(),
                     eid_[nxt].get());
  CheckHip(hipGetLastError(), "scatter launch");
  cur_ = nxt;
}

void DeviceCsr::Download(int32_t* row_offsets, int32_t* neighbors, int32_t* edge_ids, size_t capacity,
                         hipStream_t stream) const {
  ro_[cur_].Download(row_offsets, n_ + 1, stream);
  CheckHip(hipStreamSynchronize(stream), "hipStreamSynchronize");
  const size_t a = static_cast<size_t>(row_offsets[n_]);
  if (a > capacity || a > adjacency_) throw std::logic_error("device CSR larger than the buffer");
  nb_[cur_].Download(neighbors, a, stream);
  eid_[cur_].Download(edge_ids, a, stream);
  CheckHip(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

void DeviceCsr::DownloadActive(int32_t* active, hipStream_t stream) const {
  active_.Download(active, m_, stream);
  CheckHip(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

}  // namespace ee
