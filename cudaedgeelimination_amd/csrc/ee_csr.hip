// Commit on the device: clear the active flags of committed edge slots and
// compact the CSR (see ee_csr.h).  Four launches, every one spread over the
// adjacency entries (or nodes) rather than over rows, so even a small graph
// fills many workgroups: deactivate; the live mask of every 64 entries (one
// ballot) with per-block counts; a scan of the block counts; the scatter of
// the live entries and of the new row offsets.
#include "ee_csr.h"

namespace ee {
namespace {

constexpr uint32_t kThreads = 256;  // = entries per block of the count and scatter passes
constexpr uint32_t kWavesPerBlock = kThreads / 64;
constexpr uint32_t kScanThreads = 1024;

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

// masks[w] = the live entries among entries [64 w, 64 w + 64) (bit j: entry
// 64 w + j's edge is active); block_counts[b] = live entries of block b.  The
// grid covers the buffer's capacity; the CSR holds row_offsets[n] entries,
// and every entry past them (stale ids) counts as not live and is not read.
__global__ __launch_bounds__(kThreads) void LiveMaskKernel(const int32_t* __restrict__ row_offsets, int32_t n,
                                                           const int32_t* __restrict__ edge_ids,
                                                           const int32_t* __restrict__ active,
                                                           unsigned long long* __restrict__ masks,
                                                           uint32_t* __restrict__ block_counts) {
  __shared__ uint32_t wave_counts[kWavesPerBlock];
  const int32_t entries = row_offsets[n];
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  const int64_t a = static_cast<int64_t>(blockIdx.x) * kThreads + t;
  const bool live = a < entries && active[edge_ids[a]] != 0;
  const unsigned long long m = __ballot(live);
  if (lane == 0) {
    masks[static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + wave] = m;
    wave_counts[wave] = static_cast<uint32_t>(__popcll(m));
  }
  __syncthreads();
  if (t == 0) {
    uint32_t total = 0;
    for (uint32_t w = 0; w < kWavesPerBlock; ++w) total += wave_counts[w];
    block_counts[blockIdx.x] = total;
  }
}

// offsets[b] = block_counts[0] + ... + block_counts[b - 1] for b <= blocks
// (offsets[blocks] = all live entries): one workgroup, 4 counts per thread
// per round, wavefront shuffles and wavefront totals in LDS.
__global__ __launch_bounds__(kScanThreads) void BlockScanKernel(const uint32_t* __restrict__ counts, int32_t blocks,
                                                                uint32_t* __restrict__ offsets) {
  __shared__ uint32_t wave_sums[kScanThreads / 64];
  __shared__ uint32_t carry;
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  if (t == 0) carry = 0;
  __syncthreads();
  for (int64_t base = 0; base < blocks; base += 4 * kScanThreads) {
    uint32_t v[4], own = 0;
    const int64_t first = base + 4 * static_cast<int64_t>(t);
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) {
      v[j] = first + j < blocks ? counts[first + j] : 0u;
      own += v[j];
    }
    uint32_t inc = own;
    for (uint32_t d = 1; d < 64; d <<= 1) {
      const uint32_t o = __shfl_up(inc, d, 64);
      if (lane >= d) inc += o;
    }
    if (lane == 63) wave_sums[wave] = inc;
    __syncthreads();
    uint32_t run = carry + inc - own;
    for (uint32_t w = 0; w < wave; ++w) run += wave_sums[w];
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) {
      if (first + j < blocks) offsets[first + j] = run;
      run += v[j];
    }
    __syncthreads();  // every thread has read carry and wave_sums
    if (t == kScanThreads - 1) carry = run;
    __syncthreads();
  }
  if (t == 0) offsets[blocks] = carry;
}

// The live entries before entry a (a <= entries).
__device__ __forceinline__ uint32_t LiveBefore(int64_t a, int32_t entries,
                                               const unsigned long long* __restrict__ masks,
                                               const uint32_t* __restrict__ offsets) {
  const int64_t block = a / kThreads;
  if (a >= entries) return offsets[(entries + kThreads - 1) / kThreads];
  uint32_t before = offsets[block];
  const int64_t w0 = block * kWavesPerBlock, w = a / 64;
  for (int64_t q = w0; q < w; ++q) before += static_cast<uint32_t>(__popcll(masks[q]));
  const uint32_t bit = static_cast<uint32_t>(a & 63);
  return before + static_cast<uint32_t>(__popcll(masks[w] & ((1ull << bit) - 1ull)));
}

// Blocks [0, entry_blocks): live entry a to position LiveBefore(a); the
// rest: new_ro[v] = LiveBefore(row_offsets[v]) for v <= n.  Removing
// entries keeps each row's order.
__global__ __launch_bounds__(kThreads) void CompactKernel(
    const int32_t* __restrict__ row_offsets, const int32_t* __restrict__ neighbors,
    const int32_t* __restrict__ edge_ids, int32_t n, uint32_t entry_blocks,
    const unsigned long long* __restrict__ masks, const uint32_t* __restrict__ offsets,
    int32_t* __restrict__ new_ro, int32_t* __restrict__ new_nb, int32_t* __restrict__ new_eid) {
  const int32_t entries = row_offsets[n];
  const uint32_t t = threadIdx.x;
  if (blockIdx.x < entry_blocks) {
    const int64_t a = static_cast<int64_t>(blockIdx.x) * kThreads + t;
    if (a >= entries) return;
    if (((masks[a / 64] >> (a & 63)) & 1ull) == 0) return;
    const uint32_t at = LiveBefore(a, entries, masks, offsets);
    new_nb[at] = neighbors[a];
    new_eid[at] = edge_ids[a];
    return;
  }
  const int64_t v = static_cast<int64_t>(blockIdx.x - entry_blocks) * kThreads + t;
  if (v > n) return;
  new_ro[v] = static_cast<int32_t>(LiveBefore(row_offsets[v], entries, masks, offsets));
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
  const size_t blocks = (adjacency_ + kThreads - 1) / kThreads;
  masks_.Allocate(blocks * kWavesPerBlock, device);
  block_counts_.Allocate(blocks, device);
  block_offsets_.Allocate(blocks + 1, device);
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
  // the passes are sized for the capacity; the kernels read the current
  // CSR's entry count (its last row end) on the device, so one captured
  // sequence serves every commit
  const uint32_t entry_blocks = static_cast<uint32_t>((adjacency_ + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(LiveMaskKernel, dim3(entry_blocks), dim3(kThreads), 0, stream, ro_[cur_].get(), n,
                     eid_[cur_].get(), active_.get(), masks_.get(), block_counts_.get());
  CheckHip(hipGetLastError(), "live mask launch");
  hipLaunchKernelGGL(BlockScanKernel, dim3(1), dim3(kScanThreads), 0, stream, block_counts_.get(),
                     static_cast<int32_t>(entry_blocks), block_offsets_.get());
  CheckHip(hipGetLastError(), "block scan launch");
  const uint32_t node_blocks = static_cast<uint32_t>((n_ + 1 + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(CompactKernel, dim3(entry_blocks + node_blocks), dim3(kThreads), 0, stream,
                     ro_[cur_].get(), nb_[cur_].get(), eid_[cur_].get(), n, entry_blocks,
                     masks_.get(), block_offsets_.get(), ro_[nxt].get(), nb_[nxt].get(), eid_[nxt].get());
  CheckHip(hipGetLastError(), "compact launch");
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
