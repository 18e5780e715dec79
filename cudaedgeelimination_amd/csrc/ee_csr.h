// Device-resident active flags and CSR of a snapshot (or of the union of
// many snapshots), updated in place by commits.
//
// The reference rebuilds the CSR on the host after every commit and uploads
// it again for the next epoch (src/cpu/elimination_commit.cpp:64-72,
// src/common/graph.cpp:230-282, src/cuda/jv_cuda.cu:522-528).  Here a commit
// is a list of edge slots to deactivate: the device clears their active
// flags and compacts the CSR into its second buffer.  Removing entries from
// rows that are ascending by neighbour keeps them ascending, so the result
// equals RebuildCsr over the new active flags (rows are filled in edge-id
// order, which for canonical (u < v, sorted) edges is ascending neighbour
// order, and the host rebuild the session keeps beside it is the same).
#pragma once

#include "ee_runtime.h"

namespace ee {

class DeviceCsr {
 public:
  // n nodes, m edge slots, at most `adjacency` CSR entries, at most
  // `max_commits` slots per commit.
  void Allocate(size_t n, size_t m, size_t adjacency, size_t max_commits, int device);
  // Uploads active flags (0/1 per slot) and a CSR into the current buffer.
  void Upload(const int32_t* active, const int32_t* row_offsets, const int32_t* neighbors,
              const int32_t* edge_ids, size_t adjacency, hipStream_t stream);
  // Pinned staging of a commit: word 0 the count, then the edge slots.
  int32_t* StageCommits() { return h_commits_.get(); }
  // Enqueues the upload of the staged commit (count + slots) on `stream`.
  void UploadCommits(hipStream_t stream);
  // Enqueues: deactivate the slots of the device commit list, then compact
  // the current CSR into the other buffer, which becomes current.  The
  // kernels read the count from device memory, so one captured sequence
  // serves any commit of at most max_commits slots.
  void EnqueueCompact(hipStream_t stream);

  int32_t* active() { return active_.get(); }
  const int32_t* row_offsets() const { return ro_[cur_].get(); }
  const int32_t* neighbors() const { return nb_[cur_].get(); }
  const int32_t* edge_ids() const { return eid_[cur_].get(); }
  int current() const { return cur_; }
  // Selects the current buffer (after replaying a captured compaction).
  void SetCurrent(int buffer) { cur_ = buffer & 1; }
  // Downloads the current CSR (row_offsets [n + 1]; neighbours and edge ids
  // [row_offsets[n]] each) and waits.
  void Download(int32_t* row_offsets, int32_t* neighbors, int32_t* edge_ids, size_t capacity,
                hipStream_t stream) const;
  // Downloads the active flags (m entries) and waits.
  void DownloadActive(int32_t* active, hipStream_t stream) const;
  size_t nodes() const { return n_; }

 private:
  size_t n_ = 0, m_ = 0, adjacency_ = 0, max_commits_ = 0;
  int cur_ = 0;
  DeviceArray<int32_t> active_;
  DeviceArray<int32_t> ro_[2], nb_[2], eid_[2];
  DeviceArray<unsigned long long> masks_;  // live mask per 64 adjacency entries
  DeviceArray<uint32_t> block_counts_, block_offsets_;
  DeviceArray<int32_t> commits_;
  Pinned<int32_t> h_commits_;
};

}  // namespace ee
