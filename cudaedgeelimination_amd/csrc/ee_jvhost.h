// Host side of one resident JV elimination run: the snapshot, the CPU
// re-verification, the deterministic commit and the V1 proof of the epoch
// loop of RunJvElimination (src/cpu/jv_elimination.cpp:250-368).  The device
// propose step is supplied by the caller (a single-graph session or a fleet of
// graphs sharing one launch per epoch) as a per-edge witness array.
#pragma once

#include <cstdint>
#include <functional>
#include <future>
#include <string>
#include <vector>

#include "ee_hip.h"

namespace ee {

// Runs body(i) for every i in [0, n) on up to `threads` host threads (0: the
// cores this process may run on, at most 32); each index runs exactly once.
// The first exception a body throws is rethrown once every thread is done.
void ParallelFor(size_t n, const std::function<void(size_t)>& body, unsigned threads = 0);

struct JvHost {
  int32_t n = 0, m = 0;
  uint8_t type = 0;
  bool near = false;  // coordinate extents below 2^25: FP64 distances on the device
  std::vector<int64_t> x, y, weight;
  std::vector<int32_t> u, v;
  std::vector<uint8_t> active;
  std::vector<int32_t> row_offsets, neighbors, edge_ids;

  // Copies and validates a graph view (endpoints in range, weights in
  // [0, 2^62], distance-safe coordinates) and builds the CSR.
  void Load(const ee_graph_view* g);

  int32_t Degree(int32_t vtx) const { return row_offsets[vtx + 1] - row_offsets[vtx]; }
  int64_t Dist(int32_t a, int32_t b) const;
  void RebuildCsr();
  void DropInactiveFromCsr();
  // With async_hash set, the content hash of every epoch's snapshot is
  // computed on a helper thread from a copy of the active flags, so it never
  // waits on the critical path; the proof records get their hashes when the
  // run ends.  Without it, the hash is computed when the epoch commits
  // (before the snapshot changes), on the calling thread.
  bool async_hash = false;
  // Host threads of the re-verification of a large epoch (each candidate is
  // checked on its own; 1 = the calling thread only, 0 = every core).
  unsigned verify_threads = 1;
  // When set, asked for the hash of each new snapshot generation first: it
  // returns true with a future of that hash (computed elsewhere, e.g. on the
  // device), or false to leave the generation to the host.
  std::function<bool(std::shared_future<uint64_t>*)> hash_source;
  // FNV-1a state after the header and coordinates (the edge part follows).
  uint64_t prefix_hash() const { return prefix_hash_; }
  // IsJvWitness (jv_elimination.cpp:86-117) with 128-bit sums.
  bool Verify(int32_t e, int32_t c) const;

  // ---- one run
  void StartRun();
  // Active edge count at the start of `epoch`; takes the snapshot hash job.
  void BeginEpoch(uint32_t epoch);
  // Re-verifies every proposal of witnesses[0..m) (witness or -1), throws if
  // any is rejected (nothing is committed), and runs the deterministic
  // commit rule: sort by (u, v, method, witness, edge), one per edge, running
  // degree gate > 2 (src/cpu/elimination_commit.cpp:13-75).  Returns the
  // edge slots it commits, in commit order; the snapshot is unchanged.
  const std::vector<int32_t>& VerifyAndGate(const int32_t* witnesses, double propose_ms);
  // Applies the commit of the last VerifyAndGate to the snapshot (active
  // flags, CSR), records it for the proof and ends the epoch; returns the
  // committed count.  The run is done when nothing was committed.
  size_t ApplyCommit();
  size_t FinishEpoch(const int32_t* witnesses, double propose_ms) {
    VerifyAndGate(witnesses, propose_ms);
    return ApplyCommit();
  }
  // Final hash and the V1 proof text with backend name "hip"; total_ms is
  // the run's time up to this call, to which the wait for hashes is added.
  void EndRun(double total_ms);

  bool done = false;
  ee_jv_run_stats stats{};
  std::string proof;

 private:
  struct Record {
    uint32_t epoch;
    size_t snap;  // index into snaps_
    int32_t edge, a, b, witness;
  };
  struct Metrics {
    uint32_t epoch;
    size_t before, proposed, verified, rejected, committed;
    double propose_ms, verify_ms;
  };
  // The content hash of one snapshot generation: a value, or a job computing
  // it from a copy of that generation's active flags.
  struct SnapHash {
    uint64_t generation = 0;
    bool ready = false;
    uint64_t value = 0;
    std::shared_future<uint64_t> job;
    uint64_t Get();
  };
  uint64_t ComputeHash(const uint8_t* active_flags) const;
  // The hash holder of the current generation (reuses the last one when the
  // snapshot has not changed since).
  SnapHash CurrentSnap();

  std::vector<Record> records_;
  std::vector<Metrics> metrics_;
  size_t live_edges_ = 0;  // active edge slots, kept through the run's commits
  // the next CSR of a parallel DropInactiveFromCsr (swapped in; capacity kept)
  std::vector<int32_t> next_rows_, next_neighbors_, next_edge_ids_;
  std::vector<SnapHash> snaps_;  // one per epoch of the run
  Metrics cur_{};
  std::vector<std::pair<int32_t, int32_t>> gated_;  // (edge, witness) of the pending commit
  std::vector<int32_t> gated_edges_;
  uint64_t snap_generation_ = 0;
  // every change of the snapshot bumps the generation
  uint64_t generation_ = 1;
  uint64_t prefix_hash_ = 0;
};

}  // namespace ee
