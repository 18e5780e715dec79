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
  // ContentHash of the current snapshot; recomputed only after a commit
  // changed it (the point prefix is hashed once, at Load).
  uint64_t Hash() const;
  // With async_hash set, a stale snapshot hash is computed on a helper thread
  // from the start of a run or epoch until the commit needs it, overlapping
  // the device propose step (FNV-1a is byte-serial).
  bool async_hash = false;
  // IsJvWitness (jv_elimination.cpp:86-117) with 128-bit sums.
  bool Verify(int32_t e, int32_t c) const;

  // ---- one run
  void StartRun();
  // Active edge count at the start of `epoch`; the snapshot hash is taken
  // when the epoch commits (the snapshot does not change in between).
  void BeginEpoch(uint32_t epoch);
  // Re-verifies every proposal of witnesses[0..m) (witness or -1), throws if
  // any is rejected, commits and records; returns the committed count.  The
  // run is done when nothing was committed.
  size_t FinishEpoch(const int32_t* witnesses, double propose_ms);
  // Final hash and the V1 proof text with backend name "hip".
  void EndRun(double total_ms);

  bool done = false;
  ee_jv_run_stats stats{};
  std::string proof;

 private:
  struct Record {
    uint32_t epoch;
    uint64_t hash;
    int32_t edge, a, b, witness;
  };
  struct Metrics {
    uint32_t epoch;
    size_t before, proposed, verified, rejected, committed;
    double propose_ms, verify_ms;
  };
  void HashAsync();
  uint64_t ComputeHash() const;

  std::vector<Record> records_;
  std::vector<Metrics> metrics_;
  Metrics cur_{};
  uint64_t snap_generation_ = 0;
  bool initial_pending_ = false;
  // every change of the snapshot bumps the generation
  uint64_t generation_ = 1;
  uint64_t prefix_hash_ = 0;
  mutable uint64_t hash_ = 0, hash_generation_ = 0;
  // the hash being computed for snapshot generation pending_generation_
  mutable std::future<uint64_t> pending_hash_;
  mutable uint64_t pending_generation_ = 0;
};

}  // namespace ee
