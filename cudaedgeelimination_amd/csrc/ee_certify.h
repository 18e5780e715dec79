// Host certification of device k-opt cost matrices.
//
// The reference accepts a device cost matrix only after comparing it cell by
// cell with an independent exact CPU matrix of the same task rows
// (EvaluateKOptTemplateCosts, src/cpu/kopt_search.cpp:1455-1475), and its leaf
// consumer rechecks only the templates whose cell is below the row's deleted
// cost (TryReconnectFromCostRow, kopt_search.cpp:677-734).  CertifyCells does
// both in one pass over a block: it recomputes every cell on the host by the
// reference CPU scorer's rules (KOptCostTaskCpuScorer, kopt_search.cpp:962-1042),
// compares it with the device cell, and sets the row's improving-template bit.
#pragma once

#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "ee_hip.h"

namespace ee {

// A fixed set of host threads that run the iterations of one loop at a time;
// the calling thread works too.  Thread start-up is paid once, not per loop.
class HostPool {
 public:
  // threads = 0: the cores this process may run on (at most 64).
  explicit HostPool(unsigned threads);
  ~HostPool();
  HostPool(const HostPool&) = delete;
  HostPool& operator=(const HostPool&) = delete;

  unsigned size() const { return static_cast<unsigned>(workers_.size()) + 1; }
  // Runs body(i) for every i in [0, n), each exactly once; rethrows the first
  // exception a body threw once every iteration is done.  One Run at a time.
  void Run(size_t n, const std::function<void(size_t)>& body);

 private:
  // One Run: its iterations are claimed from `next`; a worker holds the job
  // it registered for, so a late worker never claims another Run's index.
  struct Job {
    const std::function<void(size_t)>* body = nullptr;
    size_t n = 0;
    std::atomic<size_t> next{0};
    std::exception_ptr error;  // under the pool's mutex
  };
  void Drain(Job& job);

  std::vector<std::thread> workers_;
  std::mutex mu_;
  std::condition_variable wake_, done_;
  std::atomic<uint64_t> generation_{0};  // written under mu_
  std::atomic<bool> stop_{false};        // written under mu_
  std::shared_ptr<Job> job_;             // the latest Run's; under mu_
  std::atomic<size_t> busy_{0};          // written under mu_
};

// A reconnect table as the CPU scorer walks it: per template, its k added
// edges as indices into the table's distinct port pairs, in ascending order
// of the first port (BuildKOptCpuReconnectPlans, kopt_search.cpp:815-842).
struct CertifyPlan {
  uint64_t serial = 0;  // unique per plan made, keys the certifier's verdict caches
  uint32_t k = 0;
  uint32_t templates = 0;
  uint32_t pairs = 0;
  uint8_t pair_p[45] = {};
  uint8_t pair_q[45] = {};
  std::vector<uint8_t> tpl;  // templates x k
};

// Throws std::invalid_argument unless every template is a perfect matching of
// the 2k ports (at most 945 templates).
CertifyPlan MakeCertifyPlan(uint32_t k, const ee_endpoint_matching* templates, uint32_t count);

struct CertifyResult {
  uint64_t mismatch = UINT64_MAX;  // first (row-major) cell differing from the CPU matrix
  uint64_t improving = 0;          // cells below their row's deleted_cost
};

// Certifies rows [0, rows) of `cells` (rows x plan.templates, row-major)
// against the exact CPU matrix of `tasks` on the n coordinates x, y (distance
// safe; the rows pass ValidateKOptCostTask); near is NearExtent(x, y, n),
// which callers that certify many blocks of one graph compute once.
// improving (may be null) receives
// per row ceil(templates / 32) words, bit j set iff the row's certified cell
// j is below its deleted_cost.  The words of rows from the first mismatching
// row on are unspecified.  side (may be null) runs once on one of the pool's
// threads alongside the rows; an exception it throws is rethrown.
CertifyResult CertifyCells(HostPool& pool, const CertifyPlan& plan, const int64_t* x, const int64_t* y,
                           int32_t n, uint8_t distance_type, bool near, const ee_kopt_task* tasks, size_t rows,
                           const int64_t* cells, uint32_t* improving,
                           const std::function<void()>* side = nullptr);

}  // namespace ee
