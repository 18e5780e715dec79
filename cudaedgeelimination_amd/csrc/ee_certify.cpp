// Host certification of device k-opt cost matrices (see ee_certify.h).
//
// Cell rules (KOptCostTaskCpuScorer::Score / ResolvePairCost,
// src/cpu/kopt_search.cpp:983-1031): the template's k added edges are
// visited in ascending order of their first port; an added edge is the
// canonical node pair of its two ports, and the cell is INT64_MAX when an
// added edge is a self loop, equals a deleted edge (canonical pair of ports
// 2e, 2e + 1), equals an earlier added edge, or the sum overflows; otherwise
// it is the sum of the k exact distances.
//
// Which of these rules fire depends only on which ports of the row hold the
// same node, so a row's node pattern (each port labelled by the first port
// holding its node) decides the invalid templates of every row with that
// pattern; the certifier caches those verdicts per pattern within a chunk of
// rows and sums the row's exact port-pair distances for the others.
#include "ee_certify.h"

#include <sched.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <exception>
#include <limits>
#include <stdexcept>

#include "ee_exact.h"
#include "ee_runtime.h"

namespace ee {

// ------------------------------------------------------------- HostPool
namespace {
using PoolClock = std::chrono::steady_clock;
// how long an idle worker, and a caller waiting for the last iterations,
// poll before blocking: the pipeline certifies a block every few hundred
// microseconds, and a wake-up through the condition variable costs tens
constexpr auto kPoolSpin = std::chrono::microseconds(100);
}  // namespace

HostPool::HostPool(unsigned threads) {
  if (threads == 0) {
    cpu_set_t set;
    CPU_ZERO(&set);
    threads = sched_getaffinity(0, sizeof(set), &set) == 0 ? static_cast<unsigned>(CPU_COUNT(&set)) : 1u;
  }
  threads = std::min(std::max(threads, 1u), 64u);
  workers_.reserve(threads - 1);
  for (unsigned t = 1; t < threads; ++t) {
    workers_.emplace_back([this] {
      uint64_t seen = 0;
      for (;;) {
        const auto idle = PoolClock::now();
        while (generation_.load(std::memory_order_acquire) == seen && !stop_.load(std::memory_order_acquire) &&
               PoolClock::now() - idle < kPoolSpin) {
          std::this_thread::yield();
        }
        std::shared_ptr<Job> job;
        {
          std::unique_lock<std::mutex> lk(mu_);
          wake_.wait(lk, [&] { return stop_.load() || generation_.load() != seen; });
          if (stop_.load()) return;
          seen = generation_.load();
          job = job_;
          busy_.fetch_add(1);
        }
        Drain(*job);
        std::lock_guard<std::mutex> lk(mu_);
        if (busy_.fetch_sub(1) == 1) done_.notify_all();
      }
    });
  }
}

HostPool::~HostPool() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_.store(true);
  }
  wake_.notify_all();
  for (std::thread& t : workers_) t.join();
}

void HostPool::Drain(Job& job) {
  for (;;) {
    const size_t i = job.next.fetch_add(1, std::memory_order_relaxed);
    if (i >= job.n) return;
    try {
      (*job.body)(i);
    } catch (...) {
      std::lock_guard<std::mutex> lk(mu_);
      if (!job.error) job.error = std::current_exception();
    }
  }
}

void HostPool::Run(size_t n, const std::function<void(size_t)>& body) {
  if (n == 0) return;
  if (workers_.empty() || n == 1) {
    // every iteration runs, as on the workers; the first exception follows
    std::exception_ptr error;
    for (size_t i = 0; i < n; ++i) {
      try {
        body(i);
      } catch (...) {
        if (!error) error = std::current_exception();
      }
    }
    if (error) std::rethrow_exception(error);
    return;
  }
  auto job = std::make_shared<Job>();
  job->body = &body;
  job->n = n;
  {
    std::lock_guard<std::mutex> lk(mu_);
    job_ = job;
    busy_.fetch_add(1);  // the caller
    generation_.fetch_add(1);
  }
  wake_.notify_all();
  Drain(*job);
  {
    std::lock_guard<std::mutex> lk(mu_);
    busy_.fetch_sub(1);
  }
  // every iteration has been claimed; wait for the workers still running one
  const auto start = PoolClock::now();
  while (busy_.load(std::memory_order_acquire) != 0 && PoolClock::now() - start < kPoolSpin) {
    std::this_thread::yield();
  }
  std::exception_ptr error;
  {
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [&] { return busy_.load() == 0; });
    error = job->error;
  }
  if (error) std::rethrow_exception(error);
}

// ---------------------------------------------------------- CertifyPlan
namespace {
std::atomic<uint64_t> g_plan_serial{0};
}  // namespace

CertifyPlan MakeCertifyPlan(uint32_t k, const ee_endpoint_matching* templates, uint32_t count) {
  if (k < 3 || k > 5) throw std::invalid_argument("k-opt certification: k must lie in [3, 5]");
  if (templates == nullptr || count == 0 || count > 945) {
    throw std::invalid_argument("k-opt certification: the reconnect table is empty or too large");
  }
  CertifyPlan plan;
  plan.serial = ++g_plan_serial;
  plan.k = k;
  plan.templates = count;
  plan.tpl.resize(static_cast<size_t>(count) * k);
  int index[10][10];
  for (auto& row : index) std::fill(row, row + 10, -1);
  const uint32_t ports = 2 * k;
  for (uint32_t t = 0; t < count; ++t) {
    const ee_endpoint_matching& m = templates[t];
    if (m.endpoint_count != ports) throw std::invalid_argument("k-opt certification: template endpoint count != 2k");
    uint32_t added = 0;
    for (uint32_t p = 0; p < ports; ++p) {
      const uint32_t q = m.mate[p];
      if (q >= ports || q == p || m.mate[q] != p) {
        throw std::invalid_argument("k-opt certification: template is not a perfect matching");
      }
      if (q < p) continue;
      if (index[p][q] < 0) {
        index[p][q] = static_cast<int>(plan.pairs);
        plan.pair_p[plan.pairs] = static_cast<uint8_t>(p);
        plan.pair_q[plan.pairs] = static_cast<uint8_t>(q);
        ++plan.pairs;
      }
      plan.tpl[static_cast<size_t>(t) * k + added++] = static_cast<uint8_t>(index[p][q]);
    }
  }
  return plan;
}

// --------------------------------------------------------- CertifyCells
namespace {

constexpr int64_t kInvalidCell = std::numeric_limits<int64_t>::max();
constexpr uint32_t kChunkRows = 64;
constexpr uint32_t kPatternSlots = 64;
constexpr uint32_t kMaxWords = (945 + 63) / 64;

// The invalid-template verdicts of one node pattern under one plan.
struct PatternEntry {
  uint64_t serial = 0;  // CertifyPlan::serial; 0 = empty
  uint64_t key = 0;
  uint64_t invalid[kMaxWords];
};

// Each thread keeps the verdicts of the patterns it met last, across rows,
// blocks and calls (a row's verdicts cost about T x k^2 steps, its cells T x k).
thread_local PatternEntry t_patterns[kPatternSlots];

// Port labels: each port labelled by the first port holding its node.
uint64_t PatternKey(const int32_t* node, uint32_t ports, uint8_t* label) {
  uint64_t key = 0;
  for (uint32_t p = 0; p < ports; ++p) {
    uint32_t l = p;
    for (uint32_t q = 0; q < p; ++q) {
      if (node[q] == node[p]) {
        l = q;
        break;
      }
    }
    label[p] = static_cast<uint8_t>(l);
    key = key << 4 | l;
  }
  return key;
}

void BuildVerdicts(const CertifyPlan& plan, const uint8_t* label, uint64_t* invalid) {
  const uint32_t k = plan.k, words = (plan.templates + 63) / 64;
  std::fill(invalid, invalid + words, uint64_t{0});
  // the deleted edges as label pairs (lo << 4 | hi)
  uint32_t deleted[5];
  for (uint32_t e = 0; e < k; ++e) {
    const uint32_t a = label[2 * e], b = label[2 * e + 1];
    deleted[e] = std::min(a, b) << 4 | std::max(a, b);
  }
  for (uint32_t t = 0; t < plan.templates; ++t) {
    const uint8_t* pairs = &plan.tpl[static_cast<size_t>(t) * k];
    uint32_t added[5];
    bool bad = false;
    for (uint32_t j = 0; j < k && !bad; ++j) {
      const uint32_t a = label[plan.pair_p[pairs[j]]], b = label[plan.pair_q[pairs[j]]];
      const uint32_t edge = std::min(a, b) << 4 | std::max(a, b);
      bad = a == b;  // self loop
      for (uint32_t e = 0; e < k && !bad; ++e) bad = edge == deleted[e];
      for (uint32_t i = 0; i < j && !bad; ++i) bad = edge == added[i];
      added[j] = edge;
    }
    if (bad) invalid[t / 64] |= uint64_t{1} << (t % 64);
  }
}

}  // namespace

namespace {

// Distances at most this large cannot overflow a sum of five.
constexpr int64_t kSafeDistance = int64_t{1} << 60;

// One row's cells from its pair distances, compared with the given cells,
// with the row's improving bits; returns the first mismatching template or
// T.  `plain` rows (every distance in [0, kSafeDistance]) cannot overflow,
// so their cells need no per-addition check.
template <uint32_t K>
uint32_t CertifyRow(const CertifyPlan& plan, const int64_t* dist, bool plain, const uint64_t* invalid,
                    const int64_t* row, int64_t deleted_cost, uint32_t* words, uint64_t* improving) {
  const uint32_t T = plan.templates;
  const uint8_t* tpl = plan.tpl.data();
  uint64_t count = 0;
  bool differs = false;
  for (uint32_t w = 0; w * 32 < T; ++w) {
    const uint32_t t_end = std::min(T, w * 32 + 32);
    uint32_t bits = 0;
    for (uint32_t t = w * 32; t < t_end; ++t) {
      const uint8_t* pairs = tpl + static_cast<size_t>(t) * K;
      int64_t sum = 0;
      bool overflow = false;
      if (plain) {
#pragma unroll
        for (uint32_t j = 0; j < K; ++j) sum += dist[pairs[j]];
      } else {
        for (uint32_t j = 0; j < K; ++j) {
          const int64_t d = dist[pairs[j]];
          overflow |= d < 0 || __builtin_add_overflow(sum, d, &sum);
        }
      }
      const bool bad = overflow || ((invalid[t / 64] >> (t % 64)) & 1u) != 0;
      const int64_t cell = bad ? kInvalidCell : sum;
      differs |= cell != row[t];
      bits |= static_cast<uint32_t>(cell < deleted_cost) << (t % 32);
    }
    count += static_cast<uint64_t>(__builtin_popcount(bits));
    if (words) words[w] = bits;
  }
  *improving += count;
  if (!differs) return T;
  // the rare row with a difference: find its first cell again, with checks
  for (uint32_t t = 0; t < T; ++t) {
    const uint8_t* pairs = tpl + static_cast<size_t>(t) * K;
    int64_t sum = 0;
    bool overflow = false;
    for (uint32_t j = 0; j < K; ++j) {
      const int64_t d = dist[pairs[j]];
      overflow |= d < 0 || __builtin_add_overflow(sum, d, &sum);
    }
    const bool bad = overflow || ((invalid[t / 64] >> (t % 64)) & 1u) != 0;
    if ((bad ? kInvalidCell : sum) != row[t]) return t;
  }
  return T;
}

}  // namespace

CertifyResult CertifyCells(HostPool& pool, const CertifyPlan& plan, const int64_t* x, const int64_t* y,
                           int32_t n, uint8_t distance_type, bool near, const ee_kopt_task* tasks, size_t rows,
                           const int64_t* cells, uint32_t* improving, const std::function<void()>* side) {
  static_cast<void>(n);
  const uint32_t k = plan.k, ports = 2 * k, T = plan.templates;
  const uint32_t mask_words = (T + 31) / 32;
  const size_t chunks = (rows + kChunkRows - 1) / kChunkRows;
  std::atomic<uint64_t> mismatch{UINT64_MAX};
  std::atomic<uint64_t> improving_total{0};
  // the side job is iteration 0, so it starts first and overlaps the rows
  const size_t lead = side != nullptr ? 1 : 0;
  pool.Run(chunks + lead, [&](size_t index) {
    if (index < lead) {
      (*side)();
      return;
    }
    const size_t chunk = index - lead;
    uint64_t improving_here = 0;
    const size_t r0 = chunk * kChunkRows, r1 = std::min(rows, r0 + kChunkRows);
    for (size_t r = r0; r < r1; ++r) {
      if (r * T >= mismatch.load(std::memory_order_relaxed)) break;
      const ee_kopt_task& task = tasks[r];
      uint8_t label[10];
      const uint64_t key = PatternKey(task.port_nodes, ports, label);
      PatternEntry& entry = t_patterns[((key * 0x9E3779B97F4A7C15ull) >> 58) % kPatternSlots];
      if (entry.serial != plan.serial || entry.key != key) {
        BuildVerdicts(plan, label, entry.invalid);
        entry.serial = plan.serial;
        entry.key = key;
      }
      int64_t dist[45];
      bool plain = true;
      for (uint32_t i = 0; i < plan.pairs; ++i) {
        const int32_t a = task.port_nodes[plan.pair_p[i]], b = task.port_nodes[plan.pair_q[i]];
        // both coordinate extents below 2^25: the FP64 form, exact there
        dist[i] = near ? ExactDistanceNear(static_cast<int32_t>(x[a] - x[b]), static_cast<int32_t>(y[a] - y[b]),
                                           distance_type)
                       : ExactDistance(x[a], y[a], x[b], y[b], distance_type);
        plain &= dist[i] >= 0 && dist[i] <= kSafeDistance;
      }
      const int64_t* row = cells + r * T;
      uint32_t* words = improving ? improving + r * mask_words : nullptr;
      uint32_t first = T;
      switch (k) {
        case 3:
          first = CertifyRow<3>(plan, dist, plain, entry.invalid, row, task.deleted_cost, words, &improving_here);
          break;
        case 4:
          first = CertifyRow<4>(plan, dist, plain, entry.invalid, row, task.deleted_cost, words, &improving_here);
          break;
        default:
          first = CertifyRow<5>(plan, dist, plain, entry.invalid, row, task.deleted_cost, words, &improving_here);
          break;
      }
      if (first != T) {
        const uint64_t at = r * T + first;
        uint64_t seen = mismatch.load(std::memory_order_relaxed);
        while (at < seen && !mismatch.compare_exchange_weak(seen, at)) {
        }
        break;
      }
    }
    improving_total.fetch_add(improving_here, std::memory_order_relaxed);
  });
  CertifyResult result;
  result.mismatch = mismatch.load();
  result.improving = improving_total.load();
  return result;
}

namespace {

HostPool& SharedPool() {
  static HostPool pool(0);
  return pool;
}

}  // namespace
}  // namespace ee

extern "C" ee_status ee_kopt_certify(const int64_t* x, const int64_t* y, int32_t dimension,
                                     uint8_t distance_type, uint32_t k,
                                     const ee_endpoint_matching* templates, uint32_t template_count,
                                     const ee_kopt_task* tasks, size_t task_count, const int64_t* cells,
                                     uint32_t* improving, uint64_t* mismatch, uint64_t* improving_count) {
  try {
    if (mismatch) *mismatch = UINT64_MAX;
    if (improving_count) *improving_count = 0;
    if (task_count == 0) return EE_OK;
    if (x == nullptr || y == nullptr || dimension <= 0 || tasks == nullptr || cells == nullptr) {
      throw std::invalid_argument("k-opt certification: NULL argument");
    }
    if (distance_type != EE_EUC_2D && distance_type != EE_CEIL_2D) {
      throw std::invalid_argument("k-opt certification: unknown distance type");
    }
    const ee::CertifyPlan plan = ee::MakeCertifyPlan(k, templates, template_count);
    ee::ValidateTasks(dimension, k, tasks, task_count);
    if (!ee::DistanceSafe(x, y, dimension)) {
      throw std::invalid_argument("k-opt certification needs distance-safe integer coordinates");
    }
    static std::mutex serial;  // one certification at a time on the shared pool
    std::lock_guard<std::mutex> lk(serial);
    const ee::CertifyResult r =
        ee::CertifyCells(ee::SharedPool(), plan, x, y, dimension, distance_type, ee::NearExtent(x, y, dimension),
                         tasks, task_count, cells, improving);
    if (mismatch) *mismatch = r.mismatch;
    if (improving_count) *improving_count = r.improving;
    if (r.mismatch != UINT64_MAX) {
      throw std::runtime_error("device k-opt cost failed CPU exact-matrix certification, cell=" +
                               std::to_string(r.mismatch));
    }
    return EE_OK;
  } catch (...) {
    return ee::StatusFromException();
  }
}
