// Host side of a resident JV run; see ee_jvhost.h.
#include "ee_jvhost.h"

#include <sched.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>
#include <cstdio>
#include <sstream>
#include <stdexcept>
#include <tuple>
#include <utility>

#include "ee_exact.h"
#include "ee_gen.h"
#include "ee_runtime.h"

namespace ee {
namespace {

void CheckStatus(ee_status s) {
  if (s != EE_OK) throw std::runtime_error(LastError());
}

std::string Hex(uint64_t h) {
  char buf[17];
  std::snprintf(buf, sizeof(buf), "%016llx", static_cast<unsigned long long>(h));
  return std::string(buf);
}

}  // namespace

namespace {

unsigned AffinityCores() {
  cpu_set_t set;
  CPU_ZERO(&set);
  const unsigned n = sched_getaffinity(0, sizeof(set), &set) == 0 ? static_cast<unsigned>(CPU_COUNT(&set)) : 1u;
  return std::max(1u, std::min(n, 32u));
}

// Worker threads kept for the life of the process.  One job runs at a time;
// the caller takes part in it.  Between jobs a worker polls for the next one
// for a short while (yielding), then blocks.
class ForPool {
 public:
  static ForPool& Get() {
    static ForPool* pool = new ForPool(AffinityCores() - 1);  // never destroyed: its threads outlive exit
    return *pool;
  }
  // False when another job holds the pool (the caller then runs serially).
  bool Run(size_t n, const std::function<void(size_t)>& body, unsigned threads) {
    std::unique_lock<std::mutex> job(job_mu_, std::try_to_lock);
    if (!job.owns_lock()) return false;
    const unsigned helpers = std::min<unsigned>(static_cast<unsigned>(workers_), threads - 1);
    {
      std::lock_guard<std::mutex> lk(mu_);
      body_ = &body;
      n_ = n;
      next_.store(0, std::memory_order_relaxed);
      first_ = nullptr;
      wanted_ = helpers;
      joined_ = 0;
      active_ = 0;
      accepting_ = true;
      generation_.fetch_add(1, std::memory_order_release);
    }
    cv_.notify_all();
    Work();
    // every index is claimed: no later helper joins; wait for those working
    std::unique_lock<std::mutex> lk(mu_);
    accepting_ = false;
    done_cv_.wait(lk, [&] { return active_ == 0; });
    body_ = nullptr;
    if (first_) std::rethrow_exception(first_);
    return true;
  }
  static bool InWorker() { return in_worker_; }

 private:
  explicit ForPool(unsigned workers) : workers_(workers) {
    for (unsigned w = 0; w < workers; ++w) std::thread([this] { Loop(); }).detach();
  }
  void Work() {
    for (size_t i = next_.fetch_add(1); i < n_; i = next_.fetch_add(1)) {
      try {
        (*body_)(i);
      } catch (...) {
        std::lock_guard<std::mutex> lk(mu_);
        if (!first_) first_ = std::current_exception();
      }
    }
  }
  void Loop() {
    in_worker_ = true;
    uint64_t seen = generation_.load(std::memory_order_acquire);
    for (;;) {
      // poll a while for the next job (back-to-back epochs), then block
      const auto until = std::chrono::steady_clock::now() + std::chrono::microseconds(kPollMicros);
      while (generation_.load(std::memory_order_acquire) == seen && std::chrono::steady_clock::now() < until) {
        std::this_thread::yield();
      }
      std::unique_lock<std::mutex> lk(mu_);
      cv_.wait(lk, [&] { return generation_.load(std::memory_order_relaxed) != seen; });
      seen = generation_.load(std::memory_order_relaxed);
      if (!accepting_ || joined_ >= wanted_) continue;  // the job is done or has its helpers
      ++joined_;
      ++active_;
      lk.unlock();
      Work();
      lk.lock();
      if (--active_ == 0) done_cv_.notify_one();
    }
  }

  static constexpr int kPollMicros = 100;
  static thread_local bool in_worker_;
  const size_t workers_;
  std::mutex job_mu_;  // one job at a time
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  std::atomic<uint64_t> generation_{0};
  const std::function<void(size_t)>* body_ = nullptr;
  size_t n_ = 0;
  std::atomic<size_t> next_{0};
  unsigned wanted_ = 0, joined_ = 0, active_ = 0;
  bool accepting_ = false;
  std::exception_ptr first_;
};
thread_local bool ForPool::in_worker_ = false;

}  // namespace

void ParallelFor(size_t n, const std::function<void(size_t)>& body, unsigned threads) {
  if (threads == 0) threads = AffinityCores();
  const size_t workers = std::min<size_t>(std::max(threads, 1u), n);
  // serially: one worker's worth, a call from inside a pool job, or a pool
  // busy with another caller's job
  if (workers <= 1 || ForPool::InWorker() || !ForPool::Get().Run(n, body, static_cast<unsigned>(workers))) {
    std::exception_ptr first;
    for (size_t i = 0; i < n; ++i) {
      try {
        body(i);
      } catch (...) {
        if (!first) first = std::current_exception();
      }
    }
    if (first) std::rethrow_exception(first);
  }
}

void JvHost::Load(const ee_graph_view* g) {
  for (SnapHash& h : snaps_) h.Get();  // pending hash jobs read the arrays replaced here
  snaps_.clear();
  n = g->dimension;
  m = g->edge_count;
  type = g->distance_type;
  const size_t nn = static_cast<size_t>(n), mm = static_cast<size_t>(m);
  x.assign(g->x, g->x + nn);
  y.assign(g->y, g->y + nn);
  if (!DistanceSafe(x.data(), y.data(), n)) {
    throw std::runtime_error("HIP JV needs integer coordinates whose squared distances fit uint64");
  }
  near = NearExtent(x.data(), y.data(), n);
  prefix_hash_ = ContentHashPrefix(n, type, x.data(), y.data());
  u.clear();
  v.clear();
  weight.clear();
  active.clear();
  if (mm > 0) {
    u.assign(g->edge_u, g->edge_u + mm);
    v.assign(g->edge_v, g->edge_v + mm);
    weight.assign(g->edge_weight, g->edge_weight + mm);
    active.assign(g->edge_active, g->edge_active + mm);
  }
  for (size_t e = 0; e < mm; ++e) {
    if (u[e] < 0 || u[e] >= n || v[e] < 0 || v[e] >= n) {
      throw std::invalid_argument("JV edge endpoint out of range");
    }
    if (weight[e] < 0 || weight[e] > (int64_t{1} << 62)) {
      throw std::invalid_argument("JV edge weight outside [0, 2^62]");
    }
    active[e] = active[e] ? 1 : 0;
  }
  ++generation_;
  RebuildCsr();
}

// Coordinate extents below 2^25 take the FP64 form (the same distance,
// without the integer root's correction loops).
int64_t JvHost::Dist(int32_t a, int32_t b) const {
  if (near) return ExactDistanceNear(static_cast<int32_t>(x[a] - x[b]), static_cast<int32_t>(y[a] - y[b]), type);
  return ExactDistance(x[a], y[a], x[b], y[b], type);
}

void JvHost::RebuildCsr() {
  row_offsets.assign(static_cast<size_t>(n) + 1, 0);
  int64_t live = 0;
  for (int32_t e = 0; e < m; ++e) live += active[e] != 0;
  neighbors.assign(static_cast<size_t>(2 * live), 0);
  edge_ids.assign(static_cast<size_t>(2 * live), 0);
  CheckStatus(ee_graph_build_csr(n, m, u.data(), v.data(), active.data(), row_offsets.data(),
                                 neighbors.data(), edge_ids.data()));
}

// Removes the entries of edges that are no longer active, row by row.  Rows
// keep their order, so the result equals RebuildCsr on the new flags (rows
// are filled in edge-slot order there too).  Large CSRs are filtered on the
// host cores in blocks of rows: each block counts what it keeps, the block
// offsets follow from a prefix sum, and each block writes its rows into
// fresh arrays, in the same order as one pass over all rows.
void JvHost::DropInactiveFromCsr() {
  constexpr size_t kParallelEntries = 65536, kBlocks = 64;
  const size_t entries = static_cast<size_t>(row_offsets[static_cast<size_t>(n)]);
  if (entries >= kParallelEntries && n >= static_cast<int32_t>(kBlocks)) {
    const size_t rows_per = (static_cast<size_t>(n) + kBlocks - 1) / kBlocks;
    std::vector<size_t> kept(kBlocks + 1, 0);
    ParallelFor(kBlocks, [&](size_t b) {
      const size_t r0 = std::min(static_cast<size_t>(n), b * rows_per);
      const size_t r1 = std::min(static_cast<size_t>(n), r0 + rows_per);
      size_t k = 0;
      for (size_t s = static_cast<size_t>(row_offsets[r0]); s < static_cast<size_t>(row_offsets[r1]); ++s) {
        k += active[static_cast<size_t>(edge_ids[s])] != 0;
      }
      kept[b + 1] = k;
    });
    for (size_t b = 0; b < kBlocks; ++b) kept[b + 1] += kept[b];
    next_rows_.resize(static_cast<size_t>(n) + 1);
    next_neighbors_.resize(kept[kBlocks]);
    next_edge_ids_.resize(kept[kBlocks]);
    ParallelFor(kBlocks, [&](size_t b) {
      const size_t r0 = std::min(static_cast<size_t>(n), b * rows_per);
      const size_t r1 = std::min(static_cast<size_t>(n), r0 + rows_per);
      size_t out = kept[b];
      for (size_t vtx = r0; vtx < r1; ++vtx) {
        next_rows_[vtx] = static_cast<int32_t>(out);
        for (size_t s = static_cast<size_t>(row_offsets[vtx]); s < static_cast<size_t>(row_offsets[vtx + 1]); ++s) {
          if (active[static_cast<size_t>(edge_ids[s])] != 0) {
            next_neighbors_[out] = neighbors[s];
            next_edge_ids_[out] = edge_ids[s];
            ++out;
          }
        }
      }
    });
    next_rows_[static_cast<size_t>(n)] = static_cast<int32_t>(kept[kBlocks]);
    row_offsets.swap(next_rows_);
    neighbors.swap(next_neighbors_);
    edge_ids.swap(next_edge_ids_);
    return;
  }
  int32_t out = 0, start = row_offsets[0];
  for (int32_t vtx = 0; vtx < n; ++vtx) {
    const int32_t end = row_offsets[static_cast<size_t>(vtx) + 1];
    row_offsets[static_cast<size_t>(vtx)] = out;
    for (int32_t s = start; s < end; ++s) {
      if (active[static_cast<size_t>(edge_ids[static_cast<size_t>(s)])] != 0) {
        neighbors[static_cast<size_t>(out)] = neighbors[static_cast<size_t>(s)];
        edge_ids[static_cast<size_t>(out)] = edge_ids[static_cast<size_t>(s)];
        ++out;
      }
    }
    start = end;
  }
  row_offsets[static_cast<size_t>(n)] = out;
  neighbors.resize(static_cast<size_t>(out));
  edge_ids.resize(static_cast<size_t>(out));
}

uint64_t JvHost::ComputeHash(const uint8_t* active_flags) const {
  return ContentHashEdges(prefix_hash_, m, u.data(), v.data(), weight.data(), active_flags);
}

uint64_t JvHost::SnapHash::Get() {
  if (!ready) {
    value = job.get();
    ready = true;
  }
  return value;
}

JvHost::SnapHash JvHost::CurrentSnap() {
  if (!snaps_.empty() && snaps_.back().generation == generation_) return snaps_.back();
  SnapHash h;
  h.generation = generation_;
  if (hash_source && hash_source(&h.job)) return h;
  if (async_hash) {
    // the job hashes a copy of this generation's flags, so commits may
    // change the snapshot while it runs
    std::vector<uint8_t> flags(active);
    h.job = std::async(std::launch::async,
                       [this, flags = std::move(flags)] { return ComputeHash(flags.data()); })
                .share();
  }
  return h;
}

bool JvHost::Verify(int32_t e, int32_t c) const {
  if (e < 0 || e >= m || !active[e]) return false;
  const int32_t a = u[e], b = v[e];
  if (Degree(a) <= 2 || Degree(b) <= 2) return false;
  if (c < 0 || c >= n || c == a || c == b) return false;
  const __int128 cab = weight[e], cac = Dist(a, c), cbc = Dist(b, c);
  for (int32_t s = row_offsets[c]; s < row_offsets[c + 1]; ++s) {
    const int32_t d = neighbors[s];
    if (d == a || d == b) continue;
    const __int128 left = cab + weight[edge_ids[s]];
    if (left <= cac + Dist(d, b) || left <= Dist(a, d) + cbc) return false;
  }
  return true;
}

void JvHost::StartRun() {
  for (SnapHash& h : snaps_) {
    if (h.job.valid()) h.Get();
  }
  live_edges_ = 0;
  for (int32_t e = 0; e < m; ++e) live_edges_ += active[e] != 0;
  records_.clear();
  metrics_.clear();
  snaps_.clear();
  stats = ee_jv_run_stats{};
  done = false;
  proof.clear();
}

void JvHost::BeginEpoch(uint32_t epoch) {
  cur_ = Metrics{};
  cur_.epoch = epoch;
  cur_.before = live_edges_;
  snap_generation_ = generation_;
  snaps_.push_back(CurrentSnap());
  gated_.clear();
  gated_edges_.clear();
}

const std::vector<int32_t>& JvHost::VerifyAndGate(const int32_t* witnesses, double propose_ms) {
  Metrics& mt = cur_;
  std::vector<std::pair<int32_t, int32_t>>& proposed = gated_;
  proposed.clear();
  gated_edges_.clear();
  for (int32_t e = 0; e < m; ++e) {
    if (witnesses[e] >= 0) proposed.emplace_back(e, witnesses[e]);
  }
  mt.propose_ms = propose_ms;
  mt.proposed = proposed.size();

  // host re-verification on the same snapshot
  auto t0 = std::chrono::steady_clock::now();
  constexpr size_t kVerifyBlock = 64, kParallelVerify = 512;
  if (verify_threads != 1 && proposed.size() >= kParallelVerify) {
    std::atomic<size_t> rejected{0};
    const size_t blocks = (proposed.size() + kVerifyBlock - 1) / kVerifyBlock;
    ParallelFor(blocks, [&](size_t b) {
      size_t r = 0;
      const size_t end = std::min(proposed.size(), (b + 1) * kVerifyBlock);
      for (size_t i = b * kVerifyBlock; i < end; ++i) r += Verify(proposed[i].first, proposed[i].second) ? 0 : 1;
      rejected += r;
    }, verify_threads);
    mt.rejected = rejected.load();
  } else {
    for (const auto& c : proposed) {
      if (!Verify(c.first, c.second)) ++mt.rejected;
    }
  }
  mt.verify_ms = MsSince(t0);
  mt.verified = proposed.size() - mt.rejected;
  stats.propose_ms += mt.propose_ms;
  stats.verify_ms += mt.verify_ms;
  if (mt.rejected != 0) {
    proposed.clear();
    throw std::runtime_error("HIP JV candidate failed CPU re-verification; epoch stopped, nothing committed");
  }

  // deterministic commit: sort by (u, v, method, witness, edge), one per edge,
  // running degree gate > 2 at both ends (src/cpu/elimination_commit.cpp:13-75)
  t0 = std::chrono::steady_clock::now();
  // the snapshot must be the one proposed on (elimination_commit.cpp:16-18);
  // it changes only through commits, which bump the generation
  if (generation_ != snap_generation_) throw std::runtime_error("snapshot changed before commit");
  const auto key_less = [&](const auto& l, const auto& r) {
    return std::make_tuple(u[l.first], v[l.first], l.second, l.first) <
           std::make_tuple(u[r.first], v[r.first], r.second, r.first);
  };
  // proposals come in edge-slot order, which for a snapshot sorted by (u, v)
  // (as the reference loads them) is already the commit order
  if (!std::is_sorted(proposed.begin(), proposed.end(), key_less)) {
    std::sort(proposed.begin(), proposed.end(), key_less);
  }
  proposed.erase(std::unique(proposed.begin(), proposed.end(),
                             [](const auto& l, const auto& r) { return l.first == r.first; }),
                 proposed.end());
  std::vector<int32_t> degree(static_cast<size_t>(n));
  for (int32_t vtx = 0; vtx < n; ++vtx) degree[vtx] = Degree(vtx);
  size_t kept = 0;
  for (const auto& c : proposed) {
    int32_t& du = degree[u[c.first]];
    int32_t& dv = degree[v[c.first]];
    if (du <= 2 || dv <= 2) continue;
    --du;
    --dv;
    proposed[kept++] = c;
    gated_edges_.push_back(c.first);
  }
  proposed.resize(kept);
  stats.commit_ms += MsSince(t0);
  return gated_edges_;
}

size_t JvHost::ApplyCommit() {
  Metrics& mt = cur_;
  const auto t0 = std::chrono::steady_clock::now();
  // records carry the hash of the snapshot proposed on
  const size_t snap = snaps_.size() - 1;
  if (!async_hash && !gated_.empty() && !snaps_[snap].ready) {
    snaps_[snap].value = ComputeHash(active.data());  // before the snapshot changes
    snaps_[snap].ready = true;
  }
  if (!gated_.empty()) {
    for (const auto& c : gated_) active[c.first] = 0;
    live_edges_ -= gated_.size();  // each was verified active
    ++generation_;
    DropInactiveFromCsr();
  }
  stats.commit_ms += MsSince(t0);
  mt.committed = gated_.size();
  for (const auto& c : gated_) {
    records_.push_back({mt.epoch, snap, c.first, u[c.first], v[c.first], c.second});
  }
  metrics_.push_back(mt);
  stats.edges_scanned += mt.before;
  stats.proposed += mt.proposed;
  stats.committed += mt.committed;
  ++stats.epochs;
  const size_t committed = gated_.size();
  if (committed == 0) done = true;
  gated_.clear();
  return committed;
}

void JvHost::EndRun(double total_ms) {
  done = true;
  const auto t0 = std::chrono::steady_clock::now();
  // the snapshot of the first epoch is the run's initial snapshot; the
  // final one is the current snapshot
  SnapHash final_snap = CurrentSnap();
  if (!async_hash && !final_snap.ready) {
    final_snap.value = ComputeHash(active.data());
    final_snap.ready = true;
  }
  if (snaps_.empty()) snaps_.push_back(final_snap);
  for (SnapHash& h : snaps_) {
    if (!h.ready && !h.job.valid()) {  // an epoch that committed nothing (synchronous mode)
      if (h.generation != generation_) throw std::logic_error("JV snapshot hash of a past generation is missing");
      h.value = final_snap.Get();
      h.ready = true;
    }
  }
  stats.initial_hash = snaps_.front().Get();
  stats.final_hash = final_snap.Get();
  stats.hash_wait_ms = MsSince(t0);
  stats.total_ms = total_ms + stats.hash_wait_ms;  // the run ends when its hashes are known
  // V1 proof text (src/cpu/jv_elimination.cpp:324-368)
  std::ostringstream out;
  out << "CUDAEE_PROOF_V1\nbackend hip\ninitial_hash " << Hex(stats.initial_hash) << '\n';
  for (const Record& r : records_) {
    out << "record " << r.epoch << ' ' << Hex(snaps_[r.snap].Get()) << ' ' << r.edge << ' ' << r.a << ' '
        << r.b << " JV " << r.witness << '\n';
  }
  for (const Metrics& mt : metrics_) {
    out << "metrics " << mt.epoch << ' ' << mt.before << ' ' << mt.proposed << ' ' << mt.verified
        << ' ' << mt.rejected << ' ' << mt.committed << ' ' << mt.propose_ms << ' ' << mt.verify_ms
        << '\n';
  }
  out << "final_hash " << Hex(stats.final_hash) << "\nEND\n";
  proof = out.str();
}

}  // namespace ee
