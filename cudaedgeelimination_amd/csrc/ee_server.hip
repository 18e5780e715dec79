// Serving loop for leaf-cost queries and JV rechecks over device-resident
// graph snapshots.
//
// Callers submit requests from any thread: leaf-cost queries (graph, k, task
// rows, output buffer) and JV rechecks (JV graph, edge slots, witness
// buffer).  One worker thread forms batches from the queue in arrival order
// under a cell limit, a request limit and an HBM budget, and issues each
// batch on one of `pipeline_depth` slots (own stream, device buffers and
// pinned staging): upload tasks, tiles and edge lists, one launch of the
// tiled k-opt kernel over all the batch's (graph, k) segments and one JV
// launch over all its listed edge slots, download the results.  While a batch
// is in flight the worker keeps admitting and issuing into a free slot, so
// requests that arrive meanwhile overlap on the device.  A second thread
// blocks on the oldest batch in flight, copies each request's results into
// the caller's buffer, wakes waiters and frees the slot; a batch issued while
// the server is otherwise idle is completed by the worker itself, which polls
// it until it finishes or another request arrives.  With one slot there is no
// second thread: the worker issues, waits and completes each batch in turn
// (the plain per-query path).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ee_jvhost.h"
#include "ee_runtime.h"

namespace ee {

void CheckGraphView(const ee_graph_view* g);

namespace {

using Clock = std::chrono::steady_clock;
constexpr uint32_t kTemplates[3] = {4, 25, 208};

double Ms(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}


struct Request {
  uint64_t id = 0;
  bool jv = false;  // a JV recheck rather than a leaf-cost query
  uint32_t graph = 0, k = 0;
  std::vector<ee_kopt_task> tasks;  // owned copy (ee_server_submit)
  const ee_kopt_task* task_data = nullptr;  // tasks.data() or the caller's rows
  uint64_t task_rows = 0;
  int64_t* out = nullptr;
  std::vector<int32_t> edges;  // JV: union edge slots
  int32_t* witnesses = nullptr;
  int32_t node_base = 0;       // JV: first union node of the graph
  uint64_t cells = 0, bytes = 0;
  Clock::time_point submit, issue, done;
  int32_t status = EE_OK;
  std::string error;
  bool complete = false;
  uint32_t batch_size = 0;
  uint64_t batch_id = 0;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t ev_begin = nullptr, ev_end = nullptr, ev_done = nullptr;
  DeviceArray<ee_kopt_task> d_tasks;
  DeviceArray<int64_t> d_costs;
  DeviceArray<KoptTile> d_tiles;
  Pinned<ee_kopt_task> h_tasks;
  Pinned<int64_t> h_costs;
  Pinned<KoptTile> h_tiles;
  std::vector<KoptSegment> segments;
  DeviceArray<int32_t> d_jv_list, d_jv_out;
  Pinned<int32_t> h_jv_list, h_jv_out;
  std::vector<std::shared_ptr<Request>> requests;
  std::vector<uint64_t> offsets;
  uint64_t bytes = 0;
  uint64_t batch_id = 0;
  bool busy = false;
};

struct Server {
  Context* ctx = nullptr;
  ee_server_config cfg{};
  KoptTuning tuning;
  std::array<KoptPlan, 3> plans;
  DeviceArray<KoptPlan> d_plans;
  // graphs
  std::vector<KoptInstance> instances;
  std::vector<longlong2> host_xy;
  DeviceArray<longlong2> d_xy;
  bool graphs_dirty = false;
  std::vector<KoptInstance> worker_instances;  // the worker thread's copy
  // JV snapshots, back to back in one union (node ids, edge slots and CSR
  // offset by graph); mutated under mu, uploaded by the worker when dirty
  struct JvUnion {
    std::vector<int2> uv;
    std::vector<int64_t> weight;
    std::vector<uint8_t> mode;
    std::vector<longlong2> xy;
    std::vector<int32_t> active, row_offsets{0}, neighbors, edge_ids;
    std::vector<int32_t> node_base, edge_base, edge_count;
  };
  JvUnion jv;
  bool jv_dirty = false;
  DeviceArray<int2> d_jv_uv;
  DeviceArray<int64_t> d_jv_weight;
  DeviceArray<uint8_t> d_jv_mode;
  DeviceArray<longlong2> d_jv_xy;
  DeviceArray<int32_t> d_jv_active, d_jv_row_offsets, d_jv_neighbors, d_jv_edge_ids;
  int32_t worker_jv_slots = 0;
  // queue
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> pending;
  std::unordered_map<uint64_t, std::shared_ptr<Request>> live;
  uint64_t next_id = 1, next_batch = 1;
  bool stopping = false;
  bool issuer_done = false;
  ee_server_stats stats{};
  uint64_t inflight_bytes = 0;
  std::vector<Slot> slots;
  std::deque<Slot*> inflight;  // issued batches, oldest first (guarded by mu)
  std::condition_variable cv_inflight;
  std::thread worker, completer;

  uint64_t SlotBudget() const { return cfg.hbm_budget_bytes / std::max<uint32_t>(cfg.pipeline_depth, 1); }

  // Batched mode: how long the issuing and completing threads poll (yielding
  // the core) for their next piece of work before they block.  A blocked
  // thread is woken by the operating system, tens of microseconds later and
  // more in the tail.  The issuer's window covers the gaps between requests
  // of a steady stream down to a few hundred requests per second; the
  // completer's the time between a batch's issue and its hand-over.
  static constexpr double kIssuerSpinMicros = 5000.0;
  static constexpr double kCompleterSpinMicros = 200.0;
  // bumped after every change the two threads wait for, so that a polling
  // thread looks at the queue (under the lock) only when something changed
  std::atomic<uint64_t> signal_seq{0};
  void Signal(std::condition_variable& cv, bool all = false) {
    signal_seq.fetch_add(1, std::memory_order_release);
    if (all) {
      cv.notify_all();
    } else {
      cv.notify_one();
    }
  }
  template <typename Pred>
  void SpinThenWait(std::unique_lock<std::mutex>& lk, std::condition_variable& cv, Pred ready,
                    double spin_micros);

  void Loop();
  void CompleteLoop();
  // Waits for a slot's batch and completes (or fails) its requests.
  // `poll`: wait for the batch by polling its event (yielding) rather than
  // blocking on it
  void Finish(Slot& slot, bool poll = false);
  bool Issue(Slot& slot, std::vector<std::shared_ptr<Request>> batch);
  void Complete(Slot& slot);
  void Fail(std::vector<std::shared_ptr<Request>>& batch, const std::string& why);
};

void Server::Fail(std::vector<std::shared_ptr<Request>>& batch, const std::string& why) {
  std::lock_guard<std::mutex> lk(mu);
  const auto now = Clock::now();
  for (auto& r : batch) {
    r->status = EE_ERR_RUNTIME;
    r->error = why;
    r->done = now;
    r->complete = true;
  }
  cv_done.notify_all();
}

// Stages one batch into a slot and enqueues upload, launch and download.
bool Server::Issue(Slot& slot, std::vector<std::shared_ptr<Request>> batch) {
  try {
    CheckHip(hipSetDevice(ctx->device), "hipSetDevice");
    std::vector<longlong2> xy_copy;
    std::vector<KoptInstance> inst_copy;
    std::unique_ptr<JvUnion> jv_copy;
    {
      std::lock_guard<std::mutex> lk(mu);
      if (graphs_dirty) {
        xy_copy = host_xy;
        inst_copy = instances;
        graphs_dirty = false;
      }
      if (jv_dirty) {
        jv_copy = std::make_unique<JvUnion>(jv);
        jv_dirty = false;
      }
    }
    if (!inst_copy.empty() || jv_copy) {
      // new graphs: wait for every batch that reads the old pools, then replace them
      for (Slot& s : slots) CheckHip(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
    }
    if (!inst_copy.empty()) {
      d_xy.Allocate(xy_copy.size(), ctx->device);
      d_xy.Upload(xy_copy.data(), xy_copy.size(), slot.stream);
      CheckHip(hipStreamSynchronize(slot.stream), "hipStreamSynchronize");
      worker_instances = std::move(inst_copy);
    }
    if (jv_copy) {
      const JvUnion& u = *jv_copy;
      const int dev = ctx->device;
      d_jv_uv.Allocate(std::max<size_t>(u.uv.size(), 1), dev);
      d_jv_weight.Allocate(std::max<size_t>(u.weight.size(), 1), dev);
      d_jv_mode.Allocate(std::max<size_t>(u.mode.size(), 1), dev);
      d_jv_active.Allocate(std::max<size_t>(u.active.size(), 1), dev);
      d_jv_xy.Allocate(std::max<size_t>(u.xy.size(), 1), dev);
      d_jv_row_offsets.Allocate(u.row_offsets.size(), dev);
      d_jv_neighbors.Allocate(std::max<size_t>(u.neighbors.size(), 1), dev);
      d_jv_edge_ids.Allocate(std::max<size_t>(u.edge_ids.size(), 1), dev);
      d_jv_uv.Upload(u.uv.data(), u.uv.size(), slot.stream);
      d_jv_weight.Upload(u.weight.data(), u.weight.size(), slot.stream);
      d_jv_mode.Upload(u.mode.data(), u.mode.size(), slot.stream);
      d_jv_active.Upload(u.active.data(), u.active.size(), slot.stream);
      d_jv_xy.Upload(u.xy.data(), u.xy.size(), slot.stream);
      d_jv_row_offsets.Upload(u.row_offsets.data(), u.row_offsets.size(), slot.stream);
      d_jv_neighbors.Upload(u.neighbors.data(), u.neighbors.size(), slot.stream);
      d_jv_edge_ids.Upload(u.edge_ids.data(), u.edge_ids.size(), slot.stream);
      CheckHip(hipStreamSynchronize(slot.stream), "hipStreamSynchronize");
      worker_jv_slots = static_cast<int32_t>(u.uv.size());
    }
    // leaf-cost requests: one tiled k-opt launch; JV rechecks: one JV launch
    uint64_t rows = 0, cells = 0, tiles = 0, items = 0;
    size_t kopt_requests = 0;
    bool present[3] = {false, false, false};
    slot.offsets.clear();
    for (const auto& r : batch) {
      if (r->jv) {
        slot.offsets.push_back(items);
        items += r->edges.size();
        continue;
      }
      // requests start on 128-byte boundaries (whole-line stores stay aligned)
      cells = (cells + kSegmentAlignCells - 1) / kSegmentAlignCells * kSegmentAlignCells;
      slot.offsets.push_back(cells);
      rows += r->task_rows;
      cells += r->cells;
      tiles += (r->task_rows + tuning.rows_per_tile[r->k - 3] - 1) / tuning.rows_per_tile[r->k - 3];
      present[r->k - 3] = true;
      ++kopt_requests;
    }
    CheckHip(hipEventRecord(slot.ev_begin, slot.stream), "event record");
    if (kopt_requests > 0) {
      slot.h_tasks.Reserve(rows);
      slot.h_costs.Reserve(cells);
      slot.segments.resize(kopt_requests);
      slot.h_tiles.Reserve(tiles);
      if (slot.d_tasks.count() < rows) slot.d_tasks.Allocate(Grow(slot.d_tasks.count(), rows), ctx->device);
      if (slot.d_costs.count() < cells) slot.d_costs.Allocate(Grow(slot.d_costs.count(), cells), ctx->device);
      if (slot.d_tiles.count() < tiles) slot.d_tiles.Allocate(Grow(slot.d_tiles.count(), tiles), ctx->device);
      uint64_t row = 0;
      size_t seg = 0;
      for (size_t i = 0; i < batch.size(); ++i) {
        const Request& r = *batch[i];
        if (r.jv) continue;
        std::memcpy(slot.h_tasks.get() + row, r.task_data, r.task_rows * sizeof(ee_kopt_task));
        slot.segments[seg++] = KoptSegment{r.graph, r.k, row, r.task_rows, slot.offsets[i], 0, 0, 0};
        row += r.task_rows;
      }
      KoptLaunch a{};
      KoptAssignTiles(slot.segments.data(), kopt_requests, worker_instances.data(), plans.data(), tuning,
                      slot.h_tiles.get(), &a);
      if (a.tile_count != tiles) throw std::logic_error("server tile count mismatch");
      slot.d_tasks.Upload(slot.h_tasks.get(), rows, slot.stream);
      if (a.inline_count == 0) slot.d_tiles.Upload(slot.h_tiles.get(), tiles, slot.stream);
      a.xy = d_xy.get();
      a.tiles = slot.d_tiles.get();
      a.tasks = slot.d_tasks.get();
      a.plans = d_plans.get();
      a.costs = slot.d_costs.get();
      bool near_only = true;
      for (const KoptSegment& sg : slot.segments) near_only = near_only && worker_instances[sg.instance].near != 0;
      const size_t shared = KoptPrepareLayout(&a, plans.data(), present, tuning, near_only);
      LaunchKopt(a, tuning, shared, slot.stream);
    }
    if (items > 0) {
      slot.h_jv_list.Reserve(items);
      slot.h_jv_out.Reserve(items);
      if (slot.d_jv_list.count() < items) slot.d_jv_list.Allocate(Grow(slot.d_jv_list.count(), items), ctx->device);
      if (slot.d_jv_out.count() < items) slot.d_jv_out.Allocate(Grow(slot.d_jv_out.count(), items), ctx->device);
      for (size_t i = 0; i < batch.size(); ++i) {
        const Request& r = *batch[i];
        if (!r.jv) continue;
        for (int32_t e : r.edges) {
          if (e < 0 || e >= worker_jv_slots) throw std::logic_error("JV recheck slot outside the resident union");
        }
        std::memcpy(slot.h_jv_list.get() + slot.offsets[i], r.edges.data(), r.edges.size() * sizeof(int32_t));
      }
      slot.d_jv_list.Upload(slot.h_jv_list.get(), items, slot.stream);
      JvLaunch j{};
      j.edge_count = static_cast<int32_t>(items);
      j.edge_uv = d_jv_uv.get();
      j.edge_weight = d_jv_weight.get();
      j.edge_active = d_jv_active.get();
      j.row_offsets = d_jv_row_offsets.get();
      j.neighbors = d_jv_neighbors.get();
      j.csr_edge_ids = d_jv_edge_ids.get();
      j.xy = d_jv_xy.get();
      j.witnesses = slot.d_jv_out.get();
      j.edge_mode = d_jv_mode.get();
      j.edge_list = slot.d_jv_list.get();
      LaunchJv(j, 0, slot.stream);
    }
    CheckHip(hipEventRecord(slot.ev_end, slot.stream), "event record");
    if (kopt_requests > 0) slot.d_costs.Download(slot.h_costs.get(), cells, slot.stream);
    if (items > 0) slot.d_jv_out.Download(slot.h_jv_out.get(), items, slot.stream);
    CheckHip(hipEventRecord(slot.ev_done, slot.stream), "event record");
    {
      std::lock_guard<std::mutex> lk(mu);
      const auto now = Clock::now();
      slot.batch_id = next_batch++;
      for (auto& r : batch) {
        r->issue = now;
        r->batch_size = static_cast<uint32_t>(batch.size());
        r->batch_id = slot.batch_id;
      }
      slot.bytes = 0;
      for (auto& r : batch) slot.bytes += r->bytes;
      inflight_bytes += slot.bytes;
      stats.peak_inflight_bytes = std::max(stats.peak_inflight_bytes, inflight_bytes);
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      slot.requests = std::move(batch);
      slot.busy = true;
    }
    return true;
  } catch (const std::exception& e) {
    Fail(batch, e.what());
    return false;
  }
}

void Server::Complete(Slot& slot) {
  float ms = 0.0f;
  static_cast<void>(hipEventElapsedTime(&ms, slot.ev_begin, slot.ev_end));
  for (size_t i = 0; i < slot.requests.size(); ++i) {
    Request& r = *slot.requests[i];
    if (r.jv) {
      // union node ids back to the graph's own ids
      const int32_t* w = slot.h_jv_out.get() + slot.offsets[i];
      for (size_t j = 0; j < r.edges.size(); ++j) r.witnesses[j] = w[j] >= 0 ? w[j] - r.node_base : -1;
    } else {
      std::memcpy(r.out, slot.h_costs.get() + slot.offsets[i], r.cells * sizeof(int64_t));
    }
  }
  std::lock_guard<std::mutex> lk(mu);
  const auto now = Clock::now();
  for (auto& r : slot.requests) {
    r->done = now;
    r->complete = true;
    if (r->jv) {
      stats.jv_edges += r->edges.size();
    } else {
      stats.cells += r->cells;
    }
  }
  stats.requests += slot.requests.size();
  stats.batches += 1;
  stats.device_ms += ms;
  inflight_bytes -= slot.bytes;
  slot.requests.clear();
  slot.busy = false;
  cv_done.notify_all();
}

template <typename Pred>
void Server::SpinThenWait(std::unique_lock<std::mutex>& lk, std::condition_variable& cv, Pred ready,
                          double spin_micros) {
  if (ready()) return;
  if (slots.size() > 1) {
    const auto until = std::chrono::steady_clock::now() + std::chrono::duration<double, std::micro>(spin_micros);
    uint64_t seen = signal_seq.load(std::memory_order_acquire);
    lk.unlock();
    while (std::chrono::steady_clock::now() < until) {
      if (signal_seq.load(std::memory_order_acquire) != seen) {
        lk.lock();
        if (ready()) return;
        seen = signal_seq.load(std::memory_order_acquire);
        lk.unlock();
      }
      std::this_thread::yield();
    }
    lk.lock();
  }
  cv.wait(lk, ready);
}

// Issuing side: whenever a slot is free and requests are pending, packs the
// next batch in arrival order and issues it.
void Server::Loop() {
  const uint64_t budget = SlotBudget();
  for (;;) {
    Slot* free_slot = nullptr;
    std::vector<std::shared_ptr<Request>> batch;
    {
      std::unique_lock<std::mutex> lk(mu);
      SpinThenWait(lk, cv_work, [&] {
        if (stopping && pending.empty()) return true;
        if (pending.empty()) return false;
        for (Slot& s : slots) {
          if (!s.busy) return true;
        }
        return false;
      }, kIssuerSpinMicros);
      if (stopping && pending.empty()) break;
      for (Slot& s : slots) {
        if (!s.busy) {
          free_slot = &s;
          break;
        }
      }
      uint64_t cells = 0, bytes = 0;
      while (!pending.empty() && batch.size() < cfg.max_batch_requests) {
        const auto& r = pending.front();
        if (!batch.empty() && (cells + r->cells > cfg.max_batch_cells || bytes + r->bytes > budget)) {
          break;
        }
        cells += r->cells;
        bytes += r->bytes;
        batch.push_back(r);
        pending.pop_front();
      }
    }
    if (Issue(*free_slot, std::move(batch))) {
      if (slots.size() == 1) {
        // one slot: the plain path, upload, launch and download in turn
        Finish(*free_slot);
        continue;
      }
      // With nothing else in flight and nothing waiting, completing here
      // saves the hand-over to the completing thread (the lone request's
      // latency is then the plain path's); the batch goes to that thread as
      // soon as another request arrives, so the issuer is free to overlap it.
      bool inline_finish = false;
      for (;;) {
        {
          std::lock_guard<std::mutex> lk(mu);
          if (!inflight.empty() || !pending.empty() || stopping) {
            inflight.push_back(free_slot);
            Signal(cv_inflight);
            break;
          }
        }
        const hipError_t q = hipEventQuery(free_slot->ev_done);
        if (q != hipErrorNotReady) {  // done, or failed: Finish reports it
          inline_finish = true;
          break;
        }
        std::this_thread::yield();
      }
      if (inline_finish) Finish(*free_slot);
    }
  }
  std::lock_guard<std::mutex> lk(mu);
  issuer_done = true;
  Signal(cv_inflight);
}

// Completing side: blocks on the oldest issued batch, hands its results to
// the callers and frees its slot.
void Server::CompleteLoop() {
  for (;;) {
    Slot* slot = nullptr;
    bool poll = false;
    {
      std::unique_lock<std::mutex> lk(mu);
      SpinThenWait(lk, cv_inflight, [&] { return !inflight.empty() || issuer_done; }, kCompleterSpinMicros);
      if (inflight.empty()) return;  // the issuer stopped and everything completed
      slot = inflight.front();
      inflight.pop_front();
      // with requests queued the issuer is busy in the HIP runtime, and
      // polling the event would contend with its calls
      poll = pending.empty();
    }
    Finish(*slot, poll);
    Signal(cv_work);
  }
}

void Server::Finish(Slot& slot, bool poll) {
  hipError_t st;
  if (poll) {
    while ((st = hipEventQuery(slot.ev_done)) == hipErrorNotReady) std::this_thread::yield();
  } else {
    st = hipEventSynchronize(slot.ev_done);
  }
  if (st == hipSuccess) {
    Complete(slot);
    return;
  }
  std::vector<std::shared_ptr<Request>> failed;
  {
    std::lock_guard<std::mutex> lk(mu);
    failed = std::move(slot.requests);
    inflight_bytes -= slot.bytes;
    slot.busy = false;
  }
  Fail(failed, std::string("batch failed: ") + hipGetErrorString(st));
}

Server& AsServer(ee_server* s) {
  if (s == nullptr) throw std::invalid_argument("server is NULL");
  return *reinterpret_cast<Server*>(s);
}

template <typename F> ee_status Guarded(F&& body) {
  try {
    body();
    return EE_OK;
  } catch (...) {
    return StatusFromException();
  }
}

}  // namespace
}  // namespace ee

using ee::Guarded;

namespace ee {
namespace {

// A validated leaf-cost request; `borrow` keeps the caller's rows instead of
// copying them (they must then stay valid until the request completes).
std::shared_ptr<Request> MakeKoptRequest(Server& s, uint32_t graph, uint32_t k, const ee_kopt_task* tasks,
                                         uint64_t count, int64_t* costs, bool borrow) {
  if (k < 3 || k > 5) throw std::invalid_argument("request k must be in [3,5]");
  if (count == 0 || tasks == nullptr || costs == nullptr) {
    throw std::invalid_argument("request needs at least one task and an output buffer");
  }
  auto r = std::make_shared<Request>();
  r->graph = graph;
  r->k = k;
  r->cells = count * kTemplates[k - 3];
  r->bytes = r->cells * sizeof(int64_t) + count * sizeof(ee_kopt_task);
  r->out = costs;
  if (r->bytes > s.SlotBudget()) throw std::invalid_argument("request exceeds the HBM budget");
  uint32_t n = 0;
  {
    std::lock_guard<std::mutex> lk(s.mu);
    if (graph >= s.instances.size()) throw std::invalid_argument("unknown graph id");
    n = s.instances[graph].n;
  }
  ValidateTasks(static_cast<int32_t>(n), k, tasks, count);
  if (borrow) {
    r->task_data = tasks;
  } else {
    r->tasks.assign(tasks, tasks + count);
    r->task_data = r->tasks.data();
  }
  r->task_rows = count;
  return r;
}

// A validated JV recheck of edge slots of a resident JV graph.
std::shared_ptr<Request> MakeJvRequest(Server& s, uint32_t jv_graph, const int32_t* edges, uint64_t count,
                                       int32_t* witnesses) {
  if (count == 0 || edges == nullptr || witnesses == nullptr) {
    throw std::invalid_argument("a JV recheck needs at least one edge slot and an output buffer");
  }
  if (count >= (uint64_t{1} << 31)) throw std::invalid_argument("JV recheck too large");
  auto r = std::make_shared<Request>();
  r->jv = true;
  r->graph = jv_graph;
  r->cells = count;  // counts against the batch limit like cells
  r->bytes = count * 2 * sizeof(int32_t);
  r->witnesses = witnesses;
  if (r->bytes > s.SlotBudget()) throw std::invalid_argument("request exceeds the HBM budget");
  int32_t eb = 0, m = 0;
  {
    std::lock_guard<std::mutex> lk(s.mu);
    if (jv_graph >= s.jv.node_base.size()) throw std::invalid_argument("unknown JV graph id");
    eb = s.jv.edge_base[jv_graph];
    m = s.jv.edge_count[jv_graph];
    r->node_base = s.jv.node_base[jv_graph];
  }
  r->edges.resize(count);
  for (uint64_t i = 0; i < count; ++i) {
    if (edges[i] < 0 || edges[i] >= m) throw std::invalid_argument("JV recheck edge slot out of range");
    r->edges[i] = eb + edges[i];
  }
  return r;
}

// Queues validated requests in order under one lock and wakes the worker.
void Enqueue(Server& s, std::vector<std::shared_ptr<Request>>& rs, uint64_t* ids) {
  {
    std::lock_guard<std::mutex> lk(s.mu);
    if (s.stopping) throw std::runtime_error("server is stopping");
    const Clock::time_point now = Clock::now();
    for (size_t i = 0; i < rs.size(); ++i) {
      rs[i]->id = s.next_id++;
      rs[i]->submit = now;
      s.live[rs[i]->id] = rs[i];
      s.pending.push_back(rs[i]);
      if (ids) ids[i] = rs[i]->id;
    }
  }
  s.Signal(s.cv_work);
}

}  // namespace
}  // namespace ee

extern "C" {

ee_status ee_server_create(ee_ctx* handle, const ee_server_config* config, ee_server** out) {
  return Guarded([&] {
    if (handle == nullptr || config == nullptr || out == nullptr) throw std::invalid_argument("NULL argument");
    *out = nullptr;
    if (config->pipeline_depth < 1 || config->pipeline_depth > 2 || config->max_batch_requests < 1 ||
        config->max_batch_cells < 1 || config->hbm_budget_bytes < 1024) {
      throw std::invalid_argument("bad server configuration");
    }
    auto* ctx = reinterpret_cast<ee::Context*>(handle);
    ctx->Activate();
    auto s = std::make_unique<ee::Server>();
    s->ctx = ctx;
    s->cfg = *config;
    for (uint32_t k = 3; k <= 5; ++k) {
      uint64_t hash = 0;
      const auto table = ee::CanonicalTemplates(k, &hash);
      s->plans[k - 3] = ee::BuildPlan(k, table.data(), static_cast<uint32_t>(table.size()));
    }
    s->d_plans.Allocate(3, ctx->device);
    s->d_plans.Upload(s->plans.data(), 3, ctx->stream);
    ctx->Sync();
    s->slots.resize(config->pipeline_depth);
    // size every slot for a full batch up front, so that serving does not
    // allocate (device frees synchronise the device; pinned allocations are slow)
    const uint64_t max_cells =
        std::max<uint64_t>(1, std::min<uint64_t>(config->max_batch_cells, s->SlotBudget() / sizeof(int64_t)));
    const uint64_t max_rows = max_cells / 4 + 1;  // k = 3 has 4 cells per row
    // plus the alignment padding between the requests of a batch
    const uint64_t max_pool = max_cells + ee::kSegmentAlignCells * config->max_batch_requests;
    for (ee::Slot& slot : s->slots) {
      slot.d_tasks.Allocate(max_rows, ctx->device);
      slot.d_costs.Allocate(max_pool, ctx->device);
      slot.d_tiles.Allocate(max_rows, ctx->device);
      slot.d_jv_list.Allocate(max_cells, ctx->device);
      slot.d_jv_out.Allocate(max_cells, ctx->device);
      slot.h_tasks.Reserve(max_rows);
      slot.h_costs.Reserve(max_pool);
      slot.h_tiles.Reserve(max_rows);
      slot.h_jv_list.Reserve(max_cells);
      slot.h_jv_out.Reserve(max_cells);
    }
    for (ee::Slot& slot : s->slots) {
      ee::CheckHip(hipStreamCreateWithFlags(&slot.stream, hipStreamNonBlocking), "hipStreamCreate");
      ee::CheckHip(hipEventCreate(&slot.ev_begin), "hipEventCreate");
      ee::CheckHip(hipEventCreate(&slot.ev_end), "hipEventCreate");
      ee::CheckHip(hipEventCreateWithFlags(&slot.ev_done, hipEventDisableTiming), "hipEventCreate");
    }
    ee::Server* raw = s.get();
    raw->worker = std::thread([raw] { raw->Loop(); });
    if (raw->slots.size() > 1) {
      raw->completer = std::thread([raw] {
        static_cast<void>(hipSetDevice(raw->ctx->device));
        raw->CompleteLoop();
      });
    }
    *out = reinterpret_cast<ee_server*>(s.release());
  });
}

void ee_server_destroy(ee_server* server) {
  if (server == nullptr) return;
  auto* s = reinterpret_cast<ee::Server*>(server);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stopping = true;
  }
  s->Signal(s->cv_work, true);
  if (s->worker.joinable()) s->worker.join();
  if (s->completer.joinable()) s->completer.join();
  static_cast<void>(hipSetDevice(s->ctx->device));
  for (ee::Slot& slot : s->slots) {
    static_cast<void>(hipStreamSynchronize(slot.stream));
    static_cast<void>(hipEventDestroy(slot.ev_begin));
    static_cast<void>(hipEventDestroy(slot.ev_end));
    static_cast<void>(hipEventDestroy(slot.ev_done));
    static_cast<void>(hipStreamDestroy(slot.stream));
  }
  delete s;
}

ee_status ee_server_add_graph(ee_server* server, const int64_t* x, const int64_t* y,
                              int32_t dimension, uint8_t distance_type, uint32_t* graph) {
  return Guarded([&] {
    ee::Server& s = ee::AsServer(server);
    if (x == nullptr || y == nullptr || dimension <= 0) throw std::invalid_argument("bad coordinates");
    if (distance_type != EE_EUC_2D && distance_type != EE_CEIL_2D) {
      throw std::invalid_argument("unsupported distance type");
    }
    if (!ee::DistanceSafe(x, y, dimension)) {
      throw std::invalid_argument("k-opt cost needs distance-safe integer coordinates");
    }
    std::lock_guard<std::mutex> lk(s.mu);
    const ee::KoptInstance inst = ee::MakeKoptInstance(s.host_xy.size(), x, y, dimension, distance_type);
    for (int32_t i = 0; i < dimension; ++i) s.host_xy.push_back(longlong2{x[i], y[i]});
    s.instances.push_back(inst);
    s.graphs_dirty = true;
    if (graph) *graph = static_cast<uint32_t>(s.instances.size() - 1);
  });
}


ee_status ee_server_submit(ee_server* server, uint32_t graph, uint32_t k,
                           const ee_kopt_task* tasks, uint64_t count, int64_t* costs,
                           uint64_t* request_id) {
  return Guarded([&] {
    ee::Server& s = ee::AsServer(server);
    std::vector<std::shared_ptr<ee::Request>> rs{ee::MakeKoptRequest(s, graph, k, tasks, count, costs, false)};
    ee::Enqueue(s, rs, request_id);
  });
}

ee_status ee_server_submit_many(ee_server* server, uint32_t count, const ee_server_request* requests,
                                uint64_t* request_ids) {
  return Guarded([&] {
    ee::Server& s = ee::AsServer(server);
    if (count > 0 && (requests == nullptr || request_ids == nullptr)) throw std::invalid_argument("NULL argument");
    std::vector<std::shared_ptr<ee::Request>> rs(count);
    // validation is independent per request: spread it over the host cores
    ee::ParallelFor(count, [&](size_t i) {
      const ee_server_request& q = requests[i];
      if (q.kind == 0) {
        rs[i] = ee::MakeKoptRequest(s, q.graph, q.k, static_cast<const ee_kopt_task*>(q.input), q.count,
                                    static_cast<int64_t*>(q.output), true);
      } else if (q.kind == 1) {
        rs[i] = ee::MakeJvRequest(s, q.graph, static_cast<const int32_t*>(q.input), q.count,
                                  static_cast<int32_t*>(q.output));
      } else {
        throw std::invalid_argument("request kind must be 0 (leaf costs) or 1 (JV recheck)");
      }
    });
    ee::Enqueue(s, rs, request_ids);
  });
}

ee_status ee_server_add_jv_graph(ee_server* server, const ee_graph_view* graph, uint32_t* jv_graph) {
  return Guarded([&] {
    ee::Server& s = ee::AsServer(server);
    ee::CheckGraphView(graph);
    ee::JvHost h;
    h.Load(graph);  // validates and builds the CSR
    std::lock_guard<std::mutex> lk(s.mu);
    ee::Server::JvUnion& u = s.jv;
    const size_t nodes = u.xy.size() + static_cast<size_t>(h.n);
    const size_t slots = u.uv.size() + static_cast<size_t>(h.m);
    const size_t adjacency = u.neighbors.size() + h.neighbors.size();
    if (nodes >= (size_t{1} << 31) || slots >= (size_t{1} << 31) || adjacency >= (size_t{1} << 31)) {
      throw std::invalid_argument("resident JV graphs exceed int32 ids");
    }
    const int32_t nb = static_cast<int32_t>(u.xy.size()), eb = static_cast<int32_t>(u.uv.size());
    const int32_t ab = static_cast<int32_t>(u.neighbors.size());
    const uint8_t md = static_cast<uint8_t>((h.type & 1u) | (h.near ? 2u : 0u));
    for (int32_t e = 0; e < h.m; ++e) {
      u.uv.push_back(make_int2(nb + h.u[e], nb + h.v[e]));
      u.weight.push_back(h.weight[e]);
      u.mode.push_back(md);
      u.active.push_back(h.active[e] ? 1 : 0);
    }
    for (int32_t v = 0; v < h.n; ++v) u.xy.push_back(longlong2{h.x[v], h.y[v]});
    for (int32_t v = 1; v <= h.n; ++v) u.row_offsets.push_back(ab + h.row_offsets[v]);
    for (size_t j = 0; j < h.neighbors.size(); ++j) {
      u.neighbors.push_back(nb + h.neighbors[j]);
      u.edge_ids.push_back(eb + h.edge_ids[j]);
    }
    u.node_base.push_back(nb);
    u.edge_base.push_back(eb);
    u.edge_count.push_back(h.m);
    s.jv_dirty = true;
    if (jv_graph) *jv_graph = static_cast<uint32_t>(u.node_base.size() - 1);
  });
}

ee_status ee_server_submit_jv(ee_server* server, uint32_t jv_graph, const int32_t* edges,
                              uint64_t count, int32_t* witnesses, uint64_t* request_id) {
  return Guarded([&] {
    ee::Server& s = ee::AsServer(server);
    std::vector<std::shared_ptr<ee::Request>> rs{ee::MakeJvRequest(s, jv_graph, edges, count, witnesses)};
    ee::Enqueue(s, rs, request_id);
  });
}

ee_status ee_server_wait(ee_server* server, uint64_t request_id, double timeout_ms,
                         ee_request_result* result) {
  return Guarded([&] {
    ee::Server& s = ee::AsServer(server);
    std::unique_lock<std::mutex> lk(s.mu);
    auto it = s.live.find(request_id);
    if (it == s.live.end()) throw std::invalid_argument("unknown or already collected request id");
    std::shared_ptr<ee::Request> r = it->second;
    const auto ready = [&] { return r->complete; };
    if (timeout_ms < 0) {
      s.cv_done.wait(lk, ready);
    } else if (!s.cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready)) {
      throw std::runtime_error("request did not complete before the timeout");
    }
    s.live.erase(it);
    if (result) {
      result->request_id = r->id;
      result->status = r->status;
      result->batch_size = r->batch_size;
      result->batch_id = r->batch_id;
      result->queue_ms = ee::Ms(r->submit, r->issue);
      result->service_ms = ee::Ms(r->issue, r->done);
      result->latency_ms = ee::Ms(r->submit, r->done);
    }
    if (r->status != EE_OK) throw std::runtime_error(r->error);
  });
}

ee_status ee_server_wait_many(ee_server* server, uint32_t count, const uint64_t* request_ids,
                              double timeout_ms, ee_request_result* results) {
  return Guarded([&] {
    if (count > 0 && request_ids == nullptr) throw std::invalid_argument("request_ids is NULL");
    const auto deadline = std::chrono::steady_clock::now() +
                          std::chrono::duration<double, std::milli>(timeout_ms < 0 ? 0.0 : timeout_ms);
    std::string first_error;
    for (uint32_t i = 0; i < count; ++i) {
      double left = -1.0;
      if (timeout_ms >= 0) {
        left = std::max(0.0, std::chrono::duration<double, std::milli>(deadline - std::chrono::steady_clock::now()).count());
      }
      const ee_status st = ee_server_wait(server, request_ids[i], left, results ? &results[i] : nullptr);
      if (st != EE_OK && first_error.empty()) first_error = ee::LastError();
    }
    if (!first_error.empty()) throw std::runtime_error(first_error);
  });
}

ee_status ee_server_stats_get(const ee_server* server, ee_server_stats* stats) {
  return Guarded([&] {
    if (server == nullptr || stats == nullptr) throw std::invalid_argument("NULL argument");
    auto* s = reinterpret_cast<ee::Server*>(const_cast<ee_server*>(server));
    std::lock_guard<std::mutex> lk(s->mu);
    *stats = s->stats;
  });
}

}  // extern "C"
