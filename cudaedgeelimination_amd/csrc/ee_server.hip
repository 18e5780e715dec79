// Serving loop for leaf-cost queries over device-resident graph snapshots.
//
// Callers submit requests (graph, k, task rows, output buffer) from any
// thread.  One worker thread forms batches from the queue in arrival order
// under a cell limit, a request limit and an HBM budget, and issues each
// batch on one of `pipeline_depth` slots (own stream, device buffers and
// pinned staging): upload tasks and tables, one launch of the tiled k-opt
// kernel over all the batch's (graph, k) segments, download the cells.  While
// a batch is in flight the worker keeps admitting and issuing into a free
// slot, so requests that arrive meanwhile overlap on the device.  Completion
// copies each request's cells into the caller's buffer and wakes waiters.
#include <algorithm>
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

#include "ee_runtime.h"

namespace ee {
namespace {

using Clock = std::chrono::steady_clock;
constexpr uint32_t kTemplates[3] = {4, 25, 208};

double Ms(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

template <typename T> class Pinned {
 public:
  ~Pinned() { Reset(); }
  void Reserve(size_t count) {
    if (count <= count_) return;
    Reset();
    void* p = nullptr;
    CheckHip(hipHostMalloc(&p, std::max<size_t>(count, 1) * sizeof(T), hipHostMallocDefault),
             "hipHostMalloc");
    ptr_ = static_cast<T*>(p);
    count_ = count;
  }
  T* get() { return ptr_; }
  size_t count() const { return count_; }

 private:
  void Reset() {
    if (ptr_ != nullptr) static_cast<void>(hipHostFree(ptr_));
    ptr_ = nullptr;
    count_ = 0;
  }
  T* ptr_ = nullptr;
  size_t count_ = 0;
};

struct Request {
  uint64_t id = 0;
  uint32_t graph = 0, k = 0;
  std::vector<ee_kopt_task> tasks;
  int64_t* out = nullptr;
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
  // queue
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> pending;
  std::unordered_map<uint64_t, std::shared_ptr<Request>> live;
  uint64_t next_id = 1, next_batch = 1;
  bool stopping = false;
  ee_server_stats stats{};
  uint64_t inflight_bytes = 0;
  std::vector<Slot> slots;
  std::thread worker;

  uint64_t SlotBudget() const { return cfg.hbm_budget_bytes / std::max<uint32_t>(cfg.pipeline_depth, 1); }

  void Loop();
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
    {
      std::lock_guard<std::mutex> lk(mu);
      if (graphs_dirty) {
        xy_copy = host_xy;
        inst_copy = instances;
        graphs_dirty = false;
      }
    }
    if (!inst_copy.empty()) {
      // new graphs: wait for every batch that reads the old pool, then replace it
      for (Slot& s : slots) {
        if (s.busy) CheckHip(hipStreamSynchronize(s.stream), "hipStreamSynchronize");
      }
      d_xy.Allocate(xy_copy.size(), ctx->device);
      d_xy.Upload(xy_copy.data(), xy_copy.size(), slot.stream);
      CheckHip(hipStreamSynchronize(slot.stream), "hipStreamSynchronize");
      worker_instances = std::move(inst_copy);
    }
    uint64_t rows = 0, cells = 0, tiles = 0;
    bool present[3] = {false, false, false};
    slot.offsets.clear();
    for (const auto& r : batch) {
      slot.offsets.push_back(cells);
      rows += r->tasks.size();
      cells += r->cells;
      tiles += (r->tasks.size() + tuning.rows_per_tile[r->k - 3] - 1) / tuning.rows_per_tile[r->k - 3];
      present[r->k - 3] = true;
    }
    slot.h_tasks.Reserve(rows);
    slot.h_costs.Reserve(cells);
    slot.segments.resize(batch.size());
    slot.h_tiles.Reserve(tiles);
    if (slot.d_tasks.count() < rows) slot.d_tasks.Allocate(Grow(slot.d_tasks.count(), rows), ctx->device);
    if (slot.d_costs.count() < cells) slot.d_costs.Allocate(Grow(slot.d_costs.count(), cells), ctx->device);
    if (slot.d_tiles.count() < tiles) slot.d_tiles.Allocate(Grow(slot.d_tiles.count(), tiles), ctx->device);
    uint64_t row = 0;
    for (size_t i = 0; i < batch.size(); ++i) {
      const Request& r = *batch[i];
      std::memcpy(slot.h_tasks.get() + row, r.tasks.data(), r.tasks.size() * sizeof(ee_kopt_task));
      slot.segments[i] = KoptSegment{r.graph, r.k, row, r.tasks.size(), slot.offsets[i], 0, 0, 0};
      row += r.tasks.size();
    }
    KoptLaunch a{};
    KoptAssignTiles(slot.segments.data(), batch.size(), worker_instances.data(), plans.data(), tuning,
                    slot.h_tiles.get(), &a);
    if (a.tile_count != tiles) throw std::logic_error("server tile count mismatch");
    slot.d_tasks.Upload(slot.h_tasks.get(), rows, slot.stream);
    slot.d_tiles.Upload(slot.h_tiles.get(), tiles, slot.stream);
    a.xy = d_xy.get();
    a.tiles = slot.d_tiles.get();
    a.tasks = slot.d_tasks.get();
    a.plans = d_plans.get();
    a.costs = slot.d_costs.get();
    const size_t shared = KoptPrepareLayout(&a, plans.data(), present, tuning);
    CheckHip(hipEventRecord(slot.ev_begin, slot.stream), "event record");
    LaunchKopt(a, tuning, shared, slot.stream);
    CheckHip(hipEventRecord(slot.ev_end, slot.stream), "event record");
    slot.d_costs.Download(slot.h_costs.get(), cells, slot.stream);
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
    slot.requests = std::move(batch);
    slot.busy = true;
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
    std::memcpy(r.out, slot.h_costs.get() + slot.offsets[i], r.cells * sizeof(int64_t));
  }
  std::lock_guard<std::mutex> lk(mu);
  const auto now = Clock::now();
  for (auto& r : slot.requests) {
    r->done = now;
    r->complete = true;
    stats.cells += r->cells;
  }
  stats.requests += slot.requests.size();
  stats.batches += 1;
  stats.device_ms += ms;
  inflight_bytes -= slot.bytes;
  slot.requests.clear();
  slot.busy = false;
  cv_done.notify_all();
}

void Server::Loop() {
  const uint64_t budget = SlotBudget();
  for (;;) {
    Slot* free_slot = nullptr;
    Slot* oldest = nullptr;
    for (Slot& s : slots) {
      if (!s.busy && free_slot == nullptr) free_slot = &s;
      if (s.busy && (oldest == nullptr || s.batch_id < oldest->batch_id)) oldest = &s;
    }
    std::vector<std::shared_ptr<Request>> batch;
    {
      std::unique_lock<std::mutex> lk(mu);
      if (free_slot != nullptr && pending.empty() && oldest == nullptr) {
        cv_work.wait(lk, [&] { return stopping || !pending.empty(); });
      }
      if (stopping && pending.empty() && oldest == nullptr) return;
      if (free_slot != nullptr && !pending.empty()) {
        uint64_t cells = 0, bytes = 0;
        while (!pending.empty() && batch.size() < cfg.max_batch_requests) {
          const auto& r = pending.front();
          if (!batch.empty() &&
              (cells + r->cells > cfg.max_batch_cells || bytes + r->bytes > budget)) {
            break;
          }
          cells += r->cells;
          bytes += r->bytes;
          batch.push_back(r);
          pending.pop_front();
        }
      }
    }
    if (!batch.empty()) {
      Issue(*free_slot, std::move(batch));
      continue;
    }
    if (oldest != nullptr) {
      // wait briefly for the oldest batch, waking early for new requests
      const hipError_t q = hipEventQuery(oldest->ev_done);
      if (q == hipSuccess) {
        Complete(*oldest);
      } else if (q == hipErrorNotReady) {
        if (free_slot == nullptr) {
          static_cast<void>(hipEventSynchronize(oldest->ev_done));
          Complete(*oldest);
        } else {
          std::unique_lock<std::mutex> lk(mu);
          cv_work.wait_for(lk, std::chrono::microseconds(20),
                           [&] { return stopping || !pending.empty(); });
        }
      } else {
        std::vector<std::shared_ptr<Request>> failed = std::move(oldest->requests);
        oldest->busy = false;
        Fail(failed, std::string("batch failed: ") + hipGetErrorString(q));
      }
    }
  }
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
    for (ee::Slot& slot : s->slots) {
      ee::CheckHip(hipStreamCreateWithFlags(&slot.stream, hipStreamNonBlocking), "hipStreamCreate");
      ee::CheckHip(hipEventCreate(&slot.ev_begin), "hipEventCreate");
      ee::CheckHip(hipEventCreate(&slot.ev_end), "hipEventCreate");
      ee::CheckHip(hipEventCreateWithFlags(&slot.ev_done, hipEventDisableTiming), "hipEventCreate");
    }
    ee::Server* raw = s.get();
    raw->worker = std::thread([raw] { raw->Loop(); });
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
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
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
    if (k < 3 || k > 5) throw std::invalid_argument("request k must be in [3,5]");
    if (count == 0 || tasks == nullptr || costs == nullptr) {
      throw std::invalid_argument("request needs at least one task and an output buffer");
    }
    auto r = std::make_shared<ee::Request>();
    r->graph = graph;
    r->k = k;
    r->cells = count * ee::kTemplates[k - 3];
    r->bytes = r->cells * sizeof(int64_t) + count * sizeof(ee_kopt_task);
    r->out = costs;
    if (r->bytes > s.SlotBudget()) throw std::invalid_argument("request exceeds the HBM budget");
    uint32_t n = 0;
    {
      std::lock_guard<std::mutex> lk(s.mu);
      if (graph >= s.instances.size()) throw std::invalid_argument("unknown graph id");
      n = s.instances[graph].n;
    }
    ee::ValidateTasks(static_cast<int32_t>(n), k, tasks, count);
    r->tasks.assign(tasks, tasks + count);
    {
      std::lock_guard<std::mutex> lk(s.mu);
      if (s.stopping) throw std::runtime_error("server is stopping");
      r->id = s.next_id++;
      r->submit = ee::Clock::now();
      s.live[r->id] = r;
      s.pending.push_back(r);
    }
    s.cv_work.notify_one();
    if (request_id) *request_id = r->id;
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

ee_status ee_server_stats_get(const ee_server* server, ee_server_stats* stats) {
  return Guarded([&] {
    if (server == nullptr || stats == nullptr) throw std::invalid_argument("NULL argument");
    auto* s = reinterpret_cast<ee::Server*>(const_cast<ee_server*>(server));
    std::lock_guard<std::mutex> lk(s->mu);
    *stats = s->stats;
  });
}

}  // extern "C"
