// Leaf-cost queries served through a device work queue.
//
// A persistent kernel (KoptQueueKernel, ee_kopt.hip) drains a ring of
// requests that lives in coherent host memory.  A submit copies the task
// rows into a task arena in the same memory, writes the request's entry into
// the ring and stores the new tail; the kernel's poller sees the tail within
// a microsecond or two and hands out one ticket per tile, its tile
// wavefronts write the cells straight into a cost arena in host memory, mark
// each tile and raise the entry's completion flag.  A completing thread
// polls the flags and marks of the requests in flight, copies each served
// request's cells to the caller's buffer and wakes its waiter.  Between a
// submit and its cells there is no launch and no copy command.
//
// The kernel ends by itself after `idle_exit_ms` without a request (and
// after a few seconds in any case); a submit that finds it ended launches it
// again, and so does the completing thread when it finds the stream idle
// with requests unserved.  Every graph's coordinates stay in one device pool
// sized at creation, so adding a graph never moves memory the kernel reads.
#include <algorithm>
#include <array>
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

#include "ee_runtime.h"

namespace ee {
namespace {

using Clock = std::chrono::steady_clock;
constexpr uint32_t kTemplates[3] = {4, 25, 208};
// the launch's hard limit on its own lifetime; a queue under steady load is
// launched again this often
constexpr double kLifeMs = 2000.0;
// how often the completing thread checks that the kernel still runs
constexpr double kLivenessMs = 1.0;
// in-flight requests the completing thread looks at per pass, oldest first
constexpr size_t kCollectWindow = 512;
// how long the completing thread polls for a submit before it blocks
constexpr double kIdleSpinMs = 5.0;

double Ms(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); }

// Coherent, device-mapped host memory: the device reads and writes it over
// the bus, and the host sees the device's stores without a copy.
template <typename T> class CoherentHost {
 public:
  CoherentHost() = default;
  ~CoherentHost() {
    if (host_ != nullptr) static_cast<void>(hipHostFree(host_));
  }
  CoherentHost(const CoherentHost&) = delete;
  CoherentHost& operator=(const CoherentHost&) = delete;
  void Allocate(size_t count) {
    void* p = nullptr;
    CheckHip(hipHostMalloc(&p, std::max<size_t>(count, 1) * sizeof(T), hipHostMallocMapped | hipHostMallocCoherent),
             "hipHostMalloc (coherent)");
    host_ = static_cast<T*>(p);
    count_ = count;
    std::memset(p, 0, std::max<size_t>(count, 1) * sizeof(T));
    void* d = nullptr;
    CheckHip(hipHostGetDevicePointer(&d, p, 0), "hipHostGetDevicePointer");
    device_ = static_cast<T*>(d);
  }
  T* host() { return host_; }
  T* device() { return device_; }
  size_t count() const { return count_; }

 private:
  T* host_ = nullptr;
  T* device_ = nullptr;
  size_t count_ = 0;
};

// First-fit allocation in arrival order over a circular range; blocks may be
// freed in any order and their space returns once every older block is free.
class RingArena {
 public:
  void Init(uint64_t capacity) { capacity_ = capacity; }
  bool Allocate(uint64_t n, uint64_t* at) {
    if (n == 0 || n > capacity_) return false;
    uint64_t where = 0;
    if (!live_.empty()) {
      const uint64_t head = live_.front().at;
      const uint64_t end = live_.back().at + live_.back().n;
      if (live_.back().at >= head) {  // not wrapped: free are [end, capacity) and [0, head)
        if (end + n <= capacity_) {
          where = end;
        } else if (n <= head) {
          where = 0;
        } else {
          return false;
        }
      } else {  // wrapped: free is [end, head)
        if (end + n > head) return false;
        where = end;
      }
    }
    live_.push_back(Block{where, n, false});
    *at = where;
    return true;
  }
  void Free(uint64_t at) {
    for (Block& b : live_) {
      if (b.at == at && !b.freed) {
        b.freed = true;
        break;
      }
    }
    while (!live_.empty() && live_.front().freed) live_.pop_front();
  }
  uint64_t capacity() const { return capacity_; }

 private:
  struct Block {
    uint64_t at, n;
    bool freed;
  };
  std::deque<Block> live_;
  uint64_t capacity_ = 0;
};

// A validated request, not yet published.
struct Prepared {
  const ee_kopt_task* tasks;
  uint64_t count, cells, cell_span, tiles;
  int64_t* out;
  uint32_t k;
  KoptInstance inst;
};

struct QRequest {
  uint64_t id = 0, entry = 0, slot = 0;
  uint64_t task_at = 0, cost_at = 0, cells = 0, tiles = 0;
  uint64_t marked = 0;  // tiles whose marks the completer has seen
  int64_t* out = nullptr;
  Clock::time_point submit, done;
  std::atomic<bool> complete{false};  // collected: cells in the caller's buffer
  int32_t status = EE_OK;
  std::string error;
};

struct Queue {
  Context* ctx = nullptr;
  ee_queue_config cfg{};
  KoptTuning tuning;
  std::array<KoptPlan, 3> plans;
  DeviceArray<KoptPlan> d_plans;
  std::mutex graphs_mu;  // instances, coords and uploads into the pool
  std::vector<KoptInstance> instances;
  DeviceArray<longlong2> d_xy;
  uint64_t coords = 0;
  CoherentHost<KoptQueueHost> host;
  CoherentHost<KoptQueueEntry> ring;
  CoherentHost<uint64_t> flags;
  CoherentHost<uint64_t> marks;  // per task-arena row: tile t of an entry marks row task_at + t
  CoherentHost<ee_kopt_task> task_arena;
  CoherentHost<int64_t> cost_arena;
  DeviceArray<KoptQueueDevice> dev;
  DeviceArray<KoptQueueEntry> mirror;
  DeviceArray<uint64_t> done;
  DeviceArray<uint64_t> tickets;
  std::vector<uint64_t> slot_tiles;                // tiles ever queued per ring slot
  std::vector<std::shared_ptr<QRequest>> by_slot;  // the request holding each slot
  // published requests the completing thread has not yet taken over
  std::vector<std::shared_ptr<QRequest>> incoming;
  std::atomic<uint64_t> incoming_count{0};  // bumped with every push to `incoming`
  uint64_t completed = 0;                   // requests collected
  KoptLaunch layout{};
  size_t lds = 0;
  uint64_t idle_ticks = 0, life_ticks = 0;
  hipStream_t stream = nullptr, copy_stream = nullptr;
  RingArena task_ring, cost_ring;
  std::mutex mu;
  std::condition_variable cv_work;  // the completer: something was published, or stop
  std::condition_variable cv_room;  // publishers: arena space or a ring slot came free
  std::condition_variable cv_done;  // waiters: a request completed
  std::thread completer;
  std::atomic<bool> stopping{false};
  uint32_t epoch = 0;  // the latest launch
  uint64_t tail = 0;   // entries published
  uint64_t next_id = 1;
  std::unordered_map<uint64_t, std::shared_ptr<QRequest>> live;
  ee_queue_stats stats{};
  uint64_t inflight_cells = 0;
  std::string broken;  // non-empty: the kernel refused an entry; the queue is out of service

  uint64_t FlagOf(const QRequest& r) { return __atomic_load_n(&flags.host()[r.slot], __ATOMIC_ACQUIRE); }
  // Served: the flag, then every tile's mark (each stored after that tile's
  // cells by the wave that wrote them).  Only the completer calls this.
  bool Ready(QRequest& r) {
    const uint64_t f = FlagOf(r);
    if (f == ((r.entry + 1) | kQueueRefused)) return true;
    if (f != r.entry + 1) return false;
    const uint64_t* m = marks.host() + r.task_at;
    while (r.marked < r.tiles && __atomic_load_n(&m[r.marked], __ATOMIC_ACQUIRE) == r.entry + 1) ++r.marked;
    return r.marked == r.tiles;
  }
  // under mu
  void Launch();
  void EnsureRunning(bool stream_check);
  Prepared Prepare(uint32_t graph, uint32_t k, const ee_kopt_task* tasks, uint64_t count, int64_t* costs);
  uint64_t Publish(const Prepared& p);
  void CompleteLoop();
  void Wait(const std::shared_ptr<QRequest>& r, double timeout_ms);
  void Drain();
};

void Queue::Launch() {
  KoptQueueArgs q{};
  q.host = host.device();
  q.ring = ring.device();
  q.flags = flags.device();
  q.marks = marks.device();
  q.dev = dev.get();
  q.mirror = mirror.get();
  q.done = done.get();
  q.tickets = tickets.get();
  q.ticket_capacity = tickets.count();
  q.task_capacity = task_arena.count();
  q.cost_capacity = cost_arena.count();
  q.coord_count = d_xy.count();
  q.idle_ticks = idle_ticks;
  q.life_ticks = life_ticks;
  q.ring_size = cfg.ring_entries;
  host.host()->stop = 0;
  q.epoch = ++epoch;
  LaunchKoptQueue(layout, q, cfg.workers, lds, stream);
  ++stats.launches;
}

// Launches the kernel when the latest launch has announced its end (or, with
// stream_check, when the stream has nothing left to run).
void Queue::EnsureRunning(bool stream_check) {
  if (epoch == 0) {
    Launch();
    return;
  }
  std::atomic_thread_fence(std::memory_order_seq_cst);
  const uint64_t ended = __atomic_load_n(&host.host()->ended_epoch, __ATOMIC_SEQ_CST);
  if (ended == epoch || (stream_check && hipStreamQuery(stream) == hipSuccess)) Launch();
}

Prepared Queue::Prepare(uint32_t graph, uint32_t k, const ee_kopt_task* tasks, uint64_t count, int64_t* costs) {
  if (k < 3 || k > 5) throw std::invalid_argument("request k must be in [3,5]");
  if (count == 0 || tasks == nullptr || costs == nullptr) {
    throw std::invalid_argument("request needs at least one task and an output buffer");
  }
  const uint64_t cells = count * kTemplates[k - 3];
  const uint64_t cell_span = (cells + kSegmentAlignCells - 1) / kSegmentAlignCells * kSegmentAlignCells;
  if (count > task_arena.count() || cell_span > cost_arena.count()) {
    throw std::invalid_argument("request exceeds the queue's arenas");
  }
  KoptInstance inst{};
  {
    std::lock_guard<std::mutex> lk(graphs_mu);
    if (graph >= instances.size()) throw std::invalid_argument("unknown graph id");
    inst = instances[graph];
  }
  ValidateTasks(static_cast<int32_t>(inst.n), k, tasks, count);
  const uint32_t rows = tuning.rows_per_tile[k - 3];
  if ((count + rows - 1) / rows > kQueueMaxTiles) throw std::invalid_argument("request has too many tiles");
  return Prepared{tasks, count, cells, cell_span, (count + rows - 1) / rows, costs, k, inst};
}

uint64_t Queue::Publish(const Prepared& p) {
  const uint64_t count = p.count, cell_span = p.cell_span, tiles = p.tiles;
  const KoptInstance& inst = p.inst;
  std::unique_lock<std::mutex> lk(mu);
  auto r = std::make_shared<QRequest>();
  for (;;) {
    if (!broken.empty()) throw std::runtime_error(broken);
    if (stopping) throw std::runtime_error("queue is stopping");
    const uint64_t slot = tail % cfg.ring_entries;
    if (!by_slot[slot]) {
      uint64_t task_at = 0, cost_at = 0;
      if (task_ring.Allocate(count, &task_at)) {
        if (cost_ring.Allocate(cell_span, &cost_at)) {
          r->slot = slot;
          r->task_at = task_at;
          r->cost_at = cost_at;
          break;
        }
        task_ring.Free(task_at);
      }
    }
    // no room: the completer frees space as requests complete
    cv_room.wait_for(lk, std::chrono::milliseconds(1));
  }
  std::memcpy(task_arena.host() + r->task_at, p.tasks, count * sizeof(ee_kopt_task));
  r->entry = tail;
  r->tiles = tiles;
  r->cells = p.cells;
  r->out = p.out;
  r->id = next_id++;
  r->submit = Clock::now();
  slot_tiles[r->slot] += tiles;
  KoptQueueEntry e{};
  e.seq = r->entry;
  e.task_offset = r->task_at;
  e.cost_offset = r->cost_at;
  e.coord_base = inst.coord_base;
  e.done_target = slot_tiles[r->slot];
  e.task_count = static_cast<uint32_t>(count);
  e.tiles = static_cast<uint32_t>(tiles);
  e.k = p.k;
  e.node_count = inst.n;
  e.distance_type = inst.distance_type;
  e.near = inst.near;
  ring.host()[r->slot] = e;
  by_slot[r->slot] = r;
  incoming.push_back(r);
  incoming_count.fetch_add(1, std::memory_order_release);
  live[r->id] = r;
  inflight_cells += p.cells;
  stats.peak_inflight_cells = std::max(stats.peak_inflight_cells, inflight_cells);
  // the entry and the task rows are in place before the tail moves past them
  __atomic_store_n(&host.host()->tail, tail + 1, __ATOMIC_RELEASE);
  ++tail;
  EnsureRunning(false);
  lk.unlock();
  cv_work.notify_one();
  return r->id;
}

// The completing thread: while requests are in flight it polls their flags
// and marks (yielding the core between looks), copies each served request's
// cells to the caller's buffer, then under the lock frees the arena blocks
// and ring slots of all it served and wakes their waiters; every
// kLivenessMs without progress it makes sure the kernel runs.  The requests
// it watches are its own list: the lock is taken to take over newly
// published requests (when the published count moved) and to release.  With
// nothing in flight it polls for kIdleSpinMs before it blocks.
void Queue::CompleteLoop() {
  static_cast<void>(hipSetDevice(ctx->device));
  std::deque<std::shared_ptr<QRequest>> mine;  // oldest first
  std::vector<std::shared_ptr<QRequest>> served;
  uint64_t taken_over = 0;
  auto check = Clock::now();
  for (;;) {
    if (mine.empty()) {
      // nothing in flight: poll for the next submit a while before blocking
      // (a blocked thread wakes tens of microseconds after the notify)
      const auto idle = Clock::now();
      while (incoming_count.load(std::memory_order_acquire) == taken_over &&
             !stopping.load(std::memory_order_acquire) && Ms(idle, Clock::now()) < kIdleSpinMs) {
        std::this_thread::yield();
      }
    }
    if (mine.empty() || incoming_count.load(std::memory_order_acquire) != taken_over) {
      std::unique_lock<std::mutex> lk(mu);
      if (mine.empty()) cv_work.wait(lk, [&] { return !incoming.empty() || stopping.load(); });
      if (mine.empty() && incoming.empty()) return;  // stopping with nothing in flight
      for (auto& r : incoming) mine.push_back(std::move(r));
      taken_over += incoming.size();
      incoming.clear();
    }
    // the oldest requests first: tickets go out in entry order
    const size_t n = std::min(mine.size(), kCollectWindow);
    for (size_t i = 0; i < n; ++i) {
      QRequest& r = *mine[i];
      if (!Ready(r)) continue;
      // the blocks stay allocated until the request is released below
      if ((FlagOf(r) & kQueueRefused) == 0) {
        std::memcpy(r.out, cost_arena.host() + r.cost_at, r.cells * sizeof(int64_t));
      }
      served.push_back(mine[i]);
    }
    const auto now = Clock::now();
    if (served.empty()) {
      if (Ms(check, now) > kLivenessMs) {
        std::lock_guard<std::mutex> lk(mu);
        EnsureRunning(true);
        check = now;
      }
      std::this_thread::yield();
      continue;
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      const auto done_at = Clock::now();
      for (auto& rp : served) {
        QRequest& r = *rp;
        if ((FlagOf(r) & kQueueRefused) != 0) {
          r.status = EE_ERR_RUNTIME;
          r.error = "the device queue refused the request's entry";
          broken = r.error;
        }
        task_ring.Free(r.task_at);
        cost_ring.Free(r.cost_at);
        by_slot[r.slot].reset();
        inflight_cells -= r.cells;
        r.done = done_at;
        ++stats.requests;
        stats.cells += r.cells;
        ++completed;
        r.complete.store(true, std::memory_order_release);
      }
    }
    cv_room.notify_all();
    cv_done.notify_all();
    mine.erase(std::remove_if(mine.begin(), mine.end(),
                              [](const std::shared_ptr<QRequest>& r) {
                                return r->complete.load(std::memory_order_relaxed);
                              }),
               mine.end());
    served.clear();
    check = now;
  }
}

// Polls the request's completion (yielding) for a short while, then blocks.
void Queue::Wait(const std::shared_ptr<QRequest>& r, double timeout_ms) {
  const auto start = Clock::now();
  while (Ms(start, Clock::now()) < 0.2) {
    if (r->complete.load(std::memory_order_acquire)) return;
    std::this_thread::yield();
  }
  std::unique_lock<std::mutex> lk(mu);
  const auto ready = [&] { return r->complete.load(std::memory_order_acquire); };
  if (timeout_ms < 0) {
    cv_done.wait(lk, ready);
  } else if (!cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready)) {
    throw std::runtime_error("request did not complete before the timeout");
  }
}

// Serves everything published, stops the completer, then the kernel.
void Queue::Drain() {
  {
    std::unique_lock<std::mutex> lk(mu);
    cv_room.wait_for(lk, std::chrono::seconds(30), [&] { return completed == tail || !broken.empty(); });
    stopping = true;
  }
  cv_work.notify_all();
  if (completer.joinable()) completer.join();
  host.host()->stop = 1;
  std::atomic_thread_fence(std::memory_order_seq_cst);
  static_cast<void>(hipStreamSynchronize(stream));
}

Queue& AsQueue(ee_queue* q) {
  if (q == nullptr) throw std::invalid_argument("queue is NULL");
  return *reinterpret_cast<Queue*>(q);
}

template <typename F> ee_status Guarded(F&& body) {
  try {
    body();
    return EE_OK;
  } catch (...) {
    return StatusFromException();
  }
}

void FillResult(const QRequest& r, ee_request_result* result) {
  if (result == nullptr) return;
  result->request_id = r.id;
  result->status = r.status;
  result->batch_size = 1;
  result->batch_id = r.entry;
  result->queue_ms = 0.0;
  result->service_ms = Ms(r.submit, r.done);
  result->latency_ms = Ms(r.submit, r.done);
}

}  // namespace
}  // namespace ee

using ee::Guarded;

extern "C" {

ee_status ee_queue_create(ee_ctx* handle, const ee_queue_config* config, ee_queue** out) {
  return Guarded([&] {
    if (handle == nullptr || config == nullptr || out == nullptr) throw std::invalid_argument("NULL argument");
    *out = nullptr;
    const ee_queue_config& c = *config;
    if (c.task_rows < 1 || c.task_rows > (uint64_t{1} << 32) || c.cost_cells < 16 || c.max_nodes < 1 ||
        c.ring_entries < 1 || c.ring_entries > (1u << 20) || c.workers < 1 || c.workers > 65536 ||
        !(c.idle_exit_ms > 0.0) || c.idle_exit_ms > 1000.0) {
      throw std::invalid_argument("bad queue configuration");
    }
    auto* ctx = reinterpret_cast<ee::Context*>(handle);
    ctx->Activate();
    auto q = std::make_unique<ee::Queue>();
    q->ctx = ctx;
    q->cfg = c;
    for (uint32_t k = 3; k <= 5; ++k) {
      uint64_t hash = 0;
      const auto table = ee::CanonicalTemplates(k, &hash);
      q->plans[k - 3] = ee::BuildPlan(k, table.data(), static_cast<uint32_t>(table.size()));
    }
    const int dev = ctx->device;
    q->d_plans.Allocate(3, dev);
    q->d_plans.Upload(q->plans.data(), 3, ctx->stream);
    q->d_xy.Allocate(c.max_nodes, dev);
    q->host.Allocate(1);
    q->ring.Allocate(c.ring_entries);
    q->flags.Allocate(c.ring_entries);
    q->marks.Allocate(c.task_rows);
    q->task_arena.Allocate(c.task_rows);
    q->cost_arena.Allocate(c.cost_cells);
    q->task_ring.Init(c.task_rows);
    q->cost_ring.Init(c.cost_cells);
    q->dev.Allocate(1, dev);
    q->mirror.Allocate(c.ring_entries, dev);
    q->done.Allocate(c.ring_entries, dev);
    q->tickets.Allocate(c.task_rows, dev);  // a tile has at least one row
    ee::CheckHip(hipMemsetAsync(q->dev.get(), 0, sizeof(ee::KoptQueueDevice), ctx->stream), "hipMemsetAsync");
    ee::CheckHip(hipMemsetAsync(q->mirror.get(), 0, q->mirror.bytes(), ctx->stream), "hipMemsetAsync");
    ee::CheckHip(hipMemsetAsync(q->done.get(), 0, q->done.bytes(), ctx->stream), "hipMemsetAsync");
    ee::CheckHip(hipMemsetAsync(q->tickets.get(), 0, q->tickets.bytes(), ctx->stream), "hipMemsetAsync");
    ctx->Sync();
    q->slot_tiles.assign(c.ring_entries, 0);
    q->by_slot.resize(c.ring_entries);
    q->layout.xy = q->d_xy.get();
    q->layout.plans = q->d_plans.get();
    q->layout.tasks = q->task_arena.device();
    q->layout.costs = q->cost_arena.device();
    q->lds = ee::KoptQueueLayout(&q->layout, q->plans.data(), q->tuning);
    int khz = 0;
    ee::CheckHip(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev), "wall clock rate");
    if (khz <= 0) throw ee::Unavailable("device reports no wall clock rate");
    q->idle_ticks = static_cast<uint64_t>(c.idle_exit_ms * khz);
    q->life_ticks = static_cast<uint64_t>(ee::kLifeMs * khz);
    ee::CheckHip(hipStreamCreateWithFlags(&q->stream, hipStreamNonBlocking), "hipStreamCreate");
    ee::CheckHip(hipStreamCreateWithFlags(&q->copy_stream, hipStreamNonBlocking), "hipStreamCreate");
    ee::Queue* raw = q.get();
    raw->completer = std::thread([raw] { raw->CompleteLoop(); });
    *out = reinterpret_cast<ee_queue*>(q.release());
  });
}

void ee_queue_destroy(ee_queue* queue) {
  if (queue == nullptr) return;
  auto* q = reinterpret_cast<ee::Queue*>(queue);
  static_cast<void>(hipSetDevice(q->ctx->device));
  if (q->stream != nullptr) {
    q->Drain();
    static_cast<void>(hipStreamDestroy(q->stream));
  }
  if (q->copy_stream != nullptr) static_cast<void>(hipStreamDestroy(q->copy_stream));
  delete q;
}

ee_status ee_queue_add_graph(ee_queue* queue, const int64_t* x, const int64_t* y, int32_t dimension,
                             uint8_t distance_type, uint32_t* graph) {
  return Guarded([&] {
    ee::Queue& q = ee::AsQueue(queue);
    if (x == nullptr || y == nullptr || dimension <= 0) throw std::invalid_argument("bad coordinates");
    if (distance_type != EE_EUC_2D && distance_type != EE_CEIL_2D) {
      throw std::invalid_argument("unsupported distance type");
    }
    if (!ee::DistanceSafe(x, y, dimension)) {
      throw std::invalid_argument("k-opt cost needs distance-safe integer coordinates");
    }
    std::vector<longlong2> xy(static_cast<size_t>(dimension));
    for (int32_t i = 0; i < dimension; ++i) xy[i] = longlong2{x[i], y[i]};
    std::lock_guard<std::mutex> lk(q.graphs_mu);
    if (static_cast<uint64_t>(dimension) > q.d_xy.count() - q.coords) {
      throw std::invalid_argument("graph exceeds the queue's coordinate pool (max_nodes)");
    }
    // into unused pool space, on a stream of its own: the kernel keeps running
    q.ctx->Activate();
    q.d_xy.UploadAt(q.coords, xy.data(), xy.size(), q.copy_stream);
    ee::CheckHip(hipStreamSynchronize(q.copy_stream), "hipStreamSynchronize");
    q.instances.push_back(ee::MakeKoptInstance(q.coords, x, y, dimension, distance_type));
    q.coords += static_cast<uint64_t>(dimension);
    if (graph) *graph = static_cast<uint32_t>(q.instances.size() - 1);
  });
}

ee_status ee_queue_submit(ee_queue* queue, uint32_t graph, uint32_t k, const ee_kopt_task* tasks,
                          uint64_t count, int64_t* costs, uint64_t* request_id) {
  return Guarded([&] {
    ee::Queue& q = ee::AsQueue(queue);
    const uint64_t id = q.Publish(q.Prepare(graph, k, tasks, count, costs));
    if (request_id) *request_id = id;
  });
}

ee_status ee_queue_submit_many(ee_queue* queue, uint32_t count, const ee_server_request* requests,
                               uint64_t* request_ids) {
  return Guarded([&] {
    ee::Queue& q = ee::AsQueue(queue);
    if (count > 0 && (requests == nullptr || request_ids == nullptr)) throw std::invalid_argument("NULL argument");
    // every request is validated before any is queued
    std::vector<ee::Prepared> prepared;
    prepared.reserve(count);
    for (uint32_t i = 0; i < count; ++i) {
      const ee_server_request& r = requests[i];
      if (r.kind != 0) throw std::invalid_argument("the device queue serves leaf-cost queries (kind 0)");
      prepared.push_back(q.Prepare(r.graph, r.k, static_cast<const ee_kopt_task*>(r.input), r.count,
                                   static_cast<int64_t*>(r.output)));
    }
    for (uint32_t i = 0; i < count; ++i) request_ids[i] = q.Publish(prepared[i]);
  });
}

ee_status ee_queue_wait(ee_queue* queue, uint64_t request_id, double timeout_ms, ee_request_result* result) {
  return Guarded([&] {
    ee::Queue& q = ee::AsQueue(queue);
    std::shared_ptr<ee::QRequest> r;
    {
      std::lock_guard<std::mutex> lk(q.mu);
      auto it = q.live.find(request_id);
      if (it == q.live.end()) throw std::invalid_argument("unknown or already collected request id");
      r = it->second;
    }
    q.Wait(r, timeout_ms);
    {
      std::lock_guard<std::mutex> lk(q.mu);
      q.live.erase(request_id);
    }
    ee::FillResult(*r, result);
    if (r->status != EE_OK) throw std::runtime_error(r->error);
  });
}

ee_status ee_queue_wait_many(ee_queue* queue, uint32_t count, const uint64_t* request_ids, double timeout_ms,
                             ee_request_result* results) {
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
      const ee_status st = ee_queue_wait(queue, request_ids[i], left, results ? &results[i] : nullptr);
      if (st != EE_OK && first_error.empty()) first_error = ee::LastError();
    }
    if (!first_error.empty()) throw std::runtime_error(first_error);
  });
}

ee_status ee_queue_stats_get(const ee_queue* queue, ee_queue_stats* stats) {
  return Guarded([&] {
    if (queue == nullptr || stats == nullptr) throw std::invalid_argument("NULL argument");
    auto* q = reinterpret_cast<ee::Queue*>(const_cast<ee_queue*>(queue));
    std::lock_guard<std::mutex> lk(q->mu);
    *stats = q->stats;
  });
}

}  // extern "C"
