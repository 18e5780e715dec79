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

#include "ee_jvhost.h"
#include "ee_runtime.h"

namespace ee {

void CheckGraphView(const ee_graph_view* g);

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
// requests published per lock hold, and the threads copying their rows
constexpr size_t kPublishRun = 1024;
constexpr unsigned kPublishCopyThreads = 8;

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

// A validated request, not yet published: a leaf-cost query (k = 3..5) or
// a JV recheck (k = 0).
struct Prepared {
  uint32_t k = 0;
  uint64_t count = 0;      // task rows, or edge slots
  uint64_t rows = 0;       // task-arena rows the request holds
  uint64_t cell_span = 0;  // cost-arena cells it holds
  uint64_t cells = 0;      // leaf: output cells
  uint64_t tiles = 0;
  const ee_kopt_task* tasks = nullptr;  // leaf
  int64_t* out = nullptr;               // leaf
  KoptInstance inst{};                  // leaf
  std::vector<int32_t> slots;           // JV: union edge slots
  int32_t* witnesses = nullptr;         // JV
  int32_t node_base = 0;                // JV: the graph's first union node
  bool corrupt = false;                 // test hook: the entry's task offset past the arena
};

struct QRequest {
  uint64_t id = 0, entry = 0, slot = 0;
  uint64_t task_at = 0, cost_at = 0, cells = 0, tiles = 0, span = 0;
  uint64_t marked = 0;  // tiles whose marks the completer has seen
  int64_t* out = nullptr;
  bool jv = false;
  uint64_t edges = 0;
  int32_t* witnesses = nullptr;
  int32_t node_base = 0;
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
  std::mutex graphs_mu;  // instances, coords, the JV union and uploads into them
  std::vector<KoptInstance> instances;
  DeviceArray<longlong2> d_xy;
  uint64_t coords = 0;
  // JV snapshots back to back (node ids, edge slots and CSR offset by graph),
  // sized at creation; zero-filled, so every slot below the capacity reads
  // in-bounds data
  DeviceArray<int2> jv_uv;
  DeviceArray<int64_t> jv_weight;
  DeviceArray<uint8_t> jv_mode;
  DeviceArray<longlong2> jv_xy;
  DeviceArray<int32_t> jv_active, jv_row_offsets, jv_neighbors, jv_edge_ids, jv_stage;
  uint64_t jv_nodes = 0, jv_slots = 0, jv_adjacency = 0;
  std::vector<int32_t> jv_node_base, jv_edge_base, jv_edge_count;
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
  void StopKernel();
  Prepared Prepare(uint32_t graph, uint32_t k, const ee_kopt_task* tasks, uint64_t count, int64_t* costs);
  Prepared PrepareJv(uint32_t jv_graph, const int32_t* edges, uint64_t count, int32_t* witnesses);
  void Publish(const Prepared* ps, size_t n, uint64_t* ids, unsigned copy_threads = 1);
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
  if (jv_uv.count() > 0) {
    q.jv.edge_uv = jv_uv.get();
    q.jv.edge_weight = jv_weight.get();
    q.jv.edge_active = jv_active.get();
    q.jv.row_offsets = jv_row_offsets.get();
    q.jv.neighbors = jv_neighbors.get();
    q.jv.csr_edge_ids = jv_edge_ids.get();
    q.jv.xy = jv_xy.get();
    q.jv.edge_mode = jv_mode.get();
    q.jv_slot_count = jv_uv.count();
    q.jv_stage = jv_stage.get();
  }
  host.host()->stop = 0;
  q.epoch = ++epoch;
  LaunchKoptQueue(layout, q, cfg.workers, lds, stream);
  ++stats.launches;
}

// Ends the running launch (its poller publishes what the host has queued and
// its tile waves serve it), so that data uploaded next is read by a new
// dispatch; the next submit launches again.
void Queue::StopKernel() {
  if (epoch == 0) return;
  host.host()->stop = 1;
  std::atomic_thread_fence(std::memory_order_seq_cst);
  CheckHip(hipStreamSynchronize(stream), "hipStreamSynchronize (queue kernel)");
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
  Prepared p;
  p.k = k;
  p.count = count;
  p.rows = count;
  p.cell_span = cell_span;
  p.cells = cells;
  p.tiles = (count + rows - 1) / rows;
  p.tasks = tasks;
  p.out = costs;
  p.inst = inst;
  return p;
}

Prepared Queue::PrepareJv(uint32_t jv_graph, const int32_t* edges, uint64_t count, int32_t* witnesses) {
  if (count == 0 || edges == nullptr || witnesses == nullptr) {
    throw std::invalid_argument("a JV recheck needs at least one edge slot and an output buffer");
  }
  Prepared p;
  p.k = 0;
  p.count = count;
  p.tiles = (count + kQueueJvTileEdges - 1) / kQueueJvTileEdges;
  p.rows = p.tiles;  // one mark per tile; 12 slots fit a row, the tile takes 4
  p.cell_span = ((count + 1) / 2 + kSegmentAlignCells - 1) / kSegmentAlignCells * kSegmentAlignCells;
  if (p.tiles > kQueueMaxTiles || p.rows > task_arena.count() || p.cell_span > cost_arena.count()) {
    throw std::invalid_argument("JV recheck exceeds the queue's arenas");
  }
  int32_t eb = 0, m = 0;
  {
    std::lock_guard<std::mutex> lk(graphs_mu);
    if (jv_graph >= jv_node_base.size()) throw std::invalid_argument("unknown JV graph id");
    eb = jv_edge_base[jv_graph];
    m = jv_edge_count[jv_graph];
    p.node_base = jv_node_base[jv_graph];
  }
  p.slots.resize(count);
  for (uint64_t i = 0; i < count; ++i) {
    if (edges[i] < 0 || edges[i] >= m) throw std::invalid_argument("JV recheck edge slot out of range");
    p.slots[i] = eb + edges[i];
  }
  p.witnesses = witnesses;
  return p;
}

// Publishes ps[0..n) in order: as many per lock hold as the ring and the
// arenas take (waiting for room otherwise), their rows copied into the task
// arena on up to `copy_threads` threads, then one tail store for the run.
void Queue::Publish(const Prepared* ps, size_t n, uint64_t* ids, unsigned copy_threads) {
  size_t done = 0;
  std::vector<std::shared_ptr<QRequest>> run;
  while (done < n) {
    std::unique_lock<std::mutex> lk(mu);
    run.clear();
    for (;;) {
      if (!broken.empty()) throw std::runtime_error(broken);
      if (stopping) throw std::runtime_error("queue is stopping");
      while (done + run.size() < n && run.size() < kPublishRun) {
        const Prepared& p = ps[done + run.size()];
        const uint64_t slot = (tail + run.size()) % cfg.ring_entries;
        if (by_slot[slot]) break;
        uint64_t task_at = 0, cost_at = 0;
        if (!task_ring.Allocate(p.rows, &task_at)) break;
        if (!cost_ring.Allocate(p.cell_span, &cost_at)) {
          task_ring.Free(task_at);
          break;
        }
        auto r = std::make_shared<QRequest>();
        r->slot = slot;
        r->task_at = task_at;
        r->cost_at = cost_at;
        by_slot[slot] = r;  // holds the slot; published below
        run.push_back(std::move(r));
      }
      if (!run.empty()) break;
      // no room: the completer frees space as requests complete
      cv_room.wait_for(lk, std::chrono::milliseconds(1));
    }
    const auto copy = [&](size_t j) {
      const Prepared& p = ps[done + j];
      if (p.k == 0) {
        std::memcpy(reinterpret_cast<int32_t*>(task_arena.host() + run[j]->task_at), p.slots.data(),
                    p.count * sizeof(int32_t));
      } else {
        std::memcpy(task_arena.host() + run[j]->task_at, p.tasks, p.count * sizeof(ee_kopt_task));
      }
    };
    if (run.size() > 1 && copy_threads > 1) {
      ParallelFor(run.size(), copy, copy_threads);
    } else {
      for (size_t j = 0; j < run.size(); ++j) copy(j);
    }
    const auto now = Clock::now();
    for (size_t j = 0; j < run.size(); ++j) {
      const Prepared& p = ps[done + j];
      QRequest& r = *run[j];
      const bool jv = p.k == 0;
      r.entry = tail + j;
      r.tiles = p.tiles;
      r.cells = p.cells;
      r.span = p.cell_span;
      r.out = p.out;
      r.jv = jv;
      r.edges = jv ? p.count : 0;
      r.witnesses = p.witnesses;
      r.node_base = p.node_base;
      r.id = next_id++;
      r.submit = now;
      slot_tiles[r.slot] += p.tiles;
      KoptQueueEntry e{};
      e.seq = r.entry;
      e.task_offset = r.task_at;
      e.cost_offset = r.cost_at;
      e.done_target = slot_tiles[r.slot];
      e.task_count = static_cast<uint32_t>(p.count);
      e.tiles = static_cast<uint32_t>(p.tiles);
      e.k = p.k;
      if (!jv) {
        e.coord_base = p.inst.coord_base;
        e.node_count = p.inst.n;
        e.distance_type = p.inst.distance_type;
        e.near = p.inst.near;
      }
      if (p.corrupt) e.task_offset = task_arena.count();
      ring.host()[r.slot] = e;
      incoming.push_back(run[j]);
      live[r.id] = run[j];
      inflight_cells += p.cell_span;
      if (ids) ids[done + j] = r.id;
    }
    stats.peak_inflight_cells = std::max(stats.peak_inflight_cells, inflight_cells);
    incoming_count.fetch_add(run.size(), std::memory_order_release);
    // the entries and their task rows are in place before the tail moves past them
    __atomic_store_n(&host.host()->tail, tail + run.size(), __ATOMIC_RELEASE);
    tail += run.size();
    EnsureRunning(false);
    lk.unlock();
    cv_work.notify_one();
    done += run.size();
  }
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
        if (r.jv) {
          // union node ids back to the graph's own ids
          const int32_t* w = reinterpret_cast<const int32_t*>(cost_arena.host() + r.cost_at);
          for (uint64_t j = 0; j < r.edges; ++j) r.witnesses[j] = w[j] >= 0 ? w[j] - r.node_base : -1;
        } else {
          std::memcpy(r.out, cost_arena.host() + r.cost_at, r.cells * sizeof(int64_t));
        }
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
        inflight_cells -= r.span;
        r.done = done_at;
        ++stats.requests;
        stats.cells += r.cells;
        stats.jv_edges += r.edges;
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
        !(c.idle_exit_ms > 0.0) || c.idle_exit_ms > 1000.0 || c.max_jv_edges >= (uint64_t{1} << 31) ||
        c.max_jv_nodes >= (uint64_t{1} << 31) || c.max_jv_adjacency >= (uint64_t{1} << 31) ||
        (c.max_jv_edges > 0 && (c.max_jv_nodes == 0 || c.max_jv_adjacency == 0))) {
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
    if (c.max_jv_edges > 0) {
      q->jv_uv.Allocate(c.max_jv_edges, dev);
      q->jv_weight.Allocate(c.max_jv_edges, dev);
      q->jv_mode.Allocate(c.max_jv_edges, dev);
      q->jv_active.Allocate(c.max_jv_edges, dev);
      q->jv_xy.Allocate(c.max_jv_nodes, dev);
      q->jv_row_offsets.Allocate(c.max_jv_nodes + 1, dev);
      q->jv_neighbors.Allocate(c.max_jv_adjacency, dev);
      q->jv_edge_ids.Allocate(c.max_jv_adjacency, dev);
      q->jv_stage.Allocate((static_cast<size_t>(c.workers) + 1) * 2 * ee::kQueueJvTileEdges, dev);
      const auto zero = [&](void* ptr, size_t bytes) {
        ee::CheckHip(hipMemsetAsync(ptr, 0, bytes, ctx->stream), "hipMemsetAsync");
      };
      zero(q->jv_uv.get(), q->jv_uv.bytes());
      zero(q->jv_weight.get(), q->jv_weight.bytes());
      zero(q->jv_mode.get(), q->jv_mode.bytes());
      zero(q->jv_active.get(), q->jv_active.bytes());
      zero(q->jv_xy.get(), q->jv_xy.bytes());
      zero(q->jv_row_offsets.get(), q->jv_row_offsets.bytes());
      zero(q->jv_neighbors.get(), q->jv_neighbors.bytes());
      zero(q->jv_edge_ids.get(), q->jv_edge_ids.bytes());
      zero(q->jv_stage.get(), q->jv_stage.bytes());
    }
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
    std::lock_guard<std::mutex> run(q.mu);
    std::lock_guard<std::mutex> lk(q.graphs_mu);
    if (static_cast<uint64_t>(dimension) > q.d_xy.count() - q.coords) {
      throw std::invalid_argument("graph exceeds the queue's coordinate pool (max_nodes)");
    }
    // into unused pool space; a running launch ends first (serving what is
    // queued), so that only a new dispatch reads the new coordinates
    q.ctx->Activate();
    q.StopKernel();
    q.d_xy.UploadAt(q.coords, xy.data(), xy.size(), q.copy_stream);
    ee::CheckHip(hipStreamSynchronize(q.copy_stream), "hipStreamSynchronize");
    q.instances.push_back(ee::MakeKoptInstance(q.coords, x, y, dimension, distance_type));
    q.coords += static_cast<uint64_t>(dimension);
    if (graph) *graph = static_cast<uint32_t>(q.instances.size() - 1);
  });
}

ee_status ee_queue_add_jv_graph(ee_queue* queue, const ee_graph_view* graph, uint32_t* jv_graph) {
  return Guarded([&] {
    ee::Queue& q = ee::AsQueue(queue);
    if (q.jv_uv.count() == 0) throw std::invalid_argument("the queue was created without JV capacity");
    ee::CheckGraphView(graph);
    ee::JvHost h;
    h.Load(graph);  // validates and builds the CSR
    std::lock_guard<std::mutex> run(q.mu);
    std::lock_guard<std::mutex> lk(q.graphs_mu);
    const uint64_t n = static_cast<uint64_t>(h.n), m = static_cast<uint64_t>(h.m), adj = h.neighbors.size();
    if (n > q.jv_xy.count() - q.jv_nodes || m > q.jv_uv.count() - q.jv_slots ||
        adj > q.jv_neighbors.count() - q.jv_adjacency) {
      throw std::invalid_argument("JV graph exceeds the queue's JV capacity");
    }
    const int32_t nb = static_cast<int32_t>(q.jv_nodes), eb = static_cast<int32_t>(q.jv_slots);
    const int32_t ab = static_cast<int32_t>(q.jv_adjacency);
    const uint8_t md = static_cast<uint8_t>((h.type & 1u) | (h.near ? 2u : 0u));
    std::vector<int2> uv(m);
    std::vector<uint8_t> mode(m, md);
    std::vector<int32_t> active(m), rows(n), nbrs(adj), eids(adj);
    std::vector<longlong2> xy(n);
    for (uint64_t e = 0; e < m; ++e) {
      uv[e] = make_int2(nb + h.u[e], nb + h.v[e]);
      active[e] = h.active[e] ? 1 : 0;
    }
    for (uint64_t v = 0; v < n; ++v) {
      xy[v] = longlong2{h.x[v], h.y[v]};
      rows[v] = ab + h.row_offsets[v + 1];
    }
    for (uint64_t j = 0; j < adj; ++j) {
      nbrs[j] = nb + h.neighbors[j];
      eids[j] = eb + h.edge_ids[j];
    }
    // into unused union space; a running launch ends first, so that only a
    // new dispatch reads the new snapshot
    q.ctx->Activate();
    q.StopKernel();
    const hipStream_t s = q.copy_stream;
    q.jv_uv.UploadAt(q.jv_slots, uv.data(), m, s);
    q.jv_weight.UploadAt(q.jv_slots, h.weight.data(), m, s);
    q.jv_mode.UploadAt(q.jv_slots, mode.data(), m, s);
    q.jv_active.UploadAt(q.jv_slots, active.data(), m, s);
    q.jv_xy.UploadAt(q.jv_nodes, xy.data(), n, s);
    q.jv_row_offsets.UploadAt(q.jv_nodes + 1, rows.data(), n, s);
    q.jv_neighbors.UploadAt(q.jv_adjacency, nbrs.data(), adj, s);
    q.jv_edge_ids.UploadAt(q.jv_adjacency, eids.data(), adj, s);
    ee::CheckHip(hipStreamSynchronize(s), "hipStreamSynchronize");
    q.jv_node_base.push_back(nb);
    q.jv_edge_base.push_back(eb);
    q.jv_edge_count.push_back(h.m);
    q.jv_nodes += n;
    q.jv_slots += m;
    q.jv_adjacency += adj;
    if (jv_graph) *jv_graph = static_cast<uint32_t>(q.jv_node_base.size() - 1);
  });
}

ee_status ee_queue_submit_jv(ee_queue* queue, uint32_t jv_graph, const int32_t* edges, uint64_t count,
                             int32_t* witnesses, uint64_t* request_id) {
  return Guarded([&] {
    ee::Queue& q = ee::AsQueue(queue);
    const ee::Prepared p = q.PrepareJv(jv_graph, edges, count, witnesses);
    uint64_t id = 0;
    q.Publish(&p, 1, &id);
    if (request_id) *request_id = id;
  });
}

ee_status ee_queue_inject_refusal(ee_queue* queue, uint32_t graph, const ee_kopt_task* task, int64_t* costs,
                                  uint64_t* request_id) {
  return Guarded([&] {
    ee::Queue& q = ee::AsQueue(queue);
    ee::Prepared p = q.Prepare(graph, 3, task, 1, costs);
    p.corrupt = true;
    uint64_t id = 0;
    q.Publish(&p, 1, &id);
    if (request_id) *request_id = id;
  });
}

ee_status ee_queue_submit(ee_queue* queue, uint32_t graph, uint32_t k, const ee_kopt_task* tasks,
                          uint64_t count, int64_t* costs, uint64_t* request_id) {
  return Guarded([&] {
    ee::Queue& q = ee::AsQueue(queue);
    const ee::Prepared p = q.Prepare(graph, k, tasks, count, costs);
    uint64_t id = 0;
    q.Publish(&p, 1, &id);
    if (request_id) *request_id = id;
  });
}

ee_status ee_queue_submit_many(ee_queue* queue, uint32_t count, const ee_server_request* requests,
                               uint64_t* request_ids) {
  return Guarded([&] {
    ee::Queue& q = ee::AsQueue(queue);
    if (count > 0 && (requests == nullptr || request_ids == nullptr)) throw std::invalid_argument("NULL argument");
    // every request is validated (spread over the host cores) before any is queued
    std::vector<ee::Prepared> prepared(count);
    ee::ParallelFor(count, [&](size_t i) {
      const ee_server_request& r = requests[i];
      if (r.kind == 0) {
        prepared[i] = q.Prepare(r.graph, r.k, static_cast<const ee_kopt_task*>(r.input), r.count,
                                static_cast<int64_t*>(r.output));
      } else if (r.kind == 1) {
        prepared[i] = q.PrepareJv(r.graph, static_cast<const int32_t*>(r.input), r.count,
                                  static_cast<int32_t*>(r.output));
      } else {
        throw std::invalid_argument("request kind must be 0 (leaf costs) or 1 (JV recheck)");
      }
    });
    q.Publish(prepared.data(), prepared.size(), request_ids, ee::kPublishCopyThreads);
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
