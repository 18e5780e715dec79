// Host-side runtime types: device arrays, per-context caches, the context.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "ee_internal.h"

namespace ee {

void SetLastError(const std::string& message);
const char* LastError();
// Maps the in-flight exception onto a status code (call inside catch).
ee_status StatusFromException();
double MsSince(std::chrono::steady_clock::time_point t0);
size_t Grow(size_t current, size_t required);
bool DistanceSafe(const int64_t* x, const int64_t* y, int32_t n);
// The graph content hash in two parts (ee_graph_content_hash = edges part
// continued from the prefix part).
uint64_t ContentHashPrefix(int32_t n, uint8_t distance_type, const int64_t* x, const int64_t* y);
uint64_t ContentHashEdges(uint64_t prefix, int32_t m, const int32_t* edge_u, const int32_t* edge_v,
                          const int64_t* edge_weight, const uint8_t* edge_active);
bool NearExtent(const int64_t* x, const int64_t* y, int32_t n);
KoptInstance MakeKoptInstance(uint64_t coord_base, const int64_t* x, const int64_t* y, int32_t n,
                              uint8_t distance_type);
int CountDevices();

// Page-locked host staging of at least `count` elements (hipHostMalloc);
// growing discards the contents.  The pages are written once when allocated:
// the first host access to fresh page-locked memory faults every page in
// (about 0.4 ms per MB), which would otherwise land on the first run that
// fills or reads the staging.
template <typename T> class Pinned {
 public:
  Pinned() = default;
  ~Pinned() { Reset(); }
  Pinned(const Pinned&) = delete;
  Pinned& operator=(const Pinned&) = delete;
  Pinned(Pinned&& o) noexcept : ptr_(o.ptr_), count_(o.count_) {
    o.ptr_ = nullptr;
    o.count_ = 0;
  }
  Pinned& operator=(Pinned&& o) noexcept {
    if (this != &o) {
      Reset();
      ptr_ = o.ptr_;
      count_ = o.count_;
      o.ptr_ = nullptr;
      o.count_ = 0;
    }
    return *this;
  }
  void Reserve(size_t count) {
    if (count <= count_) return;
    Reset();
    void* p = nullptr;
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    CheckHip(hipHostMalloc(&p, bytes, hipHostMallocDefault), "hipHostMalloc");
    std::memset(p, 0, bytes);
    ptr_ = static_cast<T*>(p);
    count_ = count;
  }
  T* get() { return ptr_; }
  const T* get() const { return ptr_; }
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

// Owning device allocation of `count` elements of T on one device.
template <typename T> class DeviceArray {
 public:
  DeviceArray() = default;
  DeviceArray(size_t count, int device) { Allocate(count, device); }
  ~DeviceArray() { Reset(); }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  DeviceArray(DeviceArray&& o) noexcept { Take(&o); }
  DeviceArray& operator=(DeviceArray&& o) noexcept {
    if (this != &o) {
      Reset();
      Take(&o);
    }
    return *this;
  }

  void Allocate(size_t count, int device) {
    Reset();
    if (count == 0) return;
    if (count > SIZE_MAX / sizeof(T)) throw std::overflow_error("device allocation size overflows");
    CheckHip(hipSetDevice(device), "hipSetDevice");
    void* p = nullptr;
    const hipError_t status = hipMalloc(&p, count * sizeof(T));
    if (status != hipSuccess) {
      throw std::runtime_error(std::string("hipMalloc(") + std::to_string(count * sizeof(T)) +
                               " bytes): " + hipGetErrorString(status));
    }
    ptr_ = static_cast<T*>(p);
    count_ = count;
    device_ = device;
  }

  void Upload(const T* host, size_t count, hipStream_t stream) { UploadAt(0, host, count, stream); }
  void UploadAt(size_t offset, const T* host, size_t count, hipStream_t stream) {
    if (count == 0) return;
    if (offset > count_ || count > count_ - offset) throw std::logic_error("H2D beyond device buffer");
    CheckHip(hipMemcpyAsync(ptr_ + offset, host, count * sizeof(T), hipMemcpyHostToDevice, stream),
             "hipMemcpyAsync H2D");
  }
  void Download(T* host, size_t count, hipStream_t stream) const {
    DownloadAt(0, host, count, stream);
  }
  void DownloadAt(size_t offset, T* host, size_t count, hipStream_t stream) const {
    if (count == 0) return;
    if (offset > count_ || count > count_ - offset) throw std::logic_error("D2H beyond device buffer");
    CheckHip(hipMemcpyAsync(host, ptr_ + offset, count * sizeof(T), hipMemcpyDeviceToHost, stream),
             "hipMemcpyAsync D2H");
  }

  T* get() { return ptr_; }
  const T* get() const { return ptr_; }
  size_t count() const { return count_; }
  uint64_t bytes() const { return static_cast<uint64_t>(count_) * sizeof(T); }

 private:
  void Reset() noexcept {
    if (ptr_ != nullptr) {
      static_cast<void>(hipSetDevice(device_));
      static_cast<void>(hipFree(ptr_));
    }
    ptr_ = nullptr;
    count_ = 0;
    device_ = -1;
  }
  void Take(DeviceArray* o) noexcept {
    ptr_ = o->ptr_;
    count_ = o->count_;
    device_ = o->device_;
    o->ptr_ = nullptr;
    o->count_ = 0;
    o->device_ = -1;
  }

  T* ptr_ = nullptr;
  size_t count_ = 0;
  int device_ = -1;
};

struct JvCache {
  bool valid = false;
  int32_t dimension = 0;
  uint8_t distance_type = 0;
  bool near = false;  // NearExtent of the cached coordinates
  // exact static key (host copies)
  std::vector<int64_t> x, y, edge_weight;
  std::vector<int32_t> edge_u, edge_v;
  std::vector<int32_t> host_active;
  // static: 8m + 8m + 16n bytes
  DeviceArray<int2> d_uv;
  DeviceArray<int64_t> d_weight;
  DeviceArray<longlong2> d_xy;
  // dynamic: 4m + 4(n+1) + 4A + 4A + 4m bytes
  DeviceArray<int32_t> d_active;
  DeviceArray<int32_t> d_row_offsets;
  DeviceArray<int32_t> d_neighbors;
  DeviceArray<int32_t> d_edge_ids;
  DeviceArray<int32_t> d_witnesses;
};

struct KoptCache {
  struct TemplateEntry {
    bool valid = false;
    uint64_t hash = 0;
    std::vector<ee_endpoint_matching> table;
  };
  bool snapshot_valid = false;
  int32_t dimension = 0;
  uint8_t distance_type = 0;
  bool near = false;  // NearExtent of the snapshot
  std::vector<int64_t> x, y;
  DeviceArray<longlong2> d_xy;
  std::array<TemplateEntry, 3> templates;
  std::array<KoptPlan, 3> host_plans;
  DeviceArray<KoptPlan> d_plans;
  DeviceArray<ee_kopt_task> d_tasks;
  DeviceArray<int64_t> d_costs;
  DeviceArray<KoptTile> d_tiles;
};

// Device copy of one tour context for the deletion-set producer, kept while
// the caller passes the same tour, edge costs and selectable positions.
struct KoptCursorCache {
  bool valid = false;
  std::vector<int32_t> tour;
  std::vector<int64_t> edge_costs;
  std::vector<uint64_t> selectable;
  DeviceArray<int32_t> d_tour;
  DeviceArray<int64_t> d_edge_costs;
  DeviceArray<uint64_t> d_selectable;
  DeviceArray<uint32_t> d_positions;
  DeviceArray<uint32_t> d_status;  // bit 0: some deleted cost overflowed int64
};

// Device copy of one snapshot's coordinates and active CSR for the HT reply
// enumerations, kept while the caller passes the same arrays, plus their
// growing workspaces.
struct HtCache {
  bool valid = false;
  int32_t dimension = 0;
  uint8_t distance_type = 0;
  std::vector<int64_t> x, y;
  std::vector<int32_t> row_offsets, neighbors;
  std::vector<uint8_t> active;  // the snapshot passed the HT checks with these flags
  bool near = false;            // NearExtent: the FP64 distance path applies
  DeviceArray<longlong2> d_xy;
  DeviceArray<int32_t> d_row_offsets, d_neighbors;
  DeviceArray<int32_t> d_items;      // centers, or (endpoint, internal) task pairs
  DeviceArray<uint64_t> d_counts;    // per item
  DeviceArray<uint64_t> d_offsets;   // per item + 1
  DeviceArray<unsigned long long> d_chunk_sums;  // reply totals per scan chunk of centers
  DeviceArray<int32_t> d_replies;    // triples or pairs, exactly the last call's replies
  int last_kind = 0;                 // last completed call: 1 Hamilton, 2 end replies
  uint64_t last_total = 0;           // its reply count
};

struct Context {
  explicit Context(int requested_device);
  ~Context();
  Context(const Context&) = delete;
  Context& operator=(const Context&) = delete;

  void Activate() const;
  void Sync() const;

  int device = -1;
  int compute_units = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev_begin = nullptr, ev_end = nullptr;
  uint32_t jv_grid = 0;  // 0: one wavefront per edge slot or pair of slots
  uint32_t jv_edges_per_wave = 2;  // 2: pairs of edge slots on 32-lane halves
  KoptTuning kopt_tuning;
  JvCache jv;
  KoptCache kopt;
  KoptCursorCache cursor;
  HtCache ht;
  // DownloadToCaller's two page-locked chunks and their events, made on
  // first use
  Pinned<char> stage[2];
  hipEvent_t stage_ready[2] = {nullptr, nullptr};
};

// Copies `bytes` of device memory to caller memory and waits, after the work
// queued before it on the context's stream.  A large copy goes through the
// context's two page-locked chunks: while one chunk is in flight over PCIe
// the other is copied out on host threads.  A copy straight into pageable
// caller memory is staged by the runtime at about half that rate.
void DownloadToCaller(Context& ctx, void* dst, const void* src, size_t bytes);

// Elements [offset, offset + count) of a device array into caller memory, as
// DownloadToCaller.
template <typename T>
void DownloadRange(Context& ctx, const DeviceArray<T>& a, size_t offset, size_t count, T* host) {
  if (count == 0) return;
  if (offset > a.count() || count > a.count() - offset) throw std::logic_error("D2H beyond device buffer");
  DownloadToCaller(ctx, host, a.get() + offset, count * sizeof(T));
}

// The deletion-set producer of ee_cursor.hip: the number of sets of size k
// the reference cursor enumerates over `selectable_count` positions, the
// checks of a tour view, and the row-building launch (sets [first, first +
// count) into d_tasks and, when not NULL, their sorted positions into
// d_positions; bit 0 of *d_status is set where a deleted cost overflows).
uint64_t CursorSetCount(uint32_t selectable_count, uint32_t k);
void CheckCursorTour(const ee_tour_view* tour, int32_t n);
void LaunchCursorRows(const int32_t* d_tour, const int64_t* d_edge_costs, const uint64_t* d_selectable,
                      uint32_t tour_length, uint32_t selectable_count, uint32_t k, uint64_t first,
                      uint32_t count, ee_kopt_task* d_tasks, uint32_t* d_positions, uint32_t* d_status,
                      hipStream_t stream);

// Looks up (once per device, with that device current) the device address of
// the no-allocation scratch array of ee_kopt_eval_costs_noalloc.
void InitNoAllocScratch(int device);

std::vector<ee_endpoint_matching> CanonicalTemplates(uint32_t k, uint64_t* hash);
KoptPlan BuildPlan(uint32_t k, const ee_endpoint_matching* templates, uint32_t count);
void ValidateTasks(int32_t n, uint32_t k, const ee_kopt_task* tasks, size_t count);
// Null arrays, dimensions, distance type and integer coordinates of a view.
void CheckGraphView(const ee_graph_view* g);

// Writes the task rows of one plain evaluation into the device task pool
// (`d_tasks`, room for the evaluation's rows), enqueued on the context's
// stream.  The rows must pass ValidateTasks.
using KoptTaskSource = std::function<void(ee_kopt_task* d_tasks)>;

// The plain synchronous k-opt evaluation (snapshot, template and workspace
// caches, one launch, download) with its task rows taken from `fill_tasks`.
// Exactly one of `costs` (task_count x template_count cells) and `masks`
// (task_count x ceil(template_count / 32) improving-template words) is
// non-null.  sync_after_fill: wait for the rows before the launch (usage's
// h2d_ms is then their transfer); rows made by a kernel on the context's
// stream need no wait, and their time falls into kernel_ms.
void EvalKoptCostsFrom(Context& ctx, const int64_t* x, const int64_t* y, int32_t n, uint8_t type,
                       uint32_t k, const ee_endpoint_matching* templates, uint32_t template_count,
                       uint64_t generator_hash, size_t task_count, const KoptTaskSource& fill_tasks,
                       bool sync_after_fill, int64_t* costs, uint32_t* masks, ee_kopt_usage* usage);

}  // namespace ee
