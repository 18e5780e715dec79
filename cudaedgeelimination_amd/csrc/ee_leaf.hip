// Certified leaf pipeline (ee_leaf_*): the per-k cost loop of the
// reference's leaf search (ProvePathSystemsByKOptImpl,
// src/cpu/kopt_search.cpp:2651-2764) with the device and the host working on
// different blocks.
//
// The reference runs, per block of task rows: device evaluation, a blocking
// download of every cell, a full CPU matrix of the same rows and a cell by
// cell comparison (EvaluateKOptTemplateCosts, kopt_search.cpp:1455-1475),
// then its consumer scans each row for templates cheaper than the deleted
// edges (TryReconnectFromCostRow, kopt_search.cpp:677-734).  Here each
// submitted block owns a slot: device task rows (uploaded, or built by the
// cursor kernel), a cost buffer, and pinned host copies of rows, positions
// and cells.  A block's launches and downloads are enqueued on one of two
// streams when it is submitted, so while ee_leaf_next certifies block i on
// the host threads the device is already computing and downloading the
// blocks after it.  Certification is the same rule (every cell recomputed on
// the host and compared); device-built rows are also rebuilt on the host from
// the cursor's enumeration order and compared, so no device output reaches
// the consumer unchecked.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "ee_certify.h"
#include "ee_runtime.h"

namespace ee {
namespace {

constexpr uint32_t kMaxLeafDepth = 4;
constexpr uint32_t kMaxLeafRows = 65536;
constexpr uint32_t kMaxTemplatesPerRow = 208;  // canonical k = 5

struct LeafSlot {
  bool queued = false;
  bool cursor = false;
  uint32_t k = 0;
  uint32_t rows = 0;
  uint64_t tag = 0;
  uint64_t first = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev_start = nullptr, ev_kernel = nullptr, ev_done = nullptr;
  DeviceArray<ee_kopt_task> d_tasks;
  DeviceArray<uint32_t> d_positions;
  DeviceArray<int64_t> d_costs;
  DeviceArray<uint32_t> d_status;
  Pinned<ee_kopt_task> h_tasks;
  Pinned<uint32_t> h_positions;
  Pinned<int64_t> h_costs;
  Pinned<uint32_t> h_status;
  std::vector<uint32_t> improving;
};

// The tour context of cursor blocks: host copy (for the row rebuild) and
// device copy (for the cursor kernel).
struct LeafTour {
  bool valid = false;
  std::vector<int32_t> tour;
  std::vector<int64_t> edge_costs;
  std::vector<uint64_t> selectable;
  DeviceArray<int32_t> d_tour;
  DeviceArray<int64_t> d_edge_costs;
  DeviceArray<uint64_t> d_selectable;
};

// C(n, j) for j <= 4 and n < 2^16 (exact in 128-bit intermediates).
uint64_t Choose(uint64_t n, uint32_t j) {
  if (n < j) return 0;
  unsigned __int128 c = 1;
  for (uint32_t i = 0; i < j; ++i) c = c * (n - i) / (i + 1);
  return static_cast<uint64_t>(c);
}

// The cursor's combination of rank `rank` (index 0 fixed at 0, the other k-1
// indices increasing in [1, items)).  Slot s after index `prev` takes the
// largest v whose smaller candidates u (each the root of C(items - 1 - u,
// rem) sets, rem = slots after s) number at most the rank left; those sum to
// C(items - 1 - prev, rem + 1) - C(items - v, rem + 1).
std::vector<uint64_t> UnrankCombination(uint64_t rank, uint32_t k, uint64_t items) {
  std::vector<uint64_t> c(k, 0);
  uint64_t prev = 0;
  for (uint32_t s = 1; s < k; ++s) {
    const uint32_t rem = k - 1 - s;
    const uint64_t all = Choose(items - 1 - prev, rem + 1);
    const auto below = [&](uint64_t v) { return all - Choose(items - v, rem + 1); };
    uint64_t lo = prev + 1, hi = items - 1 - rem;
    while (lo < hi) {
      const uint64_t mid = lo + (hi - lo + 1) / 2;
      if (below(mid) <= rank) {
        lo = mid;
      } else {
        hi = mid - 1;
      }
    }
    rank -= below(lo);
    c[s] = lo;
    prev = lo;
  }
  return c;
}

// AdvanceCombination (kopt_search.cpp:736-752): index 0 never moves.
bool Advance(std::vector<uint64_t>* c, uint64_t items) {
  const size_t k = c->size();
  size_t p = k - 1;
  while (p >= 1 && (*c)[p] == items - (k - p)) {
    if (p == 1) return false;
    --p;
  }
  ++(*c)[p];
  for (size_t q = p + 1; q < k; ++q) (*c)[q] = (*c)[q - 1] + 1;
  return true;
}

struct Leaf {
  Leaf(Context* c, uint32_t d, uint32_t rows) : ctx(c), depth(d), max_rows(rows), pool(0) {}
  ~Leaf() {
    static_cast<void>(hipSetDevice(ctx->device));
    for (hipStream_t s : streams) {
      if (s != nullptr) static_cast<void>(hipStreamSynchronize(s));
    }
    for (LeafSlot& s : slots) {
      for (hipEvent_t e : {s.ev_start, s.ev_kernel, s.ev_done}) {
        if (e != nullptr) static_cast<void>(hipEventDestroy(e));
      }
    }
    slots.clear();
    for (hipStream_t s : streams) {
      if (s != nullptr) static_cast<void>(hipStreamDestroy(s));
    }
  }

  Context* ctx;
  uint32_t depth, max_rows;
  int32_t n = 0;
  uint8_t type = 0;
  bool near = false;
  std::vector<int64_t> x, y;
  DeviceArray<longlong2> d_xy;
  DeviceArray<KoptPlan> d_plans;
  KoptPlan host_plans[3];
  CertifyPlan cert[3];
  hipStream_t streams[2] = {nullptr, nullptr};
  std::vector<LeafSlot> slots;
  std::deque<uint32_t> fifo;
  int current = -1;
  LeafTour tour;
  HostPool pool;
  ee_leaf_stats stats{};
  // a fault for tests of the fail-closed path: one cell of the next block's
  // host copy replaced before certification
  bool inject = false;
  uint64_t inject_cell = 0;
  int64_t inject_value = 0;

  void Create(const int64_t* xs, const int64_t* ys, int32_t dim, uint8_t dist_type) {
    if (xs == nullptr || ys == nullptr || dim <= 0 || (dist_type != EE_EUC_2D && dist_type != EE_CEIL_2D)) {
      throw std::invalid_argument("leaf pipeline: coordinates are invalid");
    }
    if (!DistanceSafe(xs, ys, dim)) throw std::invalid_argument("leaf pipeline needs distance-safe coordinates");
    n = dim;
    type = dist_type;
    near = NearExtent(xs, ys, dim);
    x.assign(xs, xs + dim);
    y.assign(ys, ys + dim);
    ctx->Activate();
    const int dev = ctx->device;
    std::vector<longlong2> xy(static_cast<size_t>(n));
    for (int32_t i = 0; i < n; ++i) xy[static_cast<size_t>(i)] = longlong2{x[i], y[i]};
    d_xy.Allocate(xy.size(), dev);
    d_plans.Allocate(3, dev);
    for (uint32_t k = 3; k <= 5; ++k) {
      uint64_t hash = 0;
      const std::vector<ee_endpoint_matching> tpl = CanonicalTemplates(k, &hash);
      host_plans[k - 3] = BuildPlan(k, tpl.data(), static_cast<uint32_t>(tpl.size()));
      cert[k - 3] = MakeCertifyPlan(k, tpl.data(), static_cast<uint32_t>(tpl.size()));
    }
    for (hipStream_t& s : streams) CheckHip(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    slots.resize(depth + 1);
    const size_t cells = static_cast<size_t>(max_rows) * kMaxTemplatesPerRow;
    for (size_t i = 0; i < slots.size(); ++i) {
      LeafSlot& s = slots[i];
      s.stream = streams[i % 2];
      CheckHip(hipEventCreate(&s.ev_start), "hipEventCreate");
      CheckHip(hipEventCreate(&s.ev_kernel), "hipEventCreate");
      CheckHip(hipEventCreate(&s.ev_done), "hipEventCreate");
      s.d_tasks.Allocate(max_rows, dev);
      s.d_positions.Allocate(static_cast<size_t>(max_rows) * kMaxK, dev);
      s.d_costs.Allocate(cells, dev);
      s.d_status.Allocate(1, dev);
      s.h_tasks.Reserve(max_rows);
      s.h_positions.Reserve(static_cast<size_t>(max_rows) * kMaxK);
      s.h_costs.Reserve(cells);
      s.h_status.Reserve(1);
      s.improving.resize(static_cast<size_t>(max_rows) * ((kMaxTemplatesPerRow + 31) / 32));
    }
    d_xy.Upload(xy.data(), xy.size(), ctx->stream);
    d_plans.Upload(host_plans, 3, ctx->stream);
    ctx->Sync();
    stats.certify_threads = pool.size();
  }

  void Drain() {
    for (hipStream_t s : streams) CheckHip(hipStreamSynchronize(s), "hipStreamSynchronize");
  }

  void SetTour(const ee_tour_view* t) {
    CheckCursorTour(t, n);
    ctx->Activate();
    Drain();
    const size_t L = t->tour_length, S = t->selectable_count;
    tour.valid = false;
    tour.tour.assign(t->tour, t->tour + L);
    tour.edge_costs.assign(t->edge_costs, t->edge_costs + L);
    tour.selectable.assign(t->selectable, t->selectable + S);
    const int dev = ctx->device;
    if (tour.d_tour.count() < L) tour.d_tour.Allocate(L, dev);
    if (tour.d_edge_costs.count() < L) tour.d_edge_costs.Allocate(L, dev);
    if (tour.d_selectable.count() < std::max<size_t>(S, 1)) tour.d_selectable.Allocate(std::max<size_t>(S, 1), dev);
    tour.d_tour.Upload(tour.tour.data(), L, ctx->stream);
    tour.d_edge_costs.Upload(tour.edge_costs.data(), L, ctx->stream);
    tour.d_selectable.Upload(tour.selectable.data(), S, ctx->stream);
    ctx->Sync();
    tour.valid = true;
  }

  LeafSlot& FreeSlot() {
    if (fifo.size() >= depth) throw std::logic_error("leaf pipeline: `depth` blocks are already outstanding");
    for (size_t i = 0; i < slots.size(); ++i) {
      if (!slots[i].queued && static_cast<int>(i) != current) {
        fifo.push_back(static_cast<uint32_t>(i));
        return slots[i];
      }
    }
    throw std::logic_error("leaf pipeline: no free slot");
  }

  // The cost launch of a slot whose rows are (being) written on its stream,
  // then the downloads and the completion event.
  void EnqueueCosts(LeafSlot& s) {
    KoptInstance inst{};
    inst.n = static_cast<uint32_t>(n);
    inst.distance_type = type;
    inst.near = near ? 1 : 0;
    const KoptTuning& tuning = ctx->kopt_tuning;
    KoptSegment segment{0, s.k, 0, s.rows, 0, 0, tuning.rows_per_tile[s.k - 3], 0};
    KoptLaunch a{};
    KoptAssignTiles(&segment, 1, &inst, host_plans, tuning, nullptr, &a);
    if (a.inline_count == 0) throw std::logic_error("leaf pipeline launch needs tile descriptors");
    a.xy = d_xy.get();
    a.tiles = nullptr;
    a.tasks = s.d_tasks.get();
    a.plans = d_plans.get();
    a.costs = s.d_costs.get();
    bool present[3] = {false, false, false};
    present[s.k - 3] = true;
    const size_t shared = KoptPrepareLayout(&a, host_plans, present, tuning, near);
    LaunchKopt(a, tuning, shared, s.stream);
    CheckHip(hipEventRecord(s.ev_kernel, s.stream), "hipEventRecord");
    const size_t cells = static_cast<size_t>(s.rows) * host_plans[s.k - 3].templates;
    if (s.cursor) {
      s.d_tasks.Download(s.h_tasks.get(), s.rows, s.stream);
      s.d_positions.Download(s.h_positions.get(), static_cast<size_t>(s.rows) * s.k, s.stream);
      s.d_status.Download(s.h_status.get(), 1, s.stream);
    }
    s.d_costs.Download(s.h_costs.get(), cells, s.stream);
    CheckHip(hipEventRecord(s.ev_done, s.stream), "hipEventRecord");
  }

  void SubmitTasks(uint32_t k, const ee_kopt_task* tasks, uint32_t rows, uint64_t tag) {
    if (k < 3 || k > 5) throw std::invalid_argument("leaf pipeline: k must lie in [3, 5]");
    if (rows == 0 || rows > max_rows || tasks == nullptr) {
      throw std::invalid_argument("leaf pipeline: a block needs 1..max_rows task rows");
    }
    ValidateTasks(n, k, tasks, rows);
    ctx->Activate();
    LeafSlot& s = FreeSlot();
    try {
      s.cursor = false;
      s.k = k;
      s.rows = rows;
      s.tag = tag;
      s.first = 0;
      std::memcpy(s.h_tasks.get(), tasks, rows * sizeof(ee_kopt_task));
      CheckHip(hipEventRecord(s.ev_start, s.stream), "hipEventRecord");
      s.d_tasks.Upload(s.h_tasks.get(), rows, s.stream);
      EnqueueCosts(s);
      s.queued = true;
    } catch (...) {
      fifo.pop_back();
      throw;
    }
  }

  uint32_t SubmitCursor(uint32_t k, uint64_t first, uint32_t count, uint64_t tag) {
    if (k < 3 || k > 5) throw std::invalid_argument("leaf pipeline: k must lie in [3, 5]");
    if (!tour.valid) throw std::logic_error("leaf pipeline: no tour context (ee_leaf_set_tour)");
    if (count > max_rows) throw std::invalid_argument("leaf pipeline: a block holds at most max_rows rows");
    const uint64_t total = CursorSetCount(static_cast<uint32_t>(tour.selectable.size()), k);
    const uint32_t rows = first >= total ? 0 : static_cast<uint32_t>(std::min<uint64_t>(count, total - first));
    if (rows == 0) return 0;
    ctx->Activate();
    LeafSlot& s = FreeSlot();
    try {
      s.cursor = true;
      s.k = k;
      s.rows = rows;
      s.tag = tag;
      s.first = first;
      CheckHip(hipEventRecord(s.ev_start, s.stream), "hipEventRecord");
      CheckHip(hipMemsetAsync(s.d_status.get(), 0, sizeof(uint32_t), s.stream), "hipMemsetAsync status");
      LaunchCursorRows(tour.d_tour.get(), tour.d_edge_costs.get(), tour.d_selectable.get(),
                       static_cast<uint32_t>(tour.tour.size()), static_cast<uint32_t>(tour.selectable.size()), k,
                       first, rows, s.d_tasks.get(), s.d_positions.get(), s.d_status.get(), s.stream);
      EnqueueCosts(s);
      s.queued = true;
    } catch (...) {
      fifo.pop_back();
      throw;
    }
    return rows;
  }

  // The rows of a cursor block rebuilt on the host: the cursor's enumeration
  // order from rank `first` and BuildKOptCostTask's rule
  // (kopt_search.cpp:561-584), compared with the device's rows and positions.
  void CheckCursorRows(const LeafSlot& s) const {
    const uint32_t k = s.k;
    const uint64_t items = tour.selectable.size();
    std::vector<uint64_t> c = UnrankCombination(s.first, k, items);
    const size_t L = tour.tour.size();
    for (uint32_t r = 0; r < s.rows; ++r) {
      uint64_t pos[kMaxK];
      for (uint32_t i = 0; i < k; ++i) pos[i] = tour.selectable[c[i]];
      std::sort(pos, pos + k);
      ee_kopt_task want{};
      unsigned __int128 cost = 0;
      for (uint32_t e = 0; e < k; ++e) {
        want.port_nodes[2 * e] = tour.tour[pos[e]];
        want.port_nodes[2 * e + 1] = tour.tour[(pos[e] + 1) % L];
        cost += static_cast<uint64_t>(tour.edge_costs[pos[e]]);
      }
      if (cost > static_cast<unsigned __int128>(std::numeric_limits<int64_t>::max())) {
        throw std::runtime_error("k-opt cursor: deleted cost sum overflows int64_t");
      }
      want.deleted_cost = static_cast<int64_t>(cost);
      const ee_kopt_task& got = s.h_tasks.get()[r];
      const uint32_t* got_pos = s.h_positions.get() + static_cast<size_t>(r) * k;
      bool same = got.deleted_cost == want.deleted_cost;
      for (uint32_t p = 0; p < 2 * k; ++p) same = same && got.port_nodes[p] == want.port_nodes[p];
      for (uint32_t i = 0; i < k; ++i) same = same && got_pos[i] == pos[i];
      if (!same) {
        throw std::runtime_error("leaf pipeline: device-built task row " + std::to_string(r) + " of block " +
                                 std::to_string(s.tag) + " differs from the cursor's row");
      }
      if (r + 1 < s.rows && !Advance(&c, items)) {
        throw std::runtime_error("leaf pipeline: the cursor ran out of deletion sets inside a block");
      }
    }
  }

  void Next(ee_leaf_block* out) {
    if (out == nullptr) throw std::invalid_argument("leaf pipeline: block is NULL");
    if (fifo.empty()) throw std::logic_error("leaf pipeline: no block is outstanding");
    ctx->Activate();
    const uint32_t index = fifo.front();
    fifo.pop_front();
    LeafSlot& s = slots[index];
    s.queued = false;
    current = -1;
    auto t0 = std::chrono::steady_clock::now();
    CheckHip(hipEventSynchronize(s.ev_done), "hipEventSynchronize");
    stats.wait_ms += MsSince(t0);
    float kernel_ms = 0.0f, copy_ms = 0.0f;
    CheckHip(hipEventElapsedTime(&kernel_ms, s.ev_start, s.ev_kernel), "hipEventElapsedTime");
    CheckHip(hipEventElapsedTime(&copy_ms, s.ev_kernel, s.ev_done), "hipEventElapsedTime");
    stats.kernel_ms += kernel_ms;
    stats.copy_ms += copy_ms;
    if (s.cursor && (s.h_status.get()[0] & 1u) != 0) {
      throw std::runtime_error("k-opt cursor: deleted cost sum overflows int64_t");
    }
    t0 = std::chrono::steady_clock::now();
    const CertifyPlan& plan = cert[s.k - 3];
    if (inject) {
      inject = false;
      if (inject_cell < static_cast<uint64_t>(s.rows) * plan.templates) s.h_costs.get()[inject_cell] = inject_value;
    }
    const std::function<void()> rows_check = [&] { CheckCursorRows(s); };
    const CertifyResult r = CertifyCells(pool, plan, x.data(), y.data(), n, type, near, s.h_tasks.get(), s.rows,
                                         s.h_costs.get(), s.improving.data(), s.cursor ? &rows_check : nullptr);
    stats.certify_ms += MsSince(t0);
    if (r.mismatch != UINT64_MAX) {
      throw std::runtime_error("device k-opt cost failed CPU exact-matrix certification, block " +
                               std::to_string(s.tag) + ", cell=" + std::to_string(r.mismatch));
    }
    current = static_cast<int>(index);
    ++stats.blocks;
    stats.rows += s.rows;
    stats.cells += static_cast<uint64_t>(s.rows) * plan.templates;
    stats.improving_cells += r.improving;
    ee_leaf_block b{};
    b.tag = s.tag;
    b.first = s.first;
    b.k = s.k;
    b.rows = s.rows;
    b.template_count = plan.templates;
    b.mask_words = (plan.templates + 31) / 32;
    b.tasks = s.h_tasks.get();
    b.positions = s.cursor ? s.h_positions.get() : nullptr;
    b.costs = s.h_costs.get();
    b.improving = s.improving.data();
    b.improving_cells = r.improving;
    *out = b;
  }
};

Leaf& AsLeaf(ee_leaf_pipeline* p) {
  if (p == nullptr) throw std::invalid_argument("leaf pipeline is NULL");
  return *reinterpret_cast<Leaf*>(p);
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

extern "C" {

ee_status ee_leaf_create(ee_ctx* ctx, const int64_t* x, const int64_t* y, int32_t dimension,
                         uint8_t distance_type, uint32_t depth, uint32_t max_rows, ee_leaf_pipeline** out) {
  return ee::Guarded([&] {
    if (ctx == nullptr || out == nullptr) throw std::invalid_argument("NULL argument");
    *out = nullptr;
    if (depth < 1 || depth > ee::kMaxLeafDepth) throw std::invalid_argument("leaf pipeline depth must lie in [1, 4]");
    if (max_rows < 1 || max_rows > ee::kMaxLeafRows) {
      throw std::invalid_argument("leaf pipeline max_rows must lie in [1, 65536]");
    }
    auto leaf = std::make_unique<ee::Leaf>(reinterpret_cast<ee::Context*>(ctx), depth, max_rows);
    leaf->Create(x, y, dimension, distance_type);
    *out = reinterpret_cast<ee_leaf_pipeline*>(leaf.release());
  });
}

void ee_leaf_destroy(ee_leaf_pipeline* pipeline) { delete reinterpret_cast<ee::Leaf*>(pipeline); }

ee_status ee_leaf_set_tour(ee_leaf_pipeline* pipeline, const ee_tour_view* tour) {
  return ee::Guarded([&] { ee::AsLeaf(pipeline).SetTour(tour); });
}

ee_status ee_leaf_submit_tasks(ee_leaf_pipeline* pipeline, uint32_t k, const ee_kopt_task* tasks, uint32_t rows,
                               uint64_t tag) {
  return ee::Guarded([&] { ee::AsLeaf(pipeline).SubmitTasks(k, tasks, rows, tag); });
}

ee_status ee_leaf_submit_cursor(ee_leaf_pipeline* pipeline, uint32_t k, uint64_t first, uint32_t count,
                                uint64_t tag, uint32_t* rows) {
  return ee::Guarded([&] {
    if (rows) *rows = 0;
    const uint32_t made = ee::AsLeaf(pipeline).SubmitCursor(k, first, count, tag);
    if (rows) *rows = made;
  });
}

ee_status ee_leaf_next(ee_leaf_pipeline* pipeline, ee_leaf_block* block) {
  return ee::Guarded([&] { ee::AsLeaf(pipeline).Next(block); });
}

ee_status ee_leaf_inject_cell(ee_leaf_pipeline* pipeline, uint64_t cell, int64_t value) {
  return ee::Guarded([&] {
    ee::Leaf& l = ee::AsLeaf(pipeline);
    l.inject = true;
    l.inject_cell = cell;
    l.inject_value = value;
  });
}

ee_status ee_leaf_stats_get(const ee_leaf_pipeline* pipeline, ee_leaf_stats* stats) {
  return ee::Guarded([&] {
    if (stats == nullptr) throw std::invalid_argument("stats is NULL");
    *stats = ee::AsLeaf(const_cast<ee_leaf_pipeline*>(pipeline)).stats;
  });
}

}  // extern "C"
