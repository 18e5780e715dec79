// Device-resident k-opt batches: every (instance, k) segment of a batch is
// evaluated by one launch of the tiled cost kernel, with inputs and the whole
// cost output kept in HBM between launches.  Includes the autotuner that
// picks rows per tile, grid, block and store policy by timing candidates.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "ee_runtime.h"

namespace ee {
namespace {

constexpr uint32_t kTemplateCount[3] = {4, 25, 208};
// Bumped whenever the k-opt kernel changes, so cached tunings are retuned.
constexpr int kKernelRevision = 8;

struct Batch {
  explicit Batch(Context* c) : ctx(c) {}
  ~Batch() {
    DropGraph();
    for (PipeLane& l : lanes) {
      if (l.ev_join != nullptr) static_cast<void>(hipEventDestroy(l.ev_join));
      if (l.stream != nullptr) static_cast<void>(hipStreamDestroy(l.stream));
    }
  }

  Context* ctx;
  // instances
  std::vector<KoptInstance> instances;
  std::vector<longlong2> host_xy;
  DeviceArray<longlong2> d_xy;
  bool instances_dirty = false;
  // work
  std::vector<KoptSegment> segments;
  DeviceArray<ee_kopt_task> d_tasks;
  DeviceArray<int64_t> d_costs;
  // improving-template masks instead of cells (ee_kopt_batch_set_output)
  bool mask_output = false;
  DeviceArray<uint32_t> d_masks;
  std::vector<KoptTile> tiles;
  DeviceArray<KoptTile> d_tiles;
  KoptLaunch ranges{};  // per-k tile ranges of the current tiles
  uint64_t task_count = 0, cell_count = 0, tile_count = 0;
  bool k_present[3] = {false, false, false};
  // plans for the canonical templates of k = 3, 4, 5
  std::array<KoptPlan, 3> plans;
  DeviceArray<KoptPlan> d_plans;
  KoptTuning tuning;
  bool layout_dirty = true;
  // graph replay of one launch
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // the finalized launch arguments and their device copy, which the kernel
  // reads instead of the per-launch argument segment
  KoptLaunch host_args{};
  DeviceArray<KoptLaunch> d_args;
  bool args_valid = false;
  uint32_t grid = 0;
  size_t shared = 0;
  // pipelined runs: launch i goes to lane i % depth, lane 0 being the
  // context's stream and cost pool and every other lane a stream and a pool
  // of its own, so a batch's launch can start while the previous one drains
  struct PipeLane {
    hipStream_t stream = nullptr;
    hipEvent_t ev_join = nullptr;
    DeviceArray<int64_t> costs;
    bool written = false;  // a launch of the current work wrote this pool
  };
  uint32_t pipeline_depth = 2;
  std::vector<PipeLane> lanes;  // lanes 1 .. depth - 1

  void DropGraph() {
    if (exec != nullptr) static_cast<void>(hipGraphExecDestroy(exec));
    if (graph != nullptr) static_cast<void>(hipGraphDestroy(graph));
    exec = nullptr;
    graph = nullptr;
  }

  bool NearOnly() const {
    for (const KoptInstance& inst : instances) {
      if (inst.near == 0) return false;
    }
    return true;
  }

  size_t SharedBytes() const {
    KoptLaunch probe{};
    probe.masks = mask_output ? const_cast<uint32_t*>(d_masks.get()) : nullptr;  // mask mode stages deleted costs
    probe.mask_words = mask_output ? MaskWords() : 0u;                              // and the tile's mask words
    return KoptPrepareLayout(&probe, plans.data(), k_present, tuning, NearOnly());
  }

  // Splits segments into tiles for the current tuning and uploads the maps.
  void Prepare() {
    ctx->Activate();
    if (instances_dirty) {
      d_xy.Allocate(host_xy.size(), ctx->device);
      d_xy.Upload(host_xy.data(), host_xy.size(), ctx->stream);
      ctx->Sync();
      instances_dirty = false;
    }
    if (!layout_dirty) return;
    tiles.assign(KoptCountTiles(segments.data(), segments.size(), tuning), KoptTile{});
    KoptAssignTiles(segments.data(), segments.size(), instances.data(), plans.data(), tuning,
                    tiles.data(), &ranges);
    tile_count = ranges.tile_count;
    if (SharedBytes() > 65536) throw std::invalid_argument("k-opt tuning needs more than 64 KiB of LDS");
    if (ranges.inline_count == 0) {  // otherwise the kernel derives tiles from its arguments
      d_tiles.Allocate(std::max<size_t>(tiles.size(), 1), ctx->device);
      d_tiles.Upload(tiles.data(), tiles.size(), ctx->stream);
      ctx->Sync();
    }
    layout_dirty = false;
    DropGraph();
  }

  // words per task row of the mask output: the widest present k
  uint32_t MaskWords() const {
    uint32_t w = 1;
    for (int i = 0; i < 3; ++i) {
      if (k_present[i]) w = std::max(w, (plans[i].templates + 31u) / 32u);
    }
    return w;
  }

  KoptLaunch Args() {
    KoptLaunch a = ranges;
    a.xy = d_xy.get();
    a.tiles = d_tiles.get();
    a.tasks = d_tasks.get();
    a.plans = d_plans.get();
    a.costs = d_costs.get();
    if (mask_output) {
      a.masks = d_masks.get();
      a.mask_words = MaskWords();
    }
    return a;
  }

  // Brings the device copy of the launch arguments up to date; a change
  // (new work, tuning or output mode) drops the captured graph, whose launch
  // shape belongs to the old arguments.
  void SyncArgs() {
    KoptLaunch a;
    std::memset(&a, 0, sizeof(a));  // padding included, for the comparison below
    a = Args();
    shared = KoptPrepareLayout(&a, plans.data(), k_present, tuning, NearOnly());
    grid = KoptFinalize(&a, tuning, shared);
    if (args_valid && std::memcmp(&a, &host_args, sizeof(a)) == 0) return;
    DropGraph();
    if (d_args.count() == 0) d_args.Allocate(1, ctx->device);
    host_args = a;
    d_args.Upload(&host_args, 1, ctx->stream);
    ctx->Sync();
    args_valid = true;
  }

  void LaunchOnce() {
    SyncArgs();
    LaunchKoptResident(host_args, d_args.get(), grid, tuning, shared, ctx->stream);
  }

  // Mean device time of one launch over `launches` back-to-back launches,
  // pipelined over two streams when asked.
  // Clock stamps per launch: two per wavefront of the grid.
  size_t StampsPerLaunch() const { return 2 * static_cast<size_t>(grid) * (tuning.block / 64); }

  // stamps (measurement only, may be null): StampsPerLaunch() entries per
  // launch, launch i at stamps + i x StampsPerLaunch().
  double Time(uint32_t launches, bool use_graph, bool pipelined = false,
              unsigned long long* stamps = nullptr) {
    if (use_graph && pipelined) throw std::invalid_argument("graph replay and pipelined runs do not combine");
    Prepare();
    if (tile_count == 0 || launches == 0) return 0.0;
    SyncArgs();  // any upload happens here, never inside the capture
    if (pipelined) return TimePipelined(launches, stamps);
    if (use_graph && exec == nullptr) {
      CheckHip(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "begin capture");
      LaunchOnce();
      CheckHip(hipStreamEndCapture(ctx->stream, &graph), "end capture");
      CheckHip(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "graph instantiate");
    }
    CheckHip(hipEventRecord(ctx->ev_begin, ctx->stream), "event record");
    for (uint32_t i = 0; i < launches; ++i) {
      if (use_graph) {
        CheckHip(hipGraphLaunch(exec, ctx->stream), "graph launch");
      } else if (stamps != nullptr) {
        SyncArgs();
        LaunchKoptResident(host_args, d_args.get(), grid, tuning, shared, ctx->stream,
                           stamps + static_cast<size_t>(i) * StampsPerLaunch());
      } else {
        LaunchOnce();
      }
    }
    CheckHip(hipEventRecord(ctx->ev_end, ctx->stream), "event record");
    CheckHip(hipEventSynchronize(ctx->ev_end), "event synchronize");
    float ms = 0.0f;
    CheckHip(hipEventElapsedTime(&ms, ctx->ev_begin, ctx->ev_end), "event elapsed");
    return static_cast<double>(ms) / launches;
  }

  double TimePipelined(uint32_t launches, unsigned long long* stamps = nullptr) {
    if (mask_output) throw std::invalid_argument("pipelined runs write cost cells only");
    if (lanes.size() + 1 < pipeline_depth) lanes.resize(pipeline_depth - 1);
    std::vector<KoptLaunch> lane_args(pipeline_depth, host_args);
    for (uint32_t j = 1; j < pipeline_depth; ++j) {
      PipeLane& l = lanes[j - 1];
      if (l.stream == nullptr) {
        CheckHip(hipStreamCreateWithFlags(&l.stream, hipStreamNonBlocking), "pipeline stream");
        CheckHip(hipEventCreateWithFlags(&l.ev_join, hipEventDisableTiming), "join event");
      }
      if (l.costs.count() != d_costs.count()) l.costs.Allocate(d_costs.count(), ctx->device);
      // the kernel writes through the pointer arguments, so the device copy
      // of the launch description serves every pool
      lane_args[j].costs = l.costs.get();
    }
    CheckHip(hipEventRecord(ctx->ev_begin, ctx->stream), "event record");
    for (uint32_t j = 1; j < pipeline_depth; ++j) {
      CheckHip(hipStreamWaitEvent(lanes[j - 1].stream, ctx->ev_begin, 0), "stream wait");
    }
    for (uint32_t i = 0; i < launches; ++i) {
      const uint32_t j = i % pipeline_depth;
      const hipStream_t stream = j == 0 ? ctx->stream : lanes[j - 1].stream;
      LaunchKoptResident(lane_args[j], d_args.get(), grid, tuning, shared, stream,
                         stamps != nullptr ? stamps + static_cast<size_t>(i) * StampsPerLaunch() : nullptr);
      if (j > 0) lanes[j - 1].written = true;
    }
    for (uint32_t j = 1; j < pipeline_depth; ++j) {
      CheckHip(hipEventRecord(lanes[j - 1].ev_join, lanes[j - 1].stream), "event record");
      CheckHip(hipStreamWaitEvent(ctx->stream, lanes[j - 1].ev_join, 0), "stream wait");
    }
    CheckHip(hipEventRecord(ctx->ev_end, ctx->stream), "event record");
    CheckHip(hipEventSynchronize(ctx->ev_end), "event synchronize");
    float ms = 0.0f;
    CheckHip(hipEventElapsedTime(&ms, ctx->ev_begin, ctx->ev_end), "event elapsed");
    return static_cast<double>(ms) / launches;
  }

  // Forgets which extra pools hold results (new work or output mode).
  void ClearWritten() {
    for (PipeLane& l : lanes) l.written = false;
  }

  uint64_t ResidentBytes() const {
    uint64_t pools = 0;
    for (const PipeLane& l : lanes) pools += l.costs.bytes();
    return d_xy.bytes() + d_tasks.bytes() + d_costs.bytes() + pools + d_masks.bytes() + d_tiles.bytes() +
           d_plans.bytes();
  }
};

Batch& AsBatch(ee_kopt_batch* b) {
  if (b == nullptr) throw std::invalid_argument("batch is NULL");
  return *reinterpret_cast<Batch*>(b);
}

void ApplyTuning(Batch& b, const ee_kopt_tuning& t) {
  for (int i = 0; i < 3; ++i) {
    if (t.rows_per_tile[i] == 0 || t.rows_per_tile[i] > 64) {
      throw std::invalid_argument("rows per tile must be in [1, 64]");
    }
  }
  if (t.block == 0 || t.block > 256 || t.block % 64 != 0) {
    throw std::invalid_argument("block must be 64, 128, 192 or 256");
  }
  KoptTuning next;
  for (int i = 0; i < 3; ++i) next.rows_per_tile[i] = t.rows_per_tile[i];
  next.grid = t.grid;
  next.block = t.block;
  next.non_temporal = t.non_temporal != 0;
  const KoptTuning prev = b.tuning;
  b.tuning = next;
  size_t shared = 0;
  try {
    shared = b.SharedBytes();
  } catch (...) {
    b.tuning = prev;
    throw;
  }
  if (shared > 65536) {
    b.tuning = prev;
    throw std::invalid_argument("k-opt tuning needs more than 64 KiB of LDS");
  }
  b.layout_dirty = true;
  b.DropGraph();
}

ee_kopt_tuning Export(const KoptTuning& t) {
  ee_kopt_tuning o{};
  for (int i = 0; i < 3; ++i) o.rows_per_tile[i] = t.rows_per_tile[i];
  o.grid = t.grid;
  o.block = t.block;
  o.non_temporal = t.non_temporal ? 1 : 0;
  return o;
}

std::string TuningKey(const Batch& b, bool pipelined) {
  hipDeviceProp_t prop{};
  static_cast<void>(hipGetDeviceProperties(&prop, b.ctx->device));
  uint64_t rows[3] = {0, 0, 0};
  for (const KoptSegment& s : b.segments) rows[s.k - 3] += s.task_count;
  std::ostringstream key;
  key << "kopt" << kKernelRevision << '/' << prop.gcnArchName << '/' << prop.multiProcessorCount
      << "cu/rows";
  for (uint64_t r : rows) {
    int bucket = 0;  // power-of-two bucket of the row count
    while ((uint64_t{1} << bucket) < r) ++bucket;
    key << '-' << (r == 0 ? -1 : bucket);
  }
  if (pipelined) key << "/pipelined" << b.pipeline_depth;
  if (b.mask_output) key << "/masks";  // a different store pattern, tuned on its own
  std::string s = key.str();
  std::replace(s.begin(), s.end(), ' ', '_');
  return s;
}

bool LoadCached(const char* path, const std::string& key, ee_kopt_tuning* out) {
  std::ifstream in(path);
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream f(line);
    std::string k;
    ee_kopt_tuning t{};
    if (f >> k >> t.rows_per_tile[0] >> t.rows_per_tile[1] >> t.rows_per_tile[2] >> t.grid >>
            t.block >> t.non_temporal &&
        k == key) {
      *out = t;
      return true;
    }
  }
  return false;
}

void StoreCached(const char* path, const std::string& key, const ee_kopt_tuning& t) {
  std::vector<std::string> keep;
  {
    std::ifstream in(path);
    std::string line;
    while (std::getline(in, line)) {
      if (line.rfind(key + " ", 0) != 0 && !line.empty()) keep.push_back(line);
    }
  }
  std::ostringstream entry;
  entry << key << ' ' << t.rows_per_tile[0] << ' ' << t.rows_per_tile[1] << ' '
        << t.rows_per_tile[2] << ' ' << t.grid << ' ' << t.block << ' ' << t.non_temporal;
  keep.push_back(entry.str());
  const std::string tmp = std::string(path) + ".tmp";
  {
    std::ofstream out(tmp);
    for (const std::string& l : keep) out << l << '\n';
  }
  std::rename(tmp.c_str(), path);
}

// Coordinate descent over the tuning space, then a joint search over the
// rows of the largest k present and the grid (small tiles pay only with a
// wavefront per tile, so the two move together); every candidate is timed on
// the batch itself, best of three runs of `launches` launches.
ee_kopt_tuning Autotune(Batch& b, uint32_t launches, bool pipelined) {
  ee_kopt_tuning best = Export(b.tuning);
  const auto measure = [&](const ee_kopt_tuning& t) -> double {
    try {
      ApplyTuning(b, t);
    } catch (const std::invalid_argument&) {
      return 1e30;  // does not fit LDS
    }
    b.Time(1, false);  // warm
    double ms = b.Time(launches, false, pipelined);
    for (int rep = 0; rep < 2; ++rep) ms = std::min(ms, b.Time(launches, false, pipelined));
    return ms;
  };
  // an idle device reaches its clocks only after some work: warm up first, or
  // the first candidates are timed slow and the descent starts off course
  try {
    ApplyTuning(b, best);
    b.Time(std::max<uint32_t>(20 * launches, 400), false, pipelined);
  } catch (const std::invalid_argument&) {
  }
  double best_ms = measure(best);
  // start points: one tile per wavefront with small k = 5 tiles, and its
  // neighbours, the shapes that won on the 65,536-row leaf workload and that
  // coordinate descent from the default rarely reaches
  const ee_kopt_tuning seeds[] = {{{40, 32, 5}, 0, 256, 1}, {{36, 20, 12}, 0, 256, 1}, {{32, 24, 6}, 0, 128, 1},
                                  {{40, 26, 6}, 0, 128, 1}, {{32, 16, 5}, 0, 256, 1}, {{40, 32, 16}, 0, 256, 1},
                                  {{32, 8, 5}, 0, 256, 1}};
  for (const ee_kopt_tuning& seed : seeds) {
    const double ms = measure(seed);
    if (ms < best_ms) {
      best_ms = ms;
      best = seed;
    }
  }
  const uint32_t cu = static_cast<uint32_t>(std::max(b.ctx->compute_units, 1));
  // residency cliffs (every block reserves the LDS of the launch's largest
  // k) make the time a bumpy function of the rows, hence the fine options
  const std::vector<uint32_t> rows_options[3] = {
      {16, 24, 28, 32, 36, 40, 48, 64}, {8, 16, 20, 22, 24, 26, 28, 32, 48, 64}, {4, 5, 6, 7, 8, 12, 16, 24, 32}};
  for (int pass = 0; pass < 2; ++pass) {
    for (int k = 2; k >= 0; --k) {
      if (!b.k_present[k]) continue;
      for (uint32_t r : rows_options[k]) {
        ee_kopt_tuning t = best;
        t.rows_per_tile[k] = r;
        const double ms = measure(t);
        if (ms < best_ms) {
          best_ms = ms;
          best = t;
        }
      }
    }
    for (uint32_t block : {256u, 128u, 64u}) {
      ee_kopt_tuning t = best;
      t.block = block;
      const double ms = measure(t);
      if (ms < best_ms) {
        best_ms = ms;
        best = t;
      }
    }
    for (uint32_t grid : {0u, 2 * cu, 4 * cu, 8 * cu, 16 * cu}) {
      ee_kopt_tuning t = best;
      t.grid = grid;
      const double ms = measure(t);
      if (ms < best_ms) {
        best_ms = ms;
        best = t;
      }
    }
    for (uint32_t nt : {0u, 1u}) {
      ee_kopt_tuning t = best;
      t.non_temporal = nt;
      const double ms = measure(t);
      if (ms < best_ms) {
        best_ms = ms;
        best = t;
      }
    }
  }
  // local search: one row option up or down for one k at a time, until no
  // neighbour is faster
  for (int round = 0; round < 8; ++round) {
    bool improved = false;
    for (int k = 0; k < 3; ++k) {
      if (!b.k_present[k]) continue;
      const std::vector<uint32_t>& opts = rows_options[k];
      const auto at = std::find(opts.begin(), opts.end(), best.rows_per_tile[k]);
      if (at == opts.end()) continue;
      for (int step : {-1, 1}) {
        const auto pos = (at - opts.begin()) + step;
        if (pos < 0 || pos >= static_cast<long>(opts.size())) continue;
        ee_kopt_tuning t = best;
        t.rows_per_tile[k] = opts[static_cast<size_t>(pos)];
        const double ms = measure(t);
        if (ms < best_ms) {
          best_ms = ms;
          best = t;
          improved = true;
        }
      }
    }
    if (!improved) break;
  }
  int top = -1;
  for (int k = 2; k >= 0 && top < 0; --k) {
    if (b.k_present[k]) top = k;
  }
  if (top >= 0) {
    const ee_kopt_tuning base = best;
    for (uint32_t r : rows_options[top]) {
      for (uint32_t grid : {0u, base.grid}) {
        ee_kopt_tuning t = base;
        t.rows_per_tile[top] = r;
        t.grid = grid;
        const double ms = measure(t);
        if (ms < best_ms) {
          best_ms = ms;
          best = t;
        }
      }
    }
  }
  ApplyTuning(b, best);
  return best;
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

ee_status ee_ctx_synchronize(ee_ctx* handle) {
  return Guarded([&] {
    if (handle == nullptr) throw std::invalid_argument("context is NULL");
    auto& ctx = *reinterpret_cast<ee::Context*>(handle);
    ctx.Activate();
    ctx.Sync();
  });
}

ee_status ee_kopt_batch_create(ee_ctx* handle, ee_kopt_batch** out) {
  return Guarded([&] {
    if (handle == nullptr || out == nullptr) throw std::invalid_argument("NULL argument");
    *out = nullptr;
    auto* ctx = reinterpret_cast<ee::Context*>(handle);
    ctx->Activate();
    auto batch = std::make_unique<ee::Batch>(ctx);
    for (uint32_t k = 3; k <= 5; ++k) {
      uint64_t hash = 0;
      const auto table = ee::CanonicalTemplates(k, &hash);
      batch->plans[k - 3] = ee::BuildPlan(k, table.data(), static_cast<uint32_t>(table.size()));
    }
    batch->d_plans.Allocate(3, ctx->device);
    batch->d_plans.Upload(batch->plans.data(), 3, ctx->stream);
    ctx->Sync();
    *out = reinterpret_cast<ee_kopt_batch*>(batch.release());
  });
}

void ee_kopt_batch_destroy(ee_kopt_batch* batch) {
  if (batch == nullptr) return;
  auto* b = reinterpret_cast<ee::Batch*>(batch);
  static_cast<void>(hipSetDevice(b->ctx->device));
  static_cast<void>(hipStreamSynchronize(b->ctx->stream));
  delete b;
}

ee_status ee_kopt_batch_add_instance(ee_kopt_batch* batch, const int64_t* x, const int64_t* y,
                                     int32_t dimension, uint8_t distance_type,
                                     uint32_t* instance) {
  return Guarded([&] {
    ee::Batch& b = ee::AsBatch(batch);
    if (x == nullptr || y == nullptr || dimension <= 0) throw std::invalid_argument("bad coordinates");
    if (distance_type != EE_EUC_2D && distance_type != EE_CEIL_2D) {
      throw std::invalid_argument("unsupported distance type");
    }
    if (!ee::DistanceSafe(x, y, dimension)) {
      throw std::invalid_argument("k-opt cost needs distance-safe integer coordinates");
    }
    const ee::KoptInstance inst = ee::MakeKoptInstance(b.host_xy.size(), x, y, dimension, distance_type);
    for (int32_t i = 0; i < dimension; ++i) b.host_xy.push_back(longlong2{x[i], y[i]});
    b.instances.push_back(inst);
    b.instances_dirty = true;
    if (instance) *instance = static_cast<uint32_t>(b.instances.size() - 1);
  });
}

ee_status ee_kopt_batch_set_work(ee_kopt_batch* batch, const ee_kopt_segment_desc* segments,
                                 uint32_t segment_count, const ee_kopt_task* tasks,
                                 uint64_t* cell_offsets) {
  return Guarded([&] {
    ee::Batch& b = ee::AsBatch(batch);
    if (segment_count > 0 && segments == nullptr) throw std::invalid_argument("segments is NULL");
    std::vector<ee::KoptSegment> segs;
    uint64_t rows = 0, cells = 0;
    bool present[3] = {false, false, false};
    for (uint32_t s = 0; s < segment_count; ++s) {
      const ee_kopt_segment_desc& d = segments[s];
      if (d.instance >= b.instances.size()) throw std::invalid_argument("segment instance id");
      if (d.k < 3 || d.k > 5) throw std::invalid_argument("segment k must be in [3,5]");
      if (d.task_count > 0 && tasks == nullptr) throw std::invalid_argument("tasks is NULL");
      ee::ValidateTasks(static_cast<int32_t>(b.instances[d.instance].n), d.k, tasks + rows,
                        d.task_count);
      // every segment starts on a 128-byte boundary, so whole-line stores of
      // its rows stay line-aligned
      cells = (cells + ee::kSegmentAlignCells - 1) / ee::kSegmentAlignCells * ee::kSegmentAlignCells;
      if (cell_offsets) cell_offsets[s] = cells;
      ee::KoptSegment seg{d.instance, d.k, rows, d.task_count, cells, 0, 1, 0};
      segs.push_back(seg);
      present[d.k - 3] = present[d.k - 3] || d.task_count > 0;
      rows += d.task_count;
      cells += d.task_count * ee::kTemplateCount[d.k - 3];
    }
    if (cell_offsets) cell_offsets[segment_count] = cells;
    b.ctx->Activate();
    b.DropGraph();
    b.ClearWritten();
    b.d_tasks.Allocate(std::max<uint64_t>(rows, 1), b.ctx->device);
    b.d_tasks.Upload(tasks, rows, b.ctx->stream);
    b.d_costs.Allocate(std::max<uint64_t>(cells, 1), b.ctx->device);
    b.segments = std::move(segs);
    b.task_count = rows;
    b.cell_count = cells;
    for (int i = 0; i < 3; ++i) b.k_present[i] = present[i];
    if (b.mask_output) b.d_masks.Allocate(std::max<uint64_t>(rows * b.MaskWords(), 1), b.ctx->device);
    b.ctx->Sync();
    b.layout_dirty = true;
  });
}

ee_status ee_kopt_batch_set_tuning(ee_kopt_batch* batch, const ee_kopt_tuning* tuning) {
  return Guarded([&] {
    if (tuning == nullptr) throw std::invalid_argument("tuning is NULL");
    ee::ApplyTuning(ee::AsBatch(batch), *tuning);
  });
}

ee_status ee_kopt_batch_get_tuning(const ee_kopt_batch* batch, ee_kopt_tuning* tuning) {
  return Guarded([&] {
    if (batch == nullptr || tuning == nullptr) throw std::invalid_argument("NULL argument");
    *tuning = ee::Export(reinterpret_cast<const ee::Batch*>(batch)->tuning);
  });
}

ee_status ee_kopt_batch_autotune_runs(ee_kopt_batch* batch, uint32_t launches_per_candidate, int run_flags,
                                      const char* cache_path, ee_kopt_tuning* chosen) {
  return Guarded([&] {
    if ((run_flags & ~EE_KOPT_RUN_PIPELINED) != 0) throw std::invalid_argument("autotune times plain or pipelined runs");
    ee::Batch& b = ee::AsBatch(batch);
    const bool pipelined = (run_flags & EE_KOPT_RUN_PIPELINED) != 0;
    const std::string key = ee::TuningKey(b, pipelined);
    ee_kopt_tuning t{};
    if (cache_path != nullptr && ee::LoadCached(cache_path, key, &t)) {
      ee::ApplyTuning(b, t);
    } else {
      t = ee::Autotune(b, std::max<uint32_t>(launches_per_candidate, pipelined ? 2 * b.pipeline_depth : 1), pipelined);
      if (cache_path != nullptr) ee::StoreCached(cache_path, key, t);
    }
    if (chosen) *chosen = t;
  });
}

ee_status ee_kopt_batch_autotune(ee_kopt_batch* batch, uint32_t launches_per_candidate,
                                 const char* cache_path, ee_kopt_tuning* chosen) {
  return ee_kopt_batch_autotune_runs(batch, launches_per_candidate, 0, cache_path, chosen);
}

ee_status ee_kopt_batch_run(ee_kopt_batch* batch, uint32_t launches, int flags, double* device_ms) {
  return Guarded([&] {
    if ((flags & ~(EE_KOPT_RUN_GRAPH | EE_KOPT_RUN_PIPELINED)) != 0) {
      throw std::invalid_argument("unknown k-opt run flags");
    }
    const double ms = ee::AsBatch(batch).Time(launches, (flags & EE_KOPT_RUN_GRAPH) != 0,
                                               (flags & EE_KOPT_RUN_PIPELINED) != 0);
    if (device_ms) *device_ms = ms;
  });
}

ee_status ee_kopt_batch_run_stamped(ee_kopt_batch* batch, uint32_t launches, int flags, double* start_us,
                                    double* end_us) {
  return ee::Guarded([&] {
    if (start_us == nullptr || end_us == nullptr) throw std::invalid_argument("NULL argument");
    if (launches == 0 || launches > 1024) throw std::invalid_argument("launches must lie in [1, 1024]");
    if (flags != 0 && flags != EE_KOPT_RUN_PIPELINED) throw std::invalid_argument("plain or pipelined runs only");
    ee::Batch& b = ee::AsBatch(batch);
    b.ctx->Activate();
    b.Prepare();
    b.SyncArgs();  // sets the grid the stamps are sized for
    const size_t per = b.StampsPerLaunch();
    if (per == 0) throw std::logic_error("the batch has no work");
    ee::DeviceArray<unsigned long long> d_stamps(per * launches, b.ctx->device);
    b.Time(launches, false, flags == EE_KOPT_RUN_PIPELINED, d_stamps.get());
    std::vector<unsigned long long> h(per * launches);
    d_stamps.Download(h.data(), h.size(), b.ctx->stream);
    b.ctx->Sync();
    int khz = 0;
    ee::CheckHip(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b.ctx->device),
                 "hipDeviceGetAttribute(wall clock rate)");
    if (khz <= 0) throw ee::Unavailable("device reports no wall clock rate");
    unsigned long long origin = ~0ull;
    for (size_t w = 0; w < h.size(); w += 2) origin = std::min(origin, h[w]);
    for (uint32_t i = 0; i < launches; ++i) {
      unsigned long long first = ~0ull, last = 0;
      for (size_t w = 0; w < per; w += 2) {
        first = std::min(first, h[i * per + w]);
        last = std::max(last, h[i * per + w + 1]);
      }
      start_us[i] = static_cast<double>(first - origin) * 1e3 / khz;
      end_us[i] = static_cast<double>(last - origin) * 1e3 / khz;
    }
  });
}

ee_status ee_kopt_batch_set_pipeline_depth(ee_kopt_batch* batch, uint32_t depth) {
  return Guarded([&] {
    if (depth < 2 || depth > 4) throw std::invalid_argument("pipeline depth must be in [2, 4]");
    ee::AsBatch(batch).pipeline_depth = depth;
  });
}

ee_status ee_kopt_batch_download(ee_kopt_batch* batch, uint64_t first_cell, uint64_t cell_count,
                                 int64_t* costs) {
  return Guarded([&] {
    ee::Batch& b = ee::AsBatch(batch);
    if (first_cell > b.cell_count || cell_count > b.cell_count - first_cell) {
      throw std::invalid_argument("cell range outside the batch output");
    }
    if (cell_count > 0 && costs == nullptr) throw std::invalid_argument("costs is NULL");
    b.ctx->Activate();
    ee::DownloadRange(*b.ctx, b.d_costs, first_cell, cell_count, costs);
  });
}

ee_status ee_kopt_batch_download_pool(ee_kopt_batch* batch, uint32_t pool, uint64_t first_cell,
                                      uint64_t cell_count, int64_t* costs) {
  return Guarded([&] {
    ee::Batch& b = ee::AsBatch(batch);
    if (pool > 0 && (pool > b.lanes.size() || !b.lanes[pool - 1].written)) {
      throw std::logic_error("no launch of the current work wrote this cost pool");
    }
    if (first_cell > b.cell_count || cell_count > b.cell_count - first_cell) {
      throw std::invalid_argument("cell range outside the batch output");
    }
    if (cell_count > 0 && costs == nullptr) throw std::invalid_argument("costs is NULL");
    b.ctx->Activate();
    if (pool == 0) {
      ee::DownloadRange(*b.ctx, b.d_costs, first_cell, cell_count, costs);
    } else {
      ee::Batch::PipeLane& l = b.lanes[pool - 1];
      // the lane's launches ran on its own stream; wait for them first
      ee::CheckHip(hipStreamSynchronize(l.stream), "hipStreamSynchronize");
      ee::DownloadRange(*b.ctx, l.costs, first_cell, cell_count, costs);
    }
  });
}

ee_status ee_kopt_batch_set_output(ee_kopt_batch* batch, uint32_t mode) {
  return Guarded([&] {
    ee::Batch& b = ee::AsBatch(batch);
    if (mode > 1) throw std::invalid_argument("output mode must be 0 (costs) or 1 (improving masks)");
    b.ctx->Activate();
    b.mask_output = mode == 1;
    b.ClearWritten();
    if (b.mask_output) {
      b.d_masks.Allocate(std::max<uint64_t>(b.task_count * b.MaskWords(), 1), b.ctx->device);
    }
    b.DropGraph();  // the launch arguments change
  });
}

ee_status ee_kopt_batch_mask_words(const ee_kopt_batch* batch, uint32_t* words) {
  return Guarded([&] {
    if (batch == nullptr || words == nullptr) throw std::invalid_argument("NULL argument");
    *words = reinterpret_cast<const ee::Batch*>(batch)->MaskWords();
  });
}

ee_status ee_kopt_batch_download_masks(ee_kopt_batch* batch, uint64_t first_row, uint64_t row_count,
                                       uint32_t* masks) {
  return Guarded([&] {
    ee::Batch& b = ee::AsBatch(batch);
    if (!b.mask_output) throw std::logic_error("the batch writes cost cells, not masks");
    if (first_row > b.task_count || row_count > b.task_count - first_row) {
      throw std::invalid_argument("row range outside the batch");
    }
    if (row_count > 0 && masks == nullptr) throw std::invalid_argument("masks is NULL");
    const uint64_t w = b.MaskWords();
    b.ctx->Activate();
    ee::DownloadRange(*b.ctx, b.d_masks, first_row * w, row_count * w, masks);
  });
}

uint64_t ee_kopt_batch_resident_bytes(const ee_kopt_batch* batch) {
  return batch == nullptr ? 0 : reinterpret_cast<const ee::Batch*>(batch)->ResidentBytes();
}

}  // extern "C"
