// Host runtime of the gfx950 backend: device selection, per-context
// device-resident caches and the C entry points of the reference-shaped
// (synchronous, host buffers in and out) paths.  Host-only helpers (errors,
// validation, the template generator) are in ee_host.cpp.
//
// Cache semantics follow the reference so that callers observe the same hit
// flags and resident byte counts:
//   JV    (src/cuda/jv_cuda.cu:360-478): the static part (coordinates, edge
//         endpoints, edge weights) is keyed by an exact element-wise compare;
//         the dynamic part (active flags, CSR, witnesses) lives in buffers
//         that grow to max(required, 2 x current).  Resident bytes are
//         24m + 20n + 4 + 8A on an exact first fit.
//   k-opt (src/cuda/kopt_cost.cu:205-295): coordinates keyed by exact
//         compare of (dimension, distance type, x, y); templates keyed per k
//         by (generator hash, content); task and cost buffers grow the same
//         way.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "ee_exact.h"
#include "ee_internal.h"
#include "ee_jvhost.h"
#include "ee_runtime.h"

namespace ee {

void CheckHip(hipError_t status, const char* what) {
  if (status != hipSuccess) {
    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(status));
  }
}

KoptInstance MakeKoptInstance(uint64_t coord_base, const int64_t* x, const int64_t* y, int32_t n,
                              uint8_t distance_type) {
  KoptInstance inst{};
  inst.coord_base = coord_base;
  inst.n = static_cast<uint32_t>(n);
  inst.distance_type = distance_type;
  inst.near = NearExtent(x, y, n) ? 1 : 0;
  return inst;
}

std::vector<longlong2> Interleave(const int64_t* x, const int64_t* y, int32_t n) {
  std::vector<longlong2> xy(static_cast<size_t>(n));
  for (int32_t i = 0; i < n; ++i) {
    xy[static_cast<size_t>(i)].x = x[i];
    xy[static_cast<size_t>(i)].y = y[i];
  }
  return xy;
}

// ------------------------------------------------------------------ device
int CountDevices() {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess) return 0;
  return count;
}

int SelectDevice(int requested) {
  const int count = CountDevices();
  if (count <= 0) throw Unavailable("no visible HIP device");
  if (requested >= 0) {
    if (requested >= count) throw Unavailable("requested HIP device ordinal is not visible");
    CheckHip(hipSetDevice(requested), "hipSetDevice");
    return requested;
  }
  int best = -1;
  size_t best_free = 0;
  for (int d = 0; d < count; ++d) {
    if (hipSetDevice(d) != hipSuccess) continue;
    size_t free_bytes = 0, total_bytes = 0;
    if (hipMemGetInfo(&free_bytes, &total_bytes) == hipSuccess &&
        (best < 0 || free_bytes > best_free)) {
      best = d;
      best_free = free_bytes;
    }
  }
  if (best < 0) throw Unavailable("no HIP device could report its free memory");
  CheckHip(hipSetDevice(best), "hipSetDevice");
  return best;
}

Context::Context(int requested) {
  device = SelectDevice(requested);
  InitNoAllocScratch(device);
  CheckHip(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
  CheckHip(hipEventCreate(&ev_begin), "hipEventCreate");
  CheckHip(hipEventCreate(&ev_end), "hipEventCreate");
  hipDeviceProp_t prop{};
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    compute_units = prop.multiProcessorCount;
  }
}

Context::~Context() {
  static_cast<void>(hipSetDevice(device));
  static_cast<void>(hipStreamSynchronize(stream));
  jv = JvCache{};
  kopt = KoptCache{};
  static_cast<void>(hipEventDestroy(ev_begin));
  static_cast<void>(hipEventDestroy(ev_end));
  for (hipEvent_t e : stage_ready) {
    if (e != nullptr) static_cast<void>(hipEventDestroy(e));
  }
  static_cast<void>(hipStreamDestroy(stream));
}

void DownloadToCaller(Context& ctx, void* dst, const void* src, size_t bytes) {
  // below kDirect the runtime's own copy costs no more; chunks of kChunk
  // copied out by kParts host threads kept pace with PCIe on the MI355X host
  constexpr size_t kDirect = size_t{1} << 20, kChunk = size_t{4} << 20, kParts = 4;
  if (bytes == 0) return;
  if (dst == nullptr || src == nullptr) throw std::logic_error("download with a NULL buffer");
  if (bytes < kDirect) {
    CheckHip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx.stream), "hipMemcpyAsync D2H");
    ctx.Sync();
    return;
  }
  for (int b = 0; b < 2; ++b) {
    ctx.stage[b].Reserve(kChunk);
    if (ctx.stage_ready[b] == nullptr) {
      CheckHip(hipEventCreateWithFlags(&ctx.stage_ready[b], hipEventDisableTiming), "hipEventCreate");
    }
  }
  char* const out = static_cast<char*>(dst);
  const char* const in = static_cast<const char*>(src);
  const size_t chunks = (bytes + kChunk - 1) / kChunk;
  // chunk i goes to stage i % 2; the chunk before it is copied out meanwhile,
  // so a stage is refilled only after its previous chunk has left it
  for (size_t i = 0; i <= chunks; ++i) {
    if (i < chunks) {
      const size_t at = i * kChunk;
      CheckHip(hipMemcpyAsync(ctx.stage[i % 2].get(), in + at, std::min(kChunk, bytes - at), hipMemcpyDeviceToHost,
                              ctx.stream),
               "hipMemcpyAsync D2H chunk");
      CheckHip(hipEventRecord(ctx.stage_ready[i % 2], ctx.stream), "hipEventRecord");
    }
    if (i > 0) {
      const size_t j = i - 1, at = j * kChunk, len = std::min(kChunk, bytes - at);
      CheckHip(hipEventSynchronize(ctx.stage_ready[j % 2]), "hipEventSynchronize");
      const char* const from = ctx.stage[j % 2].get();
      ParallelFor(kParts, [&](size_t p) {
        const size_t a = len * p / kParts, b = len * (p + 1) / kParts;
        std::memcpy(out + at + a, from + a, b - a);
      }, static_cast<unsigned>(kParts));
    }
  }
}

void Context::Activate() const { CheckHip(hipSetDevice(device), "hipSetDevice"); }

void Context::Sync() const { CheckHip(hipStreamSynchronize(stream), "hipStreamSynchronize"); }

// ---------------------------------------------------------------------- JV
namespace {

void ValidateStaticJv(const ee_graph_view& g) {
  const int32_t n = g.dimension, m = g.edge_count;
  for (int32_t e = 0; e < m; ++e) {
    if (g.edge_u[e] < 0 || g.edge_u[e] >= n || g.edge_v[e] < 0 || g.edge_v[e] >= n) {
      throw std::invalid_argument("JV edge endpoint out of range");
    }
    if (g.edge_weight[e] < 0 || g.edge_weight[e] > (int64_t{1} << 62)) {
      throw std::invalid_argument("JV edge weight outside [0, 2^62]");
    }
  }
  if (!DistanceSafe(g.x, g.y, n)) {
    throw std::runtime_error("HIP JV needs integer coordinates whose squared distances fit uint64");
  }
}

void ValidateDynamicJv(const ee_graph_view& g, int64_t* adjacency) {
  const int32_t n = g.dimension, m = g.edge_count;
  if (g.row_offsets[0] != 0) throw std::invalid_argument("CSR row_offsets[0] must be 0");
  for (int32_t v = 0; v < n; ++v) {
    if (g.row_offsets[v + 1] < g.row_offsets[v]) {
      throw std::invalid_argument("CSR row_offsets must be non-decreasing");
    }
  }
  const int64_t a = g.row_offsets[n];
  for (int64_t i = 0; i < a; ++i) {
    if (g.neighbors[i] < 0 || g.neighbors[i] >= n) {
      throw std::invalid_argument("CSR neighbour out of range");
    }
    if (g.csr_edge_ids[i] < 0 || g.csr_edge_ids[i] >= m) {
      throw std::invalid_argument("CSR edge id out of range");
    }
  }
  *adjacency = a;
}

bool StaticMatches(const JvCache& c, const ee_graph_view& g) {
  const size_t n = static_cast<size_t>(g.dimension), m = static_cast<size_t>(g.edge_count);
  return c.valid && c.dimension == g.dimension && c.distance_type == g.distance_type &&
         c.x.size() == n && c.edge_u.size() == m &&
         std::memcmp(c.x.data(), g.x, n * 8) == 0 && std::memcmp(c.y.data(), g.y, n * 8) == 0 &&
         std::memcmp(c.edge_u.data(), g.edge_u, m * 4) == 0 &&
         std::memcmp(c.edge_v.data(), g.edge_v, m * 4) == 0 &&
         std::memcmp(c.edge_weight.data(), g.edge_weight, m * 8) == 0;
}

}  // namespace

uint64_t JvResidentBytes(const JvCache& c) {
  return c.d_uv.bytes() + c.d_weight.bytes() + c.d_xy.bytes() + c.d_active.bytes() +
         c.d_row_offsets.bytes() + c.d_neighbors.bytes() + c.d_edge_ids.bytes() +
         c.d_witnesses.bytes();
}

// PrepareStaticGraph (jv_cuda.cu:384-430).  Returns true on a cache hit.
bool PrepareJvStatic(Context& ctx, const ee_graph_view& g) {
  JvCache& c = ctx.jv;
  if (StaticMatches(c, g)) return true;
  ValidateStaticJv(g);
  const size_t n = static_cast<size_t>(g.dimension), m = static_cast<size_t>(g.edge_count);
  std::vector<int2> uv(m);
  for (size_t e = 0; e < m; ++e) uv[e] = make_int2(g.edge_u[e], g.edge_v[e]);
  const std::vector<longlong2> xy = Interleave(g.x, g.y, g.dimension);
  DeviceArray<int2> d_uv(m, ctx.device);
  DeviceArray<int64_t> d_weight(m, ctx.device);
  DeviceArray<longlong2> d_xy(n, ctx.device);
  d_uv.Upload(uv.data(), m, ctx.stream);
  d_weight.Upload(g.edge_weight, m, ctx.stream);
  d_xy.Upload(xy.data(), n, ctx.stream);
  ctx.Sync();
  c.d_uv = std::move(d_uv);
  c.d_weight = std::move(d_weight);
  c.d_xy = std::move(d_xy);
  c.x.assign(g.x, g.x + n);
  c.y.assign(g.y, g.y + n);
  c.edge_u.assign(g.edge_u, g.edge_u + m);
  c.edge_v.assign(g.edge_v, g.edge_v + m);
  c.edge_weight.assign(g.edge_weight, g.edge_weight + m);
  c.dimension = g.dimension;
  c.distance_type = g.distance_type;
  c.near = NearExtent(g.x, g.y, g.dimension);
  c.valid = true;
  return false;
}

// PrepareWorkspace (jv_cuda.cu:442-470).  Returns true on a cache hit.
bool PrepareJvWorkspace(Context& ctx, size_t n, size_t m, size_t adjacency) {
  JvCache& c = ctx.jv;
  const bool hit = c.d_active.count() >= m && c.d_row_offsets.count() >= n + 1 &&
                   c.d_neighbors.count() >= adjacency && c.d_edge_ids.count() >= adjacency &&
                   c.d_witnesses.count() >= m;
  if (c.d_active.count() < m) c.d_active.Allocate(Grow(c.d_active.count(), m), ctx.device);
  if (c.d_row_offsets.count() < n + 1) {
    c.d_row_offsets.Allocate(Grow(c.d_row_offsets.count(), n + 1), ctx.device);
  }
  if (c.d_neighbors.count() < adjacency) {
    c.d_neighbors.Allocate(Grow(c.d_neighbors.count(), adjacency), ctx.device);
  }
  if (c.d_edge_ids.count() < adjacency) {
    c.d_edge_ids.Allocate(Grow(c.d_edge_ids.count(), adjacency), ctx.device);
  }
  if (c.d_witnesses.count() < m) c.d_witnesses.Allocate(Grow(c.d_witnesses.count(), m), ctx.device);
  return hit;
}

void CheckGraphView(const ee_graph_view* g) {
  if (g == nullptr) throw std::invalid_argument("graph view is NULL");
  if (g->dimension <= 0 || g->edge_count < 0) throw std::invalid_argument("bad graph dimensions");
  if (g->distance_type != EE_EUC_2D && g->distance_type != EE_CEIL_2D) {
    throw std::invalid_argument("unsupported distance type");
  }
  if (g->integer_coordinates == 0) {
    throw std::runtime_error("HIP JV supports integer coordinates only");
  }
  if (g->x == nullptr || g->y == nullptr || g->row_offsets == nullptr ||
      (g->edge_count > 0 && (g->edge_u == nullptr || g->edge_v == nullptr ||
                             g->edge_weight == nullptr || g->edge_active == nullptr)) ||
      (g->row_offsets[g->dimension] > 0 && (g->neighbors == nullptr || g->csr_edge_ids == nullptr))) {
    throw std::invalid_argument("graph view has a NULL array");
  }
}

JvLaunch JvArgs(Context& ctx, int32_t m, int type) {
  JvCache& c = ctx.jv;
  JvLaunch a{};
  a.edge_count = m;
  a.distance_type = type;
  a.near = c.near ? 1 : 0;
  a.edge_uv = c.d_uv.get();
  a.edge_weight = c.d_weight.get();
  a.edge_active = c.d_active.get();
  a.row_offsets = c.d_row_offsets.get();
  a.neighbors = c.d_neighbors.get();
  a.csr_edge_ids = c.d_edge_ids.get();
  a.xy = c.d_xy.get();
  a.witnesses = c.d_witnesses.get();
  a.one_per_wave = ctx.jv_edges_per_wave == 1 ? 1 : 0;
  return a;
}

// FindJvCandidatesCuda (jv_cuda.cu:488-566) without the final compaction.
void FindJvCandidates(Context& ctx, const ee_graph_view& g, int32_t* witnesses, ee_jv_usage* usage) {
  ee_jv_usage u{};
  ctx.Activate();
  const int32_t m = g.edge_count;
  if (m == 0) {
    if (usage) *usage = u;
    return;
  }
  int64_t adjacency = 0;
  ValidateDynamicJv(g, &adjacency);
  const bool static_hit = PrepareJvStatic(ctx, g);
  const bool workspace_hit =
      PrepareJvWorkspace(ctx, static_cast<size_t>(g.dimension), static_cast<size_t>(m),
                         static_cast<size_t>(adjacency));
  JvCache& c = ctx.jv;
  c.host_active.resize(static_cast<size_t>(m));
  for (int32_t e = 0; e < m; ++e) c.host_active[static_cast<size_t>(e)] = g.edge_active[e] ? 1 : 0;

  auto t0 = std::chrono::steady_clock::now();
  c.d_active.Upload(c.host_active.data(), static_cast<size_t>(m), ctx.stream);
  c.d_row_offsets.Upload(g.row_offsets, static_cast<size_t>(g.dimension) + 1, ctx.stream);
  c.d_neighbors.Upload(g.neighbors, static_cast<size_t>(adjacency), ctx.stream);
  c.d_edge_ids.Upload(g.csr_edge_ids, static_cast<size_t>(adjacency), ctx.stream);
  ctx.Sync();
  u.h2d_ms = MsSince(t0);
  u.static_hit = static_hit ? 1 : 0;
  u.workspace_hit = workspace_hit ? 1 : 0;
  u.resident_bytes = JvResidentBytes(c);

  t0 = std::chrono::steady_clock::now();
  LaunchJv(JvArgs(ctx, m, g.distance_type), ctx.jv_grid, ctx.stream);
  ctx.Sync();
  u.kernel_ms = MsSince(t0);

  t0 = std::chrono::steady_clock::now();
  c.d_witnesses.Download(witnesses, static_cast<size_t>(m), ctx.stream);
  ctx.Sync();
  u.d2h_ms = MsSince(t0);
  if (usage) *usage = u;
}

// ------------------------------------------------------------------- k-opt
uint64_t KoptResidentBytes(const KoptCache& c) {
  return c.d_xy.bytes() + c.d_plans.bytes() + c.d_tasks.bytes() + c.d_costs.bytes() +
         c.d_tiles.bytes();
}

uint32_t TileRows(const KoptTuning& tuning, uint32_t k) { return tuning.rows_per_tile[k - 3]; }

void EvalKoptCosts(Context& ctx, const int64_t* x, const int64_t* y, int32_t n, uint8_t type,
                   uint32_t k, const ee_endpoint_matching* templates, uint32_t template_count,
                   uint64_t generator_hash, const ee_kopt_task* tasks, size_t task_count,
                   int64_t* costs, ee_kopt_usage* usage) {
  ctx.Activate();
  if (task_count == 0) {
    if (usage) *usage = ee_kopt_usage{};
    return;
  }
  if (tasks == nullptr || costs == nullptr) throw std::invalid_argument("HIP k-opt NULL buffer");
  ValidateTasks(n, k, tasks, task_count);
  const KoptTaskSource upload = [&](ee_kopt_task* d_tasks) {
    CheckHip(hipMemcpyAsync(d_tasks, tasks, task_count * sizeof(ee_kopt_task), hipMemcpyHostToDevice,
                            ctx.stream),
             "hipMemcpyAsync tasks");
  };
  EvalKoptCostsFrom(ctx, x, y, n, type, k, templates, template_count, generator_hash, task_count,
                    upload, true, costs, nullptr, usage);
}

void EvalKoptCostsFrom(Context& ctx, const int64_t* x, const int64_t* y, int32_t n, uint8_t type,
                       uint32_t k, const ee_endpoint_matching* templates, uint32_t template_count,
                       uint64_t generator_hash, size_t task_count, const KoptTaskSource& fill_tasks,
                       bool sync_after_fill, int64_t* costs, uint32_t* masks, ee_kopt_usage* usage) {
  ee_kopt_usage u{};
  ctx.Activate();
  if (task_count == 0) {
    if (usage) *usage = u;
    return;
  }
  if (k < 3 || k > 5 || templates == nullptr || template_count == 0) {
    throw std::invalid_argument("HIP k-opt reconnect table is invalid");
  }
  if (x == nullptr || y == nullptr || n <= 0 || (type != EE_EUC_2D && type != EE_CEIL_2D)) {
    throw std::invalid_argument("HIP k-opt coordinates are invalid");
  }
  if ((costs == nullptr) == (masks == nullptr)) {
    throw std::invalid_argument("HIP k-opt needs exactly one of costs and masks");
  }
  if (task_count > SIZE_MAX / template_count) throw std::overflow_error("k-opt cell count overflows");
  const size_t cells = task_count * template_count;
  const uint32_t mask_words = (template_count + 31) / 32;
  KoptCache& c = ctx.kopt;

  // snapshot
  const size_t nn = static_cast<size_t>(n);
  const bool snapshot_hit = c.snapshot_valid && c.dimension == n && c.distance_type == type &&
                            c.x.size() == nn && std::memcmp(c.x.data(), x, nn * 8) == 0 &&
                            std::memcmp(c.y.data(), y, nn * 8) == 0;
  if (!snapshot_hit) {
    if (!DistanceSafe(x, y, n)) {
      throw std::invalid_argument("k-opt cost needs distance-safe integer coordinates");
    }
    const std::vector<longlong2> xy = Interleave(x, y, n);
    DeviceArray<longlong2> d_xy(nn, ctx.device);
    d_xy.Upload(xy.data(), nn, ctx.stream);
    ctx.Sync();
    c.d_xy = std::move(d_xy);
    c.x.assign(x, x + nn);
    c.y.assign(y, y + nn);
    c.dimension = n;
    c.distance_type = type;
    c.near = NearExtent(x, y, n);
    c.snapshot_valid = true;
  }

  // templates (one plan slot per k)
  KoptCache::TemplateEntry& entry = c.templates[k - 3];
  const bool template_hit =
      entry.valid && entry.hash == generator_hash && entry.table.size() == template_count &&
      std::memcmp(entry.table.data(), templates, template_count * sizeof(ee_endpoint_matching)) == 0;
  if (!template_hit) {
    const KoptPlan plan = BuildPlan(k, templates, template_count);
    if (c.d_plans.count() < 3) c.d_plans.Allocate(3, ctx.device);
    c.host_plans[k - 3] = plan;
    c.d_plans.UploadAt(k - 3, &c.host_plans[k - 3], 1, ctx.stream);
    ctx.Sync();
    entry.table.assign(templates, templates + template_count);
    entry.hash = generator_hash;
    entry.valid = true;
  }

  // workspace
  // mask output needs task_count x mask_words words, held in the cost pool
  const size_t out_cells = masks ? (task_count * mask_words + 1) / 2 : cells;
  const bool workspace_hit = c.d_tasks.count() >= task_count && c.d_costs.count() >= out_cells;
  if (c.d_tasks.count() < task_count) c.d_tasks.Allocate(Grow(c.d_tasks.count(), task_count), ctx.device);
  if (c.d_costs.count() < out_cells) c.d_costs.Allocate(Grow(c.d_costs.count(), out_cells), ctx.device);

  KoptInstance inst{};
  inst.n = static_cast<uint32_t>(n);
  inst.distance_type = type;
  inst.near = c.near ? 1 : 0;
  KoptSegment segment{0, k, 0, task_count, 0, 0, TileRows(ctx.kopt_tuning, k), 0};
  const uint64_t n_tiles = KoptCountTiles(&segment, 1, ctx.kopt_tuning);
  std::vector<KoptTile> tiles(n_tiles);
  KoptLaunch a{};
  KoptAssignTiles(&segment, 1, &inst, c.host_plans.data(), ctx.kopt_tuning, tiles.data(), &a);
  // a one-segment launch derives its tiles from the kernel arguments
  const bool descriptors = a.inline_count == 0;
  if (descriptors && c.d_tiles.count() < n_tiles) c.d_tiles.Allocate(Grow(c.d_tiles.count(), n_tiles), ctx.device);
  auto t0 = std::chrono::steady_clock::now();
  fill_tasks(c.d_tasks.get());
  if (descriptors) c.d_tiles.Upload(tiles.data(), n_tiles, ctx.stream);
  if (sync_after_fill) ctx.Sync();
  u.h2d_ms = MsSince(t0);

  a.xy = c.d_xy.get();
  a.tiles = c.d_tiles.get();
  a.tasks = c.d_tasks.get();
  a.plans = c.d_plans.get();
  a.costs = c.d_costs.get();
  if (masks) {
    a.masks = reinterpret_cast<uint32_t*>(c.d_costs.get());
    a.mask_words = mask_words;
  }
  bool present[3] = {false, false, false};
  present[k - 3] = true;
  const size_t shared = KoptPrepareLayout(&a, c.host_plans.data(), present, ctx.kopt_tuning, c.near);
  t0 = std::chrono::steady_clock::now();
  LaunchKopt(a, ctx.kopt_tuning, shared, ctx.stream);
  ctx.Sync();
  u.kernel_ms = MsSince(t0);

  t0 = std::chrono::steady_clock::now();
  if (masks) {
    CheckHip(hipMemcpyAsync(masks, c.d_costs.get(), task_count * mask_words * sizeof(uint32_t),
                            hipMemcpyDeviceToHost, ctx.stream),
             "hipMemcpyAsync masks");
  } else {
    if (cells > c.d_costs.count()) throw std::logic_error("D2H beyond device buffer");
    DownloadToCaller(ctx, costs, c.d_costs.get(), cells * sizeof(int64_t));
  }
  ctx.Sync();
  u.d2h_ms = MsSince(t0);
  u.snapshot_hit = snapshot_hit ? 1 : 0;
  u.template_hit = template_hit ? 1 : 0;
  u.workspace_hit = workspace_hit ? 1 : 0;
  u.resident_bytes = KoptResidentBytes(c);
  if (usage) *usage = u;
}

}  // namespace ee

// =================================================================== C ABI
using ee::StatusFromException;

extern "C" {

int ee_device_count(void) { return ee::CountDevices(); }

ee_status ee_backend_available(int device, char* reason, size_t reason_capacity) {
  std::string why;
  const int count = ee::CountDevices();
  if (count <= 0) {
    why = "no visible HIP device";
  } else if (device >= count) {
    why = "requested HIP device ordinal is not visible";
  }
  if (reason && reason_capacity > 0) {
    const size_t len = std::min(why.size(), reason_capacity - 1);
    std::memcpy(reason, why.data(), len);
    reason[len] = '\0';
  }
  if (!why.empty()) {
    ee::SetLastError(why);
    return EE_ERR_UNAVAILABLE;
  }
  return EE_OK;
}

ee_status ee_ctx_create(int device, ee_ctx** out) {
  try {
    if (out == nullptr) throw std::invalid_argument("ee_ctx_create: out is NULL");
    *out = nullptr;
    *out = reinterpret_cast<ee_ctx*>(new ee::Context(device));
    return EE_OK;
  } catch (...) {
    return StatusFromException();
  }
}

void ee_ctx_destroy(ee_ctx* ctx) { delete reinterpret_cast<ee::Context*>(ctx); }

int ee_ctx_device(const ee_ctx* ctx) {
  return ctx ? reinterpret_cast<const ee::Context*>(ctx)->device : -1;
}

ee_status ee_ctx_set_jv_edges_per_wave(ee_ctx* ctx, uint32_t edges) {
  try {
    if (ctx == nullptr) throw std::invalid_argument("context is NULL");
    if (edges != 1 && edges != 2) throw std::invalid_argument("JV edges per wavefront must be 1 or 2");
    reinterpret_cast<ee::Context*>(ctx)->jv_edges_per_wave = edges;
    return EE_OK;
  } catch (...) {
    return StatusFromException();
  }
}

ee_status ee_jv_find_candidates(ee_ctx* ctx, const ee_graph_view* graph, int32_t* witnesses,
                                ee_jv_usage* usage) {
  try {
    if (ctx == nullptr) throw std::invalid_argument("context is NULL");
    ee::CheckGraphView(graph);
    if (graph->edge_count > 0 && witnesses == nullptr) {
      throw std::invalid_argument("witness buffer is NULL");
    }
    ee::FindJvCandidates(*reinterpret_cast<ee::Context*>(ctx), *graph, witnesses, usage);
    return EE_OK;
  } catch (...) {
    return StatusFromException();
  }
}

ee_status ee_jv_clear_cache(ee_ctx* ctx) {
  try {
    if (ctx == nullptr) throw std::invalid_argument("context is NULL");
    auto* c = reinterpret_cast<ee::Context*>(ctx);
    c->Activate();
    c->Sync();
    c->jv = ee::JvCache{};
    return EE_OK;
  } catch (...) {
    return StatusFromException();
  }
}

ee_status ee_kopt_eval_costs(ee_ctx* ctx, const int64_t* x, const int64_t* y, int32_t dimension,
                             uint8_t distance_type, uint32_t k,
                             const ee_endpoint_matching* templates, uint32_t template_count,
                             uint64_t generator_hash, const ee_kopt_task* tasks,
                             size_t task_count, int64_t* costs, ee_kopt_usage* usage) {
  try {
    if (ctx == nullptr) throw std::invalid_argument("context is NULL");
    ee::EvalKoptCosts(*reinterpret_cast<ee::Context*>(ctx), x, y, dimension, distance_type, k,
                      templates, template_count, generator_hash, tasks, task_count, costs, usage);
    return EE_OK;
  } catch (...) {
    return StatusFromException();
  }
}

ee_status ee_kopt_clear_cache(ee_ctx* ctx) {
  try {
    if (ctx == nullptr) throw std::invalid_argument("context is NULL");
    auto* c = reinterpret_cast<ee::Context*>(ctx);
    c->Activate();
    c->Sync();
    c->kopt = ee::KoptCache{};
    return EE_OK;
  } catch (...) {
    return StatusFromException();
  }
}

}  // extern "C"
