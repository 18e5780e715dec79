// The plain k-opt cost evaluation without any allocation
// (ee_kopt_eval_costs_noalloc).
//
// ee_kopt_eval_costs stages the coordinates, the reconnect plan, the task rows
// and the cost cells in device buffers its context allocates and grows.  This
// variant stages them in `scratch`, one static device array per GPU of
// EE_NOALLOC_SCRATCH_BYTES, and runs the same kernel (KoptKernel, one
// segment, the context's tuning) on them, so its cells are the plain path's
// bit for bit.  Host side it uses only the stack: the plan is built in place
// and the x / y arrays are interleaved on the device by a small kernel.
//
// Every call on a GPU shares that GPU's scratch array, which keeps the last
// call's contents; every region a call reads is written by the same call
// first, so nothing needs clearing.  Calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <stdexcept>
#include <string>

#include "ee_runtime.h"

static __device__ __align__(256) unsigned char scratch[EE_NOALLOC_SCRATCH_BYTES];

namespace ee {
namespace {

constexpr size_t kScratchBytes = EE_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxScratchDevices = 64;

// Device address of each GPU's `scratch` (the array's name in host code is a
// host-side stand-in, not a device address).
std::mutex g_scratch_mutex;
unsigned char* g_scratch[kMaxScratchDevices] = {};

__global__ void __launch_bounds__(256) InterleaveKernel(const int64_t* __restrict__ x,
                                                        const int64_t* __restrict__ y,
                                                        longlong2* __restrict__ xy, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    xy[i] = longlong2{x[i], y[i]};
  }
}

size_t Align256(size_t v) { return (v + 255) & ~size_t{255}; }

// Byte offsets of one call's regions inside the scratch array.
struct ScratchLayout {
  size_t plans, x, y, xy, tasks, costs, end;
};

ScratchLayout Layout(size_t n, size_t task_count, size_t cells) {
  ScratchLayout s{};
  s.plans = 0;
  s.x = Align256(s.plans + 3 * sizeof(KoptPlan));
  s.y = Align256(s.x + 8 * n);
  s.xy = Align256(s.y + 8 * n);
  s.tasks = Align256(s.xy + 16 * n);
  s.costs = Align256(s.tasks + task_count * sizeof(ee_kopt_task));
  s.end = s.costs + 8 * cells;
  return s;
}

unsigned char* ScratchOf(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (device < 0 || device >= kMaxScratchDevices || g_scratch[device] == nullptr) {
    throw std::logic_error("no-allocation scratch of this device was not set up by its context");
  }
  return g_scratch[device];
}

void EvalNoAlloc(Context& ctx, const int64_t* x, const int64_t* y, int32_t n, uint8_t type, uint32_t k,
                 const ee_endpoint_matching* templates, uint32_t template_count, const ee_kopt_task* tasks,
                 size_t task_count, int64_t* costs, ee_kopt_usage* usage) {
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
  if (tasks == nullptr || costs == nullptr) throw std::invalid_argument("HIP k-opt NULL buffer");
  if (task_count > kScratchBytes || template_count > kScratchBytes) {
    throw std::invalid_argument("k-opt evaluation does not fit the no-allocation scratch");
  }
  const size_t cells = task_count * template_count;
  const ScratchLayout s = Layout(static_cast<size_t>(n), task_count, cells);
  if (s.end > kScratchBytes) {
    throw std::invalid_argument("k-opt evaluation needs " + std::to_string(s.end) +
                                " bytes of no-allocation scratch; the maximum is " +
                                std::to_string(kScratchBytes));
  }
  ValidateTasks(n, k, tasks, task_count);
  if (!DistanceSafe(x, y, n)) throw std::invalid_argument("k-opt cost needs distance-safe integer coordinates");
  const bool near = NearExtent(x, y, n);
  const KoptPlan plan = BuildPlan(k, templates, template_count);

  unsigned char* base = ScratchOf(ctx.device);
  auto* d_plans = reinterpret_cast<KoptPlan*>(base + s.plans);
  auto* d_x = reinterpret_cast<int64_t*>(base + s.x);
  auto* d_y = reinterpret_cast<int64_t*>(base + s.y);
  auto* d_xy = reinterpret_cast<longlong2*>(base + s.xy);
  auto* d_tasks = reinterpret_cast<ee_kopt_task*>(base + s.tasks);
  auto* d_costs = reinterpret_cast<int64_t*>(base + s.costs);

  auto t0 = std::chrono::steady_clock::now();
  // the kernel reads only the plan slot of its k
  CheckHip(hipMemcpyAsync(d_plans + (k - 3), &plan, sizeof(KoptPlan), hipMemcpyHostToDevice, ctx.stream),
           "hipMemcpyAsync plan");
  CheckHip(hipMemcpyAsync(d_x, x, 8 * static_cast<size_t>(n), hipMemcpyHostToDevice, ctx.stream),
           "hipMemcpyAsync x");
  CheckHip(hipMemcpyAsync(d_y, y, 8 * static_cast<size_t>(n), hipMemcpyHostToDevice, ctx.stream),
           "hipMemcpyAsync y");
  CheckHip(hipMemcpyAsync(d_tasks, tasks, task_count * sizeof(ee_kopt_task), hipMemcpyHostToDevice,
                          ctx.stream),
           "hipMemcpyAsync tasks");
  const uint32_t blocks = static_cast<uint32_t>(std::min<size_t>((static_cast<size_t>(n) + 255) / 256, 1024));
  hipLaunchKernelGGL(InterleaveKernel, dim3(blocks), dim3(256), 0, ctx.stream, d_x, d_y, d_xy,
                     static_cast<uint32_t>(n));
  CheckHip(hipGetLastError(), "interleave kernel launch");
  ctx.Sync();
  u.h2d_ms = MsSince(t0);

  KoptInstance inst{};
  inst.n = static_cast<uint32_t>(n);
  inst.distance_type = type;
  inst.near = near ? 1 : 0;
  const KoptTuning& tuning = ctx.kopt_tuning;
  KoptSegment segment{0, k, 0, task_count, 0, 0, tuning.rows_per_tile[k - 3], 0};
  KoptPlan host_plans[3];
  host_plans[k - 3] = plan;
  KoptLaunch a{};
  // one segment: the kernel derives its tiles from the inline segment table
  KoptAssignTiles(&segment, 1, &inst, host_plans, tuning, nullptr, &a);
  if (a.inline_count == 0) throw std::logic_error("no-allocation k-opt launch needs tile descriptors");
  a.xy = d_xy;
  a.tiles = nullptr;
  a.tasks = d_tasks;
  a.plans = d_plans;
  a.costs = d_costs;
  bool present[3] = {false, false, false};
  present[k - 3] = true;
  const size_t shared = KoptPrepareLayout(&a, host_plans, present, tuning, near);
  t0 = std::chrono::steady_clock::now();
  LaunchKopt(a, tuning, shared, ctx.stream);
  ctx.Sync();
  u.kernel_ms = MsSince(t0);

  t0 = std::chrono::steady_clock::now();
  CheckHip(hipMemcpyAsync(costs, d_costs, cells * sizeof(int64_t), hipMemcpyDeviceToHost, ctx.stream),
           "hipMemcpyAsync costs");
  ctx.Sync();
  u.d2h_ms = MsSince(t0);
  u.resident_bytes = s.end;
  if (usage) *usage = u;
}

}  // namespace

void InitNoAllocScratch(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (device < 0 || device >= kMaxScratchDevices) {
    throw Unavailable("device ordinal beyond the no-allocation scratch table");
  }
  if (g_scratch[device] != nullptr) return;
  CheckHip(hipSetDevice(device), "hipSetDevice");
  void* p = nullptr;
  CheckHip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
  if (p == nullptr) throw Unavailable("no-allocation scratch has no device address");
  g_scratch[device] = static_cast<unsigned char*>(p);
}

}  // namespace ee

extern "C" ee_status ee_kopt_eval_costs_noalloc(ee_ctx* ctx, const int64_t* x, const int64_t* y,
                                                int32_t dimension, uint8_t distance_type, uint32_t k,
                                                const ee_endpoint_matching* templates,
                                                uint32_t template_count, uint64_t generator_hash,
                                                const ee_kopt_task* tasks, size_t task_count,
                                                int64_t* costs, ee_kopt_usage* usage) {
  try {
    if (ctx == nullptr) throw std::invalid_argument("context is NULL");
    static_cast<void>(generator_hash);  // only the cached path keys on it
    ee::EvalNoAlloc(*reinterpret_cast<ee::Context*>(ctx), x, y, dimension, distance_type, k, templates,
                    template_count, tasks, task_count, costs, usage);
    return EE_OK;
  } catch (...) {
    return ee::StatusFromException();
  }
}
