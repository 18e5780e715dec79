// Element-wise square roots and distances on the device, so the exact
// distance code the kernels inline can be checked on its own.
#include <vector>

#include "ee_exact.h"
#include "ee_gen.h"
#include "ee_runtime.h"

namespace ee {
namespace {

__global__ void IsqrtKernel(const uint64_t* s, uint64_t* out, size_t count) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < count;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    out[i] = IsqrtFloor(s[i]);
  }
}

// `near`: coordinate extents below 2^25, the FP64 path the k-opt kernel takes.
__global__ void DistanceKernel(const longlong2* xy, const int32_t* a, const int32_t* b,
                               size_t count, int type, int near, int64_t* out) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < count;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const longlong2 p = xy[a[i]], q = xy[b[i]];
    if (near) {
      const int32_t dx = static_cast<int32_t>(static_cast<uint32_t>(p.x) - static_cast<uint32_t>(q.x));
      const int32_t dy = static_cast<int32_t>(static_cast<uint32_t>(p.y) - static_cast<uint32_t>(q.y));
      out[i] = ExactDistanceNear(dx, dy, type);
    } else {
      out[i] = ExactDistance(p.x, p.y, q.x, q.y, type);
    }
  }
}

uint32_t Blocks(size_t count) {
  const size_t blocks = (count + 255) / 256;
  return static_cast<uint32_t>(blocks < 65536 ? blocks : 65536);
}

}  // namespace
}  // namespace ee

extern "C" {

void ee_isqrt_host(const uint64_t* s, uint64_t* root, size_t count) {
  for (size_t i = 0; i < count; ++i) root[i] = ee::IsqrtFloor(s[i]);
}

ee_status ee_isqrt_device(ee_ctx* handle, const uint64_t* s, uint64_t* root, size_t count) {
  try {
    if (handle == nullptr || (count > 0 && (s == nullptr || root == nullptr))) {
      throw std::invalid_argument("ee_isqrt_device: NULL argument");
    }
    if (count > (size_t{1} << 28)) throw std::invalid_argument("ee_isqrt_device: count too large");
    if (count == 0) return EE_OK;
    auto& ctx = *reinterpret_cast<ee::Context*>(handle);
    ctx.Activate();
    ee::DeviceArray<uint64_t> d_in(count, ctx.device), d_out(count, ctx.device);
    d_in.Upload(s, count, ctx.stream);
    hipLaunchKernelGGL(ee::IsqrtKernel, dim3(ee::Blocks(count)), dim3(256), 0, ctx.stream,
                       d_in.get(), d_out.get(), count);
    ee::CheckHip(hipGetLastError(), "isqrt kernel launch");
    d_out.Download(root, count, ctx.stream);
    ctx.Sync();
    return EE_OK;
  } catch (...) {
    return ee::StatusFromException();
  }
}

ee_status ee_distances_device(ee_ctx* handle, const int64_t* x, const int64_t* y, int32_t n,
                              uint8_t distance_type, const int32_t* a, const int32_t* b,
                              size_t count, int64_t* out) {
  try {
    if (handle == nullptr || x == nullptr || y == nullptr || n <= 0 ||
        (count > 0 && (a == nullptr || b == nullptr || out == nullptr))) {
      throw std::invalid_argument("ee_distances_device: bad argument");
    }
    if (distance_type != EE_EUC_2D && distance_type != EE_CEIL_2D) {
      throw std::invalid_argument("ee_distances_device: unsupported distance type");
    }
    if (!ee::DistanceSafe(x, y, n)) throw std::invalid_argument("coordinates are not distance safe");
    for (size_t i = 0; i < count; ++i) {
      if (a[i] < 0 || a[i] >= n || b[i] < 0 || b[i] >= n) {
        throw std::invalid_argument("ee_distances_device: node out of range");
      }
    }
    if (count == 0) return EE_OK;
    auto& ctx = *reinterpret_cast<ee::Context*>(handle);
    ctx.Activate();
    std::vector<longlong2> xy(static_cast<size_t>(n));
    for (int32_t i = 0; i < n; ++i) xy[static_cast<size_t>(i)] = longlong2{x[i], y[i]};
    ee::DeviceArray<longlong2> d_xy(static_cast<size_t>(n), ctx.device);
    ee::DeviceArray<int32_t> d_a(count, ctx.device), d_b(count, ctx.device);
    ee::DeviceArray<int64_t> d_out(count, ctx.device);
    d_xy.Upload(xy.data(), static_cast<size_t>(n), ctx.stream);
    d_a.Upload(a, count, ctx.stream);
    d_b.Upload(b, count, ctx.stream);
    hipLaunchKernelGGL(ee::DistanceKernel, dim3(ee::Blocks(count)), dim3(256), 0, ctx.stream,
                       d_xy.get(), d_a.get(), d_b.get(), count, static_cast<int>(distance_type),
                       ee::NearExtent(x, y, n) ? 1 : 0, d_out.get());
    ee::CheckHip(hipGetLastError(), "distance kernel launch");
    d_out.Download(out, count, ctx.stream);
    ctx.Sync();
    return EE_OK;
  } catch (...) {
    return ee::StatusFromException();
  }
}

}  // extern "C"
