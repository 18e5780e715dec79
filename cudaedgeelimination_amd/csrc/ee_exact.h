// Exact TSPLIB integer distances, shared by host and device code.
//
// The reference computes r = floor(sqrt(S)) with a 32-step integer restoring
// loop (src/common/distance.cpp:9-27, src/cuda/jv_cuda.cu:141-158).  Here r
// comes from the IEEE FP64 square root of S, followed by an integer
// correction that makes r*r <= S < (r+1)^2 hold exactly.  For S < 2^53 the
// conversion (double)S is exact and the correctly rounded root is within one
// unit of the answer; for larger S the correction loops still terminate at the
// exact root because r is clamped to the largest value whose square fits in
// 64 bits.  The rounding rules are the reference's (src/common/distance.cpp:44-52):
// EUC_2D adds one when S - r^2 > r, CEIL_2D adds one when r^2 != S.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define EE_HD __host__ __device__ inline
#else
#define EE_HD inline
#include <cmath>
#endif

namespace ee {

constexpr uint64_t kRootMax = 0xFFFFFFFFull;  // largest r with r*r < 2^64

EE_HD uint64_t IsqrtFloor(uint64_t s) {
  uint64_t r = static_cast<uint64_t>(sqrt(static_cast<double>(s)));
  if (r > kRootMax) r = kRootMax;
  while (r * r > s) --r;
  while (r < kRootMax && (r + 1) * (r + 1) <= s) ++r;
  return r;
}

EE_HD uint64_t AbsDiff(int64_t a, int64_t b) {
  return a >= b ? static_cast<uint64_t>(a) - static_cast<uint64_t>(b)
                : static_cast<uint64_t>(b) - static_cast<uint64_t>(a);
}

// Caller guarantees dx^2 + dy^2 fits in uint64 (integer_distance_safe).
EE_HD int64_t ExactDistance(int64_t ax, int64_t ay, int64_t bx, int64_t by, int type) {
  const uint64_t dx = AbsDiff(ax, bx);
  const uint64_t dy = AbsDiff(ay, by);
  const uint64_t s = dx * dx + dy * dy;
  const uint64_t r = IsqrtFloor(s);
  const uint64_t rr = r * r;
  // rr <= s, so for CEIL_2D "rr != s" is "s - rr > 0"
  const uint64_t up = static_cast<uint64_t>(s - rr > (type == 0 ? r : 0));
  return static_cast<int64_t>(r + up);
}

// Largest coordinate extent (max - min, per axis) for ExactDistanceNear.
constexpr int64_t kNearExtent = int64_t{1} << 25;

// Same result as ExactDistance for |dx|, |dy| < 2^25, entirely in FP64:
// s = dx^2 + dy^2 < 2^51, so s, r*r and s - r*r are exact doubles for every
// r below 2^26.  With m = floor(sqrt(s)) and an estimate of sqrt(s) within
// e < 1/4 of it, r = floor(estimate) is m, or m - 1 when sqrt(s) < m + e, or
// m + 1 when sqrt(s) > m + 1 - e.  The rounding then needs no step onto m:
//  - EUC_2D is m + (s - m^2 > m); with rem = s - r^2 it is r + (rem > r):
//    r = m - 1 gives rem = s - m^2 + 2m - 1 > r and m + 1 is not taken since
//    s - m^2 < 2me + e^2 <= m; r = m + 1 gives rem < 0 and indeed
//    s - m^2 > 2m + 1 - 2e(m + 1) > m.
//  - CEIL_2D is m + (s != m^2); it is r + (rem > 0) + (rem > 2r + 1): for
//    r = m - 1, rem = s - m^2 + 2r + 1; for r = m + 1, rem < 0 and s > m^2.
// On the device the estimate is the hardware reciprocal square root
// (relative error about 2^-23) refined by one Newton step (about 2^-45, i.e.
// below 2^-19 absolute for roots below 2^26); s = 0 is evaluated as 1/4,
// whose root floors to 0 as well.
EE_HD int64_t ExactDistanceNear(int32_t dx, int32_t dy, int type) {
  const double fx = static_cast<double>(dx), fy = static_cast<double>(dy);
#if defined(__HIP_DEVICE_COMPILE__)
  const double s = fma(fx, fx, fy * fy);
  const double s4 = fmax(s, 0.25);  // rsq(0) is +inf
  const double y = __builtin_amdgcn_rsq(s4);
  const double g = s4 * y, h = 0.5 * y;
  const double r = floor(fma(g, fma(-h, g, 0.5), g));
  const double rem = fma(-r, r, s);
  const double twice_r1 = fma(2.0, r, 1.0);
#else
  // Every value is an integer below 2^53 (squares below 2^50, s below 2^51,
  // r * r below 2^52), so plain products and sums are exact and give what
  // the fused forms give; without FMA code generation for the host, fma()
  // would be a library call per use.
  const double s = fx * fx + fy * fy;
  const double r = floor(sqrt(s));
  const double rem = s - r * r;
  const double twice_r1 = 2.0 * r + 1.0;
#endif
  const uint32_t up = type == 0 ? static_cast<uint32_t>(rem > r)
                                : static_cast<uint32_t>(rem > 0.0) + static_cast<uint32_t>(rem > twice_r1);
  return static_cast<int64_t>(static_cast<uint32_t>(r) + up);
}

}  // namespace ee
