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
  const uint64_t up = type == 0 ? static_cast<uint64_t>(s - rr > r) : static_cast<uint64_t>(rr != s);
  return static_cast<int64_t>(r + up);
}

// Largest coordinate extent (max - min, per axis) for ExactDistanceNear.
constexpr int64_t kNearExtent = int64_t{1} << 25;

// Same result as ExactDistance for |dx|, |dy| < 2^25, entirely in FP64:
// s = dx^2 + dy^2 < 2^51, so s, r*r, (r+1)^2 and s - r*r are all exact
// doubles, and the two integer-valued corrections move any root estimate
// within one unit of floor(sqrt(s)) onto it.
EE_HD int64_t ExactDistanceNear(int32_t dx, int32_t dy, int type) {
  const double fx = static_cast<double>(dx), fy = static_cast<double>(dy);
  const double s = fma(fx, fx, fy * fy);
  double r = floor(sqrt(s));
  if (r * r > s) r -= 1.0;
  if ((r + 1.0) * (r + 1.0) <= s) r += 1.0;
  const double rem = s - r * r;
  const bool up = type == 0 ? rem > r : rem != 0.0;
  return static_cast<int64_t>(r) + (up ? 1 : 0);
}

}  // namespace ee
