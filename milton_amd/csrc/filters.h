// Reconstruction filters shared by the film (device) and the environment
// lookup (host): Filter2D::evaluate of filters/{Box,Triangle,Gaussian,
// Mitchell,LanczosSinc}Filter.cpp in double precision, for the device and
// the host alike.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/milton_pt.h"

namespace mpt_filter {

// Resolved parameters: support and (p0, p1) with the reference's defaults.
struct Params {
  uint32_t type;   // MPT_FILTER_*
  double support, p0, p1;
};

__host__ __device__ inline Params resolve(uint32_t type, uint32_t support, float p0, float p1) {
  Params p{type, support ? (double)support : 2.0, p0, p1};
  if (type == MPT_FILTER_GAUSSIAN && p.p0 == 0) p.p0 = 1.0;                      // sigma
  if (type == MPT_FILTER_MITCHELL && p.p0 == 0 && p.p1 == 0) p.p0 = p.p1 = 1.0 / 3.0;   // B, C
  if (type == MPT_FILTER_LANCZOS && p.p0 == 0) p.p0 = 3.0;                       // tau
  return p;
}

__host__ __device__ inline double mitchell1(const Params &p, double v) {   // MitchellFilter.cpp:31-43
  v = fabs(2.0 * v / p.support);
  const double B = p.p0, C = p.p1;
  if (v > 1.0)
    return (((-B / 6.0 - C) * v + (B + 5.0 * C)) * v + (-2.0 * B - 8.0 * C)) * v + (4.0 / 3.0 * B + 4.0 * C);
  return v * v * ((2.0 - 1.5 * B - C) * v + (-3.0 + 2.0 * B + C)) + (1.0 - B / 3.0);
}

__host__ __device__ inline double lanczos1(const Params &p, double v) {   // LanczosSincFilter.cpp:30-45
  v = fabs(v / p.support);
  if (v < 1e-7) return 1.0;
  if (v > 1) return 0.0;
  v *= 3.14159265358979323846;
  return sin(v * p.p0) / (v * p.p0) * (sin(v) / v);
}

// The weight at offset (x, y) from the sample.
__host__ __device__ inline double weight(const Params &p, double x, double y) {
  switch (p.type) {
    case MPT_FILTER_BOX:   // inside the disc of diameter support (BoxFilter.cpp:14-16)
      return 2.0 * sqrt(x * x + y * y) < p.support ? 1.0 : 0.0;
    case MPT_FILTER_TRIANGLE: {   // signed offsets, as the reference evaluates them (TriangleFilter.cpp:14-18)
      const double r = p.support / 2.0;
      return fmax(0.0, 1.0 - x / r) * fmax(0.0, 1.0 - y / r);
    }
    case MPT_FILTER_GAUSSIAN: {   // GaussianFilter.cpp:22-31
      const double e = (-0.5) / (p.p0 * p.p0) * (x * x + y * y);
      return fmax(0.0, 1.0 / (sqrt(2.0 * 3.14159265358979323846) * p.p0) * exp(e));
    }
    case MPT_FILTER_MITCHELL:
      return mitchell1(p, x) * mitchell1(p, y);
    case MPT_FILTER_LANCZOS:
      return lanczos1(p, x) * lanczos1(p, y);
    default:
      return 0.0;
  }
}

}  // namespace mpt_filter
