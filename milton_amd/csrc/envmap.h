// Environment-image support shared by the scene upload and the scene-file
// loader: the reconstruction kernel an EnvironmentMap lookup applies, built
// on the host in double precision.
#pragma once

#include <cmath>
#include <string>
#include <vector>

#include "../../include/milton_pt.h"

namespace mpt_env {

constexpr uint32_t kMaxKernelWidth = 15;

// One filter's weight at the sample offset (x, y), for the formulas of
// filters/{Box,Triangle,Gaussian,Mitchell,LanczosSinc}Filter.cpp.
inline double filter_weight(uint32_t filter, double support, const double p[2], double x, double y) {
  switch (filter) {
    case MPT_FILTER_BOX:   // inside the disc of diameter support
      return 2.0 * std::sqrt(x * x + y * y) < support ? 1.0 : 0.0;
    case MPT_FILTER_TRIANGLE: {   // signed offsets, as the reference evaluates them
      const double r = support / 2.0;
      return std::max(0.0, 1.0 - x / r) * std::max(0.0, 1.0 - y / r);
    }
    case MPT_FILTER_GAUSSIAN: {
      const double sigma = p[0];
      const double e = (-0.5) / (sigma * sigma) * (x * x + y * y);
      return std::max(0.0, 1.0 / (std::sqrt(2.0 * M_PI) * sigma) * std::exp(e));
    }
    case MPT_FILTER_MITCHELL: {
      const double B = p[0], C = p[1];
      auto m1 = [&](double v) {
        v = std::fabs(2.0 * v / support);
        if (v > 1.0)
          return (((-B / 6.0 - C) * v + (B + 5.0 * C)) * v + (-2.0 * B - 8.0 * C)) * v + (4.0 / 3.0 * B + 4.0 * C);
        return v * v * ((2.0 - 1.5 * B - C) * v + (-3.0 + 2.0 * B + C)) + (1.0 - B / 3.0);
      };
      return m1(x) * m1(y);
    }
    case MPT_FILTER_LANCZOS: {
      const double tau = p[0];
      auto l1 = [&](double v) {
        v = std::fabs(v / support);
        if (v < 1e-7) return 1.0;
        if (v > 1) return 0.0;
        v *= M_PI;
        return std::sin(v * tau) / (v * tau) * (std::sin(v) / v);
      };
      return l1(x) * l1(y);
    }
    default:
      return 0.0;
  }
}

// The kernel KernelFilter::_computeKernel (filters/KernelFilter.cpp:162-176)
// tabulates: m = ceil(support) | 1 taps per side, weight [i][j] at the offsets
// (i - m/2, j - m/2) with m/2 taken as a real (so the taps sit half a texel
// off centre), row-major; row i pairs with image row y + i - m/2 (integer
// half) in the lookup.  Appends the sum of the weights in lookup order.
// Returns false with a message on a bad filter description.
inline bool kernel(const mpt_env_map &e, std::vector<double> &k, uint32_t &m, std::string &err) {
  const uint32_t support = e.support ? e.support : 2;
  m = ((uint32_t)std::ceil((double)support)) | 1u;
  if (m > kMaxKernelWidth) {
    err = "environment filter support too large (kernel at most 15 x 15)";
    return false;
  }
  double p[2] = {e.params[0], e.params[1]};
  switch (e.filter) {
    case MPT_FILTER_BOX:
    case MPT_FILTER_TRIANGLE:
      break;
    case MPT_FILTER_GAUSSIAN:
      if (p[0] == 0) p[0] = 1.0;
      if (!(p[0] > 0)) {
        err = "gaussian filter sigma must be positive";
        return false;
      }
      break;
    case MPT_FILTER_MITCHELL:
      if (p[0] == 0 && p[1] == 0) p[0] = p[1] = 1.0 / 3.0;
      break;
    case MPT_FILTER_LANCZOS:
      if (p[0] == 0) p[0] = 3.0;
      break;
    default:
      err = "unknown environment filter";
      return false;
  }
  const double half = m / 2.0;
  k.assign((size_t)m * m + 1, 0.0);
  for (uint32_t i = 0; i < m; ++i)
    for (uint32_t j = 0; j < m; ++j) k[(size_t)i * m + j] = filter_weight(e.filter, support, p, i - half, j - half);
  double sum = 0;
  for (size_t t = 0; t < (size_t)m * m; ++t) sum += k[t];
  if (!(sum != 0) || !std::isfinite(sum)) {
    err = "environment filter weights sum to zero";
    return false;
  }
  k[(size_t)m * m] = sum;
  return true;
}

}  // namespace mpt_env
