// Environment-image support shared by the scene upload and the scene-file
// loader: the reconstruction kernel an EnvironmentMap lookup applies, built
// on the host in double precision.
#pragma once

#include <cmath>
#include <string>
#include <vector>

#include "../../include/milton_pt.h"
#include "filters.h"

namespace mpt_env {

constexpr uint32_t kMaxKernelWidth = 15;

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
  switch (e.filter) {
    case MPT_FILTER_BOX:
    case MPT_FILTER_TRIANGLE:
    case MPT_FILTER_MITCHELL:
    case MPT_FILTER_LANCZOS:
      break;
    case MPT_FILTER_GAUSSIAN:
      if (e.params[0] < 0) {
        err = "gaussian filter sigma must be positive";
        return false;
      }
      break;
    default:
      err = "unknown environment filter";
      return false;
  }
  const mpt_filter::Params fp = mpt_filter::resolve(e.filter, support, e.params[0], e.params[1]);
  const double half = m / 2.0;
  k.assign((size_t)m * m + 1, 0.0);
  for (uint32_t i = 0; i < m; ++i)
    for (uint32_t j = 0; j < m; ++j) k[(size_t)i * m + j] = mpt_filter::weight(fp, i - half, j - half);
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
