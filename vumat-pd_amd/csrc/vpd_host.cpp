// Host-side helpers of the case runner: the lattice, the boundary bands, the regional
// constraint codes and the smooth ramp. No GPU work; these set up what the kernels consume.
#include "vpd.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <string>
#include <vector>

namespace {

// ParticleSets.cpp:12-50. side: 0 left, 1 right, 2 bottom, 3 top.
bool selectBand(const double *x, int64_t n, int side, double width, std::vector<char> &mask) {
  if (width <= 0.0 || n == 0)
    return false;
  const int comp = side < 2 ? 0 : 1;
  const bool minimum = side == 0 || side == 2;
  double extreme = minimum ? std::numeric_limits<double>::infinity()
                           : -std::numeric_limits<double>::infinity();
  for (int64_t i = 0; i < n; ++i)
    extreme = minimum ? std::min(extreme, x[3 * i + comp]) : std::max(extreme, x[3 * i + comp]);
  const double tolerance = 1.0e-10 * std::max(1.0, std::abs(extreme));
  mask.assign(n, 0);
  bool any = false;
  for (int64_t i = 0; i < n; ++i) {
    const double d = minimum ? x[3 * i + comp] - extreme : extreme - x[3 * i + comp];
    if (d <= width + tolerance) {
      mask[i] = 1;
      any = true;
    }
  }
  return any;
}

void setCode(uint8_t &code, int axis, int value) {
  code = (uint8_t)((code & ~(3u << (2 * axis))) | ((unsigned)value << (2 * axis)));
}

} // namespace

extern "C" {

// selectBoundaryRegion (ParticleSets.cpp:12-50): side "left", "right", "bottom" or "top".
int64_t vpd_select_boundary_region(const double *x, int64_t n, const char *side, double width,
                                   uint8_t *mask) {
  if (!x || !side || !mask || n <= 0 || !(width > 0.0))
    return -VPD_ERR_ARGUMENT;
  const std::string s(side);
  const int code = s == "left" ? 0 : s == "right" ? 1 : s == "bottom" ? 2 : s == "top" ? 3 : -1;
  if (code < 0)
    return -VPD_ERR_ARGUMENT;
  std::vector<char> m;
  if (!selectBand(x, n, code, width, m))
    return -VPD_ERR_ARGUMENT;
  int64_t count = 0;
  for (int64_t i = 0; i < n; ++i) {
    mask[i] = m[i] ? 1 : 0;
    count += m[i] ? 1 : 0;
  }
  return count;
}

// Geometry.cpp:10-34
int vpd_make_patch(int nx, int ny, int nz, double h, double *x, double *volume) {
  if (nx < 1 || ny < 1 || nz < 1 || !(h > 0.0) || !x || !volume)
    return VPD_ERR_ARGUMENT;
  const double yOff = 0.5 * static_cast<double>(ny - 1) * h;
  const double zOff = 0.5 * static_cast<double>(nz - 1) * h;
  const double v = h * h * h;
  int64_t p = 0;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k, ++p) {
        x[3 * p] = static_cast<double>(i) * h;
        x[3 * p + 1] = static_cast<double>(j) * h - yOff;
        x[3 * p + 2] = static_cast<double>(k) * h - zOff;
        volume[p] = v * ((i == 0 || i == nx - 1) ? 0.5 : 1.0) *
                    ((j == 0 || j == ny - 1) ? 0.5 : 1.0) * ((k == 0 || k == nz - 1) ? 0.5 : 1.0);
      }
  return VPD_OK;
}

// BoundaryConditions.cpp:54-68 (sets, anchor) and :81-115 (the order of writes in apply):
// support band → load component 0; load band → ramp; anchor lateral → 0; out of plane,
// anchor z → 0 and the support point above the anchor with the greatest y → z 0; plane
// strain holds every z at 0.
int vpd_regional_constraint_codes(const double *x, int64_t n, int component, int layers,
                                  double spacing, int planeStrain, uint8_t *codes,
                                  int64_t *anchorOut, int64_t *pinOut) {
  if (!x || !codes || n <= 0 || (component != 0 && component != 1) || layers < 1)
    return VPD_ERR_ARGUMENT;
  const double width = static_cast<double>(layers - 1) * spacing + 1.0e-9 * spacing;
  std::vector<char> support, loaded;
  if (!selectBand(x, n, component == 0 ? 0 : 2, width, support) ||
      !selectBand(x, n, component == 0 ? 1 : 3, width, loaded))
    return VPD_ERR_ARGUMENT;
  double centre[3] = {0.0, 0.0, 0.0};
  for (int64_t i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d)
      centre[d] += x[3 * i + d];
  for (double &c : centre)
    c /= static_cast<double>(n);
  // ParticleSets.cpp:52-71: first strictly closer support particle in index order.
  int64_t anchor = -1;
  double best = std::numeric_limits<double>::infinity();
  for (int64_t i = 0; i < n; ++i) {
    if (!support[i])
      continue;
    if (anchor < 0)
      anchor = i;
    double dist = 0.0;
    for (int d = 0; d < 3; ++d) {
      const double delta = x[3 * i + d] - centre[d];
      dist += delta * delta;
    }
    if (dist < best) {
      best = dist;
      anchor = i;
    }
  }
  const int lateral = component == 0 ? 1 : 0;
  std::fill(codes, codes + n, (uint8_t)0);
  for (int64_t i = 0; i < n; ++i)
    if (support[i])
      setCode(codes[i], component, VPD_ZERO);
  for (int64_t i = 0; i < n; ++i)
    if (loaded[i])
      setCode(codes[i], component, VPD_RAMP);
  setCode(codes[anchor], lateral, VPD_ZERO);
  int64_t pin = anchor;
  if (!planeStrain) {
    setCode(codes[anchor], 2, VPD_ZERO);
    double greatestY = -1.0e100;
    for (int64_t i = 0; i < n; ++i) {
      if (!support[i])
        continue;
      if (std::abs(x[3 * i] - x[3 * anchor]) < 1e-8 &&
          std::abs(x[3 * i + 2] - x[3 * anchor + 2]) < 1e-8 && x[3 * i + 1] > greatestY) {
        greatestY = x[3 * i + 1];
        pin = i;
      }
    }
    setCode(codes[pin], 2, VPD_ZERO);
  } else {
    for (int64_t i = 0; i < n; ++i)
      setCode(codes[i], 2, VPD_ZERO);
  }
  if (anchorOut)
    *anchorOut = anchor;
  if (pinOut)
    *pinOut = pin;
  return VPD_OK;
}

// BoundaryConditions.cpp:70-79
void vpd_smooth_ramp(double amplitude, double riseTime, double time, double out[3]) {
  if (time >= riseTime) {
    out[0] = amplitude;
    out[1] = 0.0;
    out[2] = 0.0;
    return;
  }
  const double pi = 3.14159265358979323846;
  const double phase = pi * time / riseTime;
  out[0] = 0.5 * amplitude * (1.0 - std::cos(phase));
  out[1] = 0.5 * amplitude * pi / riseTime * std::sin(phase);
  out[2] = 0.5 * amplitude * pi * pi / (riseTime * riseTime) * std::cos(phase);
}

} // extern "C"
