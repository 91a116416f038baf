// Device linear interpolation on a regular sizing grid.
//
// Follows SeismicMesh/sizing/mesh_size_function.py:391-408 and :514-523:
// scipy's RegularGridInterpolator(method="linear", bounds_error=False,
// fill_value=None) over float32 linspace axes (z, x[, y]) -- the interval of
// a query is the one searchsorted(axis, x) - 1 selects, clipped to the grid,
// so queries outside the box are linearly extrapolated.  Index and weight
// arithmetic is FP64; grid values are FP32.  The corner sums are taken in
// scipy's order (2D: its written formula left to right; 3D: the corners of
// the hypercube in itertools.product order, each weight a left-to-right
// product), which reproduces its float64 results exactly.
#pragma once

#include "common.h"

namespace smk {

struct GridView {
  int dim;
  int n[3];
  const double *axis[3];
  const float *values;
  double hconst;
};

__device__ inline int grid_interval(const double *ax, int n, double x) {
  // largest i in [0, n-2] with ax[i] < x (searchsorted 'left' minus one),
  // starting from the uniform-spacing guess
  const double a0 = ax[0], a1 = ax[n - 1];
  double t = (x - a0) / (a1 - a0) * (double)(n - 1);
  int i;
  if (!(t > 0.0)) i = 0;
  else if (t >= (double)(n - 2)) i = n - 2;
  else i = (int)t;
  while (i > 0 && !(ax[i] < x)) --i;
  while (i < n - 2 && ax[i + 1] < x) ++i;
  return i;
}

template <int DIM>
__device__ inline double grid_eval(const GridView &g, const double *x) {
  if (g.values == nullptr) return g.hconst;
  int i[3];
  double w[3];
#pragma unroll
  for (int k = 0; k < DIM; ++k) {
    const double *ax = g.axis[k];
    i[k] = grid_interval(ax, g.n[k], x[k]);
    double lo = ax[i[k]], hi = ax[i[k] + 1];
    w[k] = (x[k] - lo) / (hi - lo);
  }
  if (DIM == 2) {
    const float *v = g.values + (int64_t)i[0] * g.n[1] + i[1];
    double v00 = v[0], v01 = v[1];
    double v10 = v[g.n[1]], v11 = v[g.n[1] + 1];
    const double a = 1.0 - w[0], b = w[0], c = 1.0 - w[1], d = w[1];
    // scipy's 2D linear formula, left to right (its results bit for bit)
    return v00 * a * c + v01 * a * d + v10 * b * c + v11 * b * d;
  } else {
    const int64_t s1 = g.n[2], s0 = (int64_t)g.n[1] * g.n[2];
    const float *v = g.values + (int64_t)i[0] * s0 + (int64_t)i[1] * s1 + i[2];
    double acc = 0.0;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      int b0 = (c >> 2) & 1, b1 = (c >> 1) & 1, b2 = c & 1;
      double wt = (b0 ? w[0] : 1.0 - w[0]) * (b1 ? w[1] : 1.0 - w[1]) * (b2 ? w[2] : 1.0 - w[2]);
      acc += (double)v[b0 * s0 + b1 * s1 + b2] * wt;
    }
    return acc;
  }
}

}  // namespace smk
