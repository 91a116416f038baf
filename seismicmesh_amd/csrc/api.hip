// Stateless C entry points: errors, device query, SDF and sizing evaluation,
// 3D quality kernels and gradient limiting.
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "common.h"
#include "sdf_eval.h"
#include "quality3d.h"
#include "sizing_eval.h"

static thread_local char g_err[512] = "";

int sm_set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

extern "C" int sm_version(void) { return 100; }

extern "C" int sm_last_error(char *buf, size_t len) {
  if (!buf || len == 0) return SM_ERR_ARG;
  std::strncpy(buf, g_err, len - 1);
  buf[len - 1] = 0;
  return SM_OK;
}

extern "C" int sm_device_count(int *count) {
  if (!count) return SM_ERR_ARG;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  int usable = 0;
  for (int d = 0; d < n; ++d) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, d) != hipSuccess) continue;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) == 0) ++usable;
  }
  *count = usable;
  return SM_OK;
}

namespace {

using namespace smk;

template <int DIM>
__global__ void k_sdf(const sm_node *tape, int ntape, const double *x, int64_t n, double *d) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double q[3];
#pragma unroll
  for (int k = 0; k < DIM; ++k) q[k] = x[i * DIM + k];
  d[i] = sdf_eval<DIM>(tape, ntape, q);
}

template <int DIM>
__global__ void k_sizing(GridView g, const double *x, int64_t n, float *h) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double q[3];
#pragma unroll
  for (int k = 0; k < DIM; ++k) q[k] = x[i * DIM + k];
  h[i] = (float)grid_eval<DIM>(g, q);
}

__global__ void k_dihedral(const double *p, const int32_t *cells, int64_t nc, double *out) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  int32_t v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) v[k] = cells[4 * c + k];
#pragma unroll
  for (int i = 0; i < 6; ++i) out[6 * c + i] = dihedral_angle(p, v, i);
}

__global__ void k_circumgrad(const double *q1, const double *q2, const double *q3,
                             const double *q4, int64_t k, double *g) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k) return;
  circumsphere_grad(q1 + 3 * i, q2 + 3 * i, q3 + 3 * i, q4 + 3 * i, g + 3 * i);
}

// Gradient limiting: the fixed point of FastHJ.cpp's relaxation
// f_i <= f_j + elen*dfdx over the 4/6-neighbour grid stencil is the
// distance transform min_j (f0_j + elen*dfdx * |i - j|_1), which separates
// into a forward and a backward sweep along each grid line of each axis.
__global__ void k_limgrad_axis(double *f, int n0, int n1, int n2, int axis, double step) {
  int64_t line = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t len, stride, base;
  if (axis == 0) {
    if (line >= (int64_t)n1 * n2) return;
    len = n0;
    stride = 1;
    base = line * n0;
  } else if (axis == 1) {
    if (line >= (int64_t)n0 * n2) return;
    int64_t i = line % n0, k = line / n0;
    len = n1;
    stride = n0;
    base = i + (int64_t)n0 * n1 * k;
  } else {
    if (line >= (int64_t)n0 * n1) return;
    len = n2;
    stride = (int64_t)n0 * n1;
    base = line;
  }
  double prev = f[base];
  for (int64_t j = 1; j < len; ++j) {
    double cur = f[base + j * stride];
    double lim = prev + step;
    if (cur > lim) cur = lim;
    f[base + j * stride] = cur;
    prev = cur;
  }
  prev = f[base + (len - 1) * stride];
  for (int64_t j = len - 2; j >= 0; --j) {
    double cur = f[base + j * stride];
    double lim = prev + step;
    if (cur > lim) cur = lim;
    f[base + j * stride] = cur;
    prev = cur;
  }
}

}  // namespace

extern "C" int sm_sdf_eval(const sm_node *tape, int ntape, int dim, const double *x, int64_t n,
                           double *d, void *stream) {
  SM_CHECK(dim == 2 || dim == 3, SM_ERR_ARG, "sdf_eval: dim must be 2 or 3");
  SM_CHECK(n >= 0 && ntape > 0, SM_ERR_ARG, "sdf_eval: bad size");
  if (n == 0) return SM_OK;
  SM_CHECK(tape && x && d, SM_ERR_ARG, "sdf_eval: null pointer");
  hipStream_t s = stream_of(stream);
  if (dim == 2)
    hipLaunchKernelGGL(k_sdf<2>, dim3(blocks_for(n, 256)), dim3(256), 0, s, tape, ntape, x, n, d);
  else
    hipLaunchKernelGGL(k_sdf<3>, dim3(blocks_for(n, 256)), dim3(256), 0, s, tape, ntape, x, n, d);
  SM_LAUNCH_CHECK();
  return SM_OK;
}

extern "C" int sm_sizing_eval(const sm_grid *grid, const double *x, int64_t n, float *h,
                              void *stream) {
  SM_CHECK(grid && (grid->dim == 2 || grid->dim == 3), SM_ERR_ARG, "sizing_eval: bad grid");
  SM_CHECK(n >= 0, SM_ERR_ARG, "sizing_eval: bad size");
  if (n == 0) return SM_OK;
  SM_CHECK(x && h, SM_ERR_ARG, "sizing_eval: null pointer");
  GridView g{};
  g.dim = grid->dim;
  g.values = grid->values;
  g.hconst = grid->hconst;
  for (int k = 0; k < 3; ++k) {
    g.n[k] = grid->n[k];
    g.axis[k] = grid->axis[k];
  }
  if (g.values)
    for (int k = 0; k < g.dim; ++k)
      SM_CHECK(g.n[k] >= 2 && g.axis[k], SM_ERR_ARG, "sizing_eval: axis %d invalid", k);
  hipStream_t s = stream_of(stream);
  if (g.dim == 2)
    hipLaunchKernelGGL(k_sizing<2>, dim3(blocks_for(n, 256)), dim3(256), 0, s, g, x, n, h);
  else
    hipLaunchKernelGGL(k_sizing<3>, dim3(blocks_for(n, 256)), dim3(256), 0, s, g, x, n, h);
  SM_LAUNCH_CHECK();
  return SM_OK;
}

extern "C" int sm_dihedral_angles(const double *p, const int32_t *cells, int64_t ncells,
                                  double *angles, void *stream) {
  SM_CHECK(ncells >= 0, SM_ERR_ARG, "dihedral_angles: bad size");
  if (ncells == 0) return SM_OK;
  SM_CHECK(p && cells && angles, SM_ERR_ARG, "dihedral_angles: null pointer");
  hipLaunchKernelGGL(k_dihedral, dim3(blocks_for(ncells, 256)), dim3(256), 0, stream_of(stream),
                     p, cells, ncells, angles);
  SM_LAUNCH_CHECK();
  return SM_OK;
}

extern "C" int sm_circumsphere_grad(const double *p0, const double *p1, const double *p2,
                                    const double *p3, int64_t k, double *grad, void *stream) {
  SM_CHECK(k >= 0, SM_ERR_ARG, "circumsphere_grad: bad size");
  if (k == 0) return SM_OK;
  SM_CHECK(p0 && p1 && p2 && p3 && grad, SM_ERR_ARG, "circumsphere_grad: null pointer");
  hipLaunchKernelGGL(k_circumgrad, dim3(blocks_for(k, 256)), dim3(256), 0, stream_of(stream),
                     p0, p1, p2, p3, k, grad);
  SM_LAUNCH_CHECK();
  return SM_OK;
}

extern "C" int sm_limgrad(double *f, int nz, int nx, int ny, double elen, double dfdx,
                          void *stream) {
  SM_CHECK(f && nz > 0 && nx > 0 && ny > 0, SM_ERR_ARG, "limgrad: bad argument");
  const double step = elen * dfdx;
  hipStream_t s = stream_of(stream);
  const int64_t lines[3] = {(int64_t)nx * ny, (int64_t)nz * ny, (int64_t)nz * nx};
  const int dims[3] = {nz, nx, ny};
  for (int axis = 0; axis < 3; ++axis) {
    if (dims[axis] < 2) continue;
    hipLaunchKernelGGL(k_limgrad_axis, dim3(blocks_for(lines[axis], 256)), dim3(256), 0, s, f,
                       nz, nx, ny, axis, step);
    SM_LAUNCH_CHECK();
  }
  return SM_OK;
}
