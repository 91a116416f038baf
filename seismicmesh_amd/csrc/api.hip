// Stateless C entry points: errors, device query, SDF and sizing evaluation,
// 3D quality kernels and gradient limiting.
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "common.h"
#include "sdf_eval.h"
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

// fast_geometry.cpp:351-418 edge table: for angle i the edge (e[i][0],
// e[i][1]) and its opposite edge (e[5-i][0], e[5-i][1])
__constant__ int kDhEdges[6][2] = {{2, 3}, {1, 3}, {1, 2}, {0, 3}, {0, 2}, {0, 1}};

__device__ inline double norm3(const double *u) {
  return sqrt((u[0] * u[0] + u[1] * u[1]) + u[2] * u[2]);
}

__global__ void k_dihedral(const double *p, const int32_t *cells, int64_t nc, double *out) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  int v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) v[k] = cells[4 * c + k];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    const double *p0 = p + 3 * (int64_t)v[kDhEdges[i][0]];
    const double *p1 = p + 3 * (int64_t)v[kDhEdges[i][1]];
    const double *p2 = p + 3 * (int64_t)v[kDhEdges[5 - i][0]];
    const double *p3 = p + 3 * (int64_t)v[kDhEdges[5 - i][1]];
    double a[3], b[3], e[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      a[k] = p1[k] - p0[k];
      b[k] = p2[k] - p0[k];
      e[k] = p3[k] - p0[k];
    }
    double na = norm3(a), nb = norm3(b), ne = norm3(e);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      a[k] /= na;
      b[k] /= nb;
      e[k] /= ne;
    }
    // angle between the planes (a, b) and (a, e) along the edge a
    double be = (b[0] * e[0] + b[1] * e[1]) + b[2] * e[2];
    double ab = (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
    double ae = (a[0] * e[0] + a[1] * e[1]) + a[2] * e[2];
    double cab[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    double cae[3] = {a[1] * e[2] - a[2] * e[1], a[2] * e[0] - a[0] * e[2], a[0] * e[1] - a[1] * e[0]};
    out[6 * c + i] = acos((be - ab * ae) / (norm3(cab) * norm3(cae)));
  }
}

// Gradient of the squared circumradius of (q1, q2, q3, q4) with respect to
// q1.  With p_i = q_i - q4, u = p2 x p3, a = p1 . u (6 x volume) and
// w = |p2|^2 p3 - |p3|^2 p2, the circumcentre is -D / (2a) with
// D = -|p1|^2 u + p1 x w, so R^2 = |D|^2 / (4 a^2) and
//   dR^2/dp1 = (w x D - 2 (u . D) p1) / (2 a^2) - |D|^2 u / (2 a^3).
__global__ void k_circumgrad(const double *q1, const double *q2, const double *q3,
                             const double *q4, int64_t k, double *g) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k) return;
  double p1[3], p2[3], p3[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    p1[c] = q1[3 * i + c] - q4[3 * i + c];
    p2[c] = q2[3 * i + c] - q4[3 * i + c];
    p3[c] = q3[3 * i + c] - q4[3 * i + c];
  }
  double s1 = p1[0] * p1[0] + p1[1] * p1[1] + p1[2] * p1[2];
  double s2 = p2[0] * p2[0] + p2[1] * p2[1] + p2[2] * p2[2];
  double s3 = p3[0] * p3[0] + p3[1] * p3[1] + p3[2] * p3[2];
  double u[3] = {p2[1] * p3[2] - p2[2] * p3[1], p2[2] * p3[0] - p2[0] * p3[2],
                 p2[0] * p3[1] - p2[1] * p3[0]};
  double w[3] = {s2 * p3[0] - s3 * p2[0], s2 * p3[1] - s3 * p2[1], s2 * p3[2] - s3 * p2[2]};
  double a = p1[0] * u[0] + p1[1] * u[1] + p1[2] * u[2];
  double D[3] = {-s1 * u[0] + (p1[1] * w[2] - p1[2] * w[1]),
                 -s1 * u[1] + (p1[2] * w[0] - p1[0] * w[2]),
                 -s1 * u[2] + (p1[0] * w[1] - p1[1] * w[0])};
  double uD = u[0] * D[0] + u[1] * D[1] + u[2] * D[2];
  double DD = D[0] * D[0] + D[1] * D[1] + D[2] * D[2];
  double wxD[3] = {w[1] * D[2] - w[2] * D[1], w[2] * D[0] - w[0] * D[2], w[0] * D[1] - w[1] * D[0]};
  double a2 = 2.0 * a * a, a3 = 2.0 * a * a * a;
#pragma unroll
  for (int c = 0; c < 3; ++c) g[3 * i + c] = (wxD[c] - 2.0 * uD * p1[c]) / a2 - DD * u[c] / a3;
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
