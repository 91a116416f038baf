// Tetrahedron quality arithmetic shared by the stateless quality kernels and
// the sliver-removal step: the six dihedral angles and the gradient of the
// squared circumradius with respect to the first vertex.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace smk {

// fast_geometry.cpp:351-418 edge table: for angle i the edge (e[i][0],
// e[i][1]) and its opposite edge (e[5-i][0], e[5-i][1])
__device__ inline int dh_edge(int i, int j) {
  // {{2, 3}, {1, 3}, {1, 2}, {0, 3}, {0, 2}, {0, 1}}
  constexpr int tab[12] = {2, 3, 1, 3, 1, 2, 0, 3, 0, 2, 0, 1};
  return tab[2 * i + j];
}

__device__ inline double norm3(const double *u) {
  return sqrt((u[0] * u[0] + u[1] * u[1]) + u[2] * u[2]);
}

// Dihedral angle i (radians) of the tetrahedron with vertex ids v in p (n, 3).
__device__ inline double dihedral_angle(const double *p, const int32_t *v, int i) {
  const double *p0 = p + 3 * (int64_t)v[dh_edge(i, 0)];
  const double *p1 = p + 3 * (int64_t)v[dh_edge(i, 1)];
  const double *p2 = p + 3 * (int64_t)v[dh_edge(5 - i, 0)];
  const double *p3 = p + 3 * (int64_t)v[dh_edge(5 - i, 1)];
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
  return acos((be - ab * ae) / (norm3(cab) * norm3(cae)));
}

// Gradient of the squared circumradius of (q1, q2, q3, q4) with respect to
// q1.  With p_i = q_i - q4, u = p2 x p3, a = p1 . u (6 x volume) and
// w = |p2|^2 p3 - |p3|^2 p2, the circumcentre is -D / (2a) with
// D = -|p1|^2 u + p1 x w, so R^2 = |D|^2 / (4 a^2) and
//   dR^2/dp1 = (w x D - 2 (u . D) p1) / (2 a^2) - |D|^2 u / (2 a^3).
__device__ inline void circumsphere_grad(const double *q1, const double *q2, const double *q3,
                                         const double *q4, double *g) {
  double p1[3], p2[3], p3[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    p1[c] = q1[c] - q4[c];
    p2[c] = q2[c] - q4[c];
    p3[c] = q3[c] - q4[c];
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
  for (int c = 0; c < 3; ++c) g[c] = (wxD[c] - 2.0 * uD * p1[c]) / a2 - DD * u[c] / a3;
}

}  // namespace smk
