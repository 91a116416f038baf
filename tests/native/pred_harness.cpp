// Test harness: exposes the exact and filtered predicate stages of
// seismicmesh_amd/csrc/predicates.h to the CPU test suite (ctypes).
#include "predicates.h"

extern "C" {
int h_orient2d_exact(const double *a, const double *b, const double *c) {
  return smpred::orient2d_exact(a, b, c);
}
int h_incircle_exact96(const double *a, const double *b, const double *c, const double *d) {
  return smpred::incircle_exact<96>(a, b, c, d);
}
int h_incircle_h(const double *a, const double *b, const double *c, const double *d) {
  return smpred::incircle_h(a, b, c, d);
}
int h_orient3d_h(const double *a, const double *b, const double *c, const double *d) {
  return smpred::orient3d_h(a, b, c, d);
}
int h_insphere_h(const double *a, const double *b, const double *c, const double *d,
                 const double *e) {
  return smpred::insphere_h(a, b, c, d, e);
}
int h_insphere_exact160(const double *a, const double *b, const double *c, const double *d,
                        const double *e) {
  return smpred::insphere_exact<160>(a, b, c, d, e);
}
}

extern "C" {
int h_orient2d_if_exact(const double *a, const double *b, const double *c) {
  return smpred::orient2d_if_exact(a, b, c);
}
int h_incircle_if_exact(const double *a, const double *b, const double *c, const double *d) {
  return smpred::incircle_if_exact(a, b, c, d);
}
}

extern "C" {
int h_orient2d_fix(const double *a, const double *b, const double *c) {
  return smpred::orient2d_fix(a, b, c);
}
int h_incircle_fix(const double *a, const double *b, const double *c, const double *d) {
  return smpred::incircle_fix(a, b, c, d);
}
int h_orient3d_fix(const double *a, const double *b, const double *c, const double *d) {
  return smpred::orient3d_fix(a, b, c, d);
}
int h_insphere_fix(const double *a, const double *b, const double *c, const double *d,
                   const double *e) {
  return smpred::insphere_fix(a, b, c, d, e);
}
}
