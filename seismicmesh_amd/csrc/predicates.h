// Filtered exact geometric predicates for host and device (gfx950).
//
// Method: Shewchuk, "Adaptive Precision Floating-Point Arithmetic and Fast
// Robust Geometric Predicates" (Discrete Comput. Geom. 18, 1997).  Each
// predicate first evaluates its determinant in double precision and accepts
// the sign when it exceeds the paper's forward error bound (stage A); else it
// re-evaluates the determinant exactly with floating-point expansions
// (sums of non-overlapping doubles).  Products use FMA for the exact product
// tail.  This file must be compiled without floating-point contraction
// (-ffp-contract=off) so that the double-precision stage matches the error
// analysis.
//
// The exact stage keeps expansions in fixed-size arrays.  A predicate whose
// exact evaluation would not fit returns SM_PRED_UNRESOLVED; callers treat
// that as "cannot decide here" and fall back to the host rebuild.  Before
// it, the 2D predicates have a register-only exact stage in 128/256-bit
// integer arithmetic that settles nearby points of similar magnitude (the
// fixed-point stage below).
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define SM_HD __host__ __device__
#else
#define SM_HD
#endif

#define SM_PRED_UNRESOLVED 0x7fffffff

namespace smpred {

// 2^-53: half an ulp of 1.0.
constexpr double kEps = 1.1102230246251565e-16;
constexpr double kCcwBoundA = (3.0 + 16.0 * kEps) * kEps;
constexpr double kIccBoundA = (10.0 + 96.0 * kEps) * kEps;
constexpr double kO3dBoundA = (7.0 + 56.0 * kEps) * kEps;
constexpr double kIspBoundA = (16.0 + 224.0 * kEps) * kEps;

// ---------------------------------------------------------------- EFTs
SM_HD inline void two_sum(double a, double b, double &x, double &y) {
  x = a + b;
  double bv = x - a;
  double av = x - bv;
  y = (a - av) + (b - bv);
}
SM_HD inline void two_diff(double a, double b, double &x, double &y) {
  x = a - b;
  double bv = a - x;
  double av = x + bv;
  y = (a - av) + (bv - b);
}
SM_HD inline void two_prod(double a, double b, double &x, double &y) {
  x = a * b;
  y = fma(a, b, -x);
}

// An expansion: components in increasing magnitude, non-overlapping,
// zero-free.  N is the capacity.
template <int N> struct Exp {
  double c[N];
  int n = 0;
  bool ovf = false;
};

// h = e + b  (Grow-Expansion with zero elimination)
template <int N, int M>
SM_HD inline void grow(const Exp<M> &e, double b, Exp<N> &h) {
  h.n = 0;
  h.ovf = e.ovf;
  double q = b;
  for (int i = 0; i < e.n; ++i) {
    double s, t;
    two_sum(q, e.c[i], s, t);
    q = s;
    if (t != 0.0) {
      if (h.n >= N) { h.ovf = true; return; }
      h.c[h.n++] = t;
    }
  }
  if (q != 0.0 || h.n == 0) {
    if (h.n >= N) { h.ovf = true; return; }
    h.c[h.n++] = q;
  }
}

// h = e + f in time linear in the lengths: the components of both are
// merged by increasing magnitude and accumulated with exact sums, dropping
// zeros (Shewchuk's Fast-Expansion-Sum with zero elimination, "Adaptive
// Precision Floating-Point Arithmetic and Fast Robust Geometric
// Predicates", 1997, Theorem 13; valid in round-to-nearest-even).
template <int N, int M, int K>
SM_HD inline void add(const Exp<M> &e, const Exp<K> &f, Exp<N> &h) {
  h.n = 0;
  h.ovf = e.ovf || f.ovf;
  if (h.ovf) return;
  const int en = e.n, fn = f.n;
  if (en + fn == 0) return;
  int i = 0, j = 0;
  // next component by magnitude (ties take e's)
  auto take = [&](double &x) {
    if (j >= fn || (i < en && (fabs(e.c[i]) <= fabs(f.c[j])))) x = e.c[i++];
    else x = f.c[j++];
  };
  double q;
  take(q);
  while (i < en || j < fn) {
    double x, s, t;
    take(x);
    two_sum(q, x, s, t);
    q = s;
    if (t != 0.0) {
      if (h.n >= N) { h.ovf = true; return; }
      h.c[h.n++] = t;
    }
  }
  if (q != 0.0 || h.n == 0) {
    if (h.n >= N) { h.ovf = true; return; }
    h.c[h.n++] = q;
  }
}

template <int N, int M>
SM_HD inline void negate(const Exp<M> &e, Exp<N> &h) {
  h.n = e.n;
  h.ovf = e.ovf;
  for (int i = 0; i < e.n; ++i) h.c[i] = -e.c[i];
}

// h = e * b  (Scale-Expansion with zero elimination)
template <int N, int M>
SM_HD inline void scale(const Exp<M> &e, double b, Exp<N> &h) {
  h.n = 0;
  h.ovf = e.ovf;
  if (e.n == 0 || b == 0.0) return;
  double q, hh;
  two_prod(e.c[0], b, q, hh);
  if (hh != 0.0) h.c[h.n++] = hh;
  for (int i = 1; i < e.n; ++i) {
    double p1, p0, s, t;
    two_prod(e.c[i], b, p1, p0);
    two_sum(q, p0, s, t);
    if (t != 0.0) {
      if (h.n >= N) { h.ovf = true; return; }
      h.c[h.n++] = t;
    }
    double s2, t2;
    two_sum(p1, s, s2, t2);  // p1 dominates s: fast path is exact too
    q = s2;
    if (t2 != 0.0) {
      if (h.n >= N) { h.ovf = true; return; }
      h.c[h.n++] = t2;
    }
  }
  if (q != 0.0 || h.n == 0) {
    if (h.n >= N) { h.ovf = true; return; }
    h.c[h.n++] = q;
  }
}

// h = e * f: the sum over f's components of e scaled by each, accumulated
// with the linear-time sum above in two alternating buffers.
template <int N, int M, int K>
SM_HD inline void mul(const Exp<M> &e, const Exp<K> &f, Exp<N> &h) {
  h.n = 0;
  h.ovf = e.ovf || f.ovf;
  if (h.ovf || e.n == 0 || f.n == 0) return;
  if (f.n == 1) {
    scale(e, f.c[0], h);
    return;
  }
  Exp<N> b0, b1, part;
  Exp<N> *cur = &b0, *nxt = &b1;
  scale(e, f.c[0], *cur);
  for (int j = 1; j < f.n; ++j) {
    scale(e, f.c[j], part);
    if (part.ovf || cur->ovf) { h.ovf = true; return; }
    Exp<N> *out = (j + 1 == f.n) ? &h : nxt;
    add(*cur, part, *out);
    if (out->ovf) { h.ovf = true; return; }
    nxt = cur;
    cur = out;
  }
}

template <int N> SM_HD inline int sign_of(const Exp<N> &e) {
  if (e.ovf) return SM_PRED_UNRESOLVED;
  if (e.n == 0) return 0;
  double top = e.c[e.n - 1];
  return (top > 0.0) - (top < 0.0);
}

template <int N> SM_HD inline void from_diff(double a, double b, Exp<N> &h) {
  double x, y;
  two_diff(a, b, x, y);
  h.n = 0;
  h.ovf = false;
  if (y != 0.0) h.c[h.n++] = y;
  if (x != 0.0) h.c[h.n++] = x;
}

// ---------------------------------------------------------------- orient2d
// Sign of det[[ax-cx, ay-cy], [bx-cx, by-cy]]; > 0 when a, b, c turn CCW.
SM_HD inline int orient2d_exact(const double *a, const double *b, const double *c) {
  Exp<2> acx, acy, bcx, bcy;
  from_diff(a[0], c[0], acx);
  from_diff(a[1], c[1], acy);
  from_diff(b[0], c[0], bcx);
  from_diff(b[1], c[1], bcy);
  Exp<8> l, r, nr;
  mul(acx, bcy, l);
  mul(acy, bcx, r);
  negate(r, nr);
  Exp<16> d;
  add(l, nr, d);
  return sign_of(d);
}

SM_HD inline int orient2d(const double *a, const double *b, const double *c) {
  double detleft = (a[0] - c[0]) * (b[1] - c[1]);
  double detright = (a[1] - c[1]) * (b[0] - c[0]);
  double det = detleft - detright;
  double detsum;
  if (detleft > 0.0) {
    if (detright <= 0.0) return (det > 0.0) - (det < 0.0);
    detsum = detleft + detright;
  } else if (detleft < 0.0) {
    if (detright >= 0.0) return (det > 0.0) - (det < 0.0);
    detsum = -detleft - detright;
  } else {
    // detleft is an exact zero: det = -detright carries the exact sign
    return (det > 0.0) - (det < 0.0);
  }
  double err = kCcwBoundA * detsum;
  if (det > err) return 1;
  if (-det > err) return -1;
  return orient2d_exact(a, b, c);
}

// ---------------------------------------------------------------- incircle
// > 0 when d lies strictly inside the circle through the CCW triangle abc.
template <int N = 96>
SM_HD inline int incircle_exact(const double *a, const double *b, const double *c,
                                const double *d) {
  Exp<2> adx, ady, bdx, bdy, cdx, cdy;
  from_diff(a[0], d[0], adx);
  from_diff(a[1], d[1], ady);
  from_diff(b[0], d[0], bdx);
  from_diff(b[1], d[1], bdy);
  from_diff(c[0], d[0], cdx);
  from_diff(c[1], d[1], cdy);
  Exp<8> t1, t2, nt;
  Exp<16> bc, ca, ab, al, bl, cl;
  mul(bdx, cdy, t1); mul(cdx, bdy, t2); negate(t2, nt); add(t1, nt, bc);
  mul(cdx, ady, t1); mul(adx, cdy, t2); negate(t2, nt); add(t1, nt, ca);
  mul(adx, bdy, t1); mul(bdx, ady, t2); negate(t2, nt); add(t1, nt, ab);
  mul(adx, adx, t1); mul(ady, ady, t2); add(t1, t2, al);
  mul(bdx, bdx, t1); mul(bdy, bdy, t2); add(t1, t2, bl);
  mul(cdx, cdx, t1); mul(cdy, cdy, t2); add(t1, t2, cl);
  Exp<N> p1, p2, p3, s12, det;
  mul(al, bc, p1);
  mul(bl, ca, p2);
  mul(cl, ab, p3);
  add(p1, p2, s12);
  add(s12, p3, det);
  return sign_of(det);
}

SM_HD inline int incircle(const double *a, const double *b, const double *c,
                          const double *d) {
  double adx = a[0] - d[0], bdx = b[0] - d[0], cdx = c[0] - d[0];
  double ady = a[1] - d[1], bdy = b[1] - d[1], cdy = c[1] - d[1];
  double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
  double alift = adx * adx + ady * ady;
  double cdxady = cdx * ady, adxcdy = adx * cdy;
  double blift = bdx * bdx + bdy * bdy;
  double adxbdy = adx * bdy, bdxady = bdx * ady;
  double clift = cdx * cdx + cdy * cdy;
  double det = alift * (bdxcdy - cdxbdy) + blift * (cdxady - adxcdy) +
               clift * (adxbdy - bdxady);
  double perm = (fabs(bdxcdy) + fabs(cdxbdy)) * alift +
                (fabs(cdxady) + fabs(adxcdy)) * blift +
                (fabs(adxbdy) + fabs(bdxady)) * clift;
  double err = kIccBoundA * perm;
  if (det > err) return 1;
  if (-det > err) return -1;
  return incircle_exact(a, b, c, d);
}

// Double-precision stage only: returns SM_PRED_UNRESOLVED when the sign is
// not certain.  Device kernels use these and hand uncertain cases to a
// separate exact pass.
SM_HD inline int orient2d_filter(const double *a, const double *b, const double *c) {
  double detleft = (a[0] - c[0]) * (b[1] - c[1]);
  double detright = (a[1] - c[1]) * (b[0] - c[0]);
  double det = detleft - detright;
  double detsum;
  if (detleft > 0.0) {
    if (detright <= 0.0) return (det > 0.0) - (det < 0.0);
    detsum = detleft + detright;
  } else if (detleft < 0.0) {
    if (detright >= 0.0) return (det > 0.0) - (det < 0.0);
    detsum = -detleft - detright;
  } else {
    return (det > 0.0) - (det < 0.0);
  }
  double err = kCcwBoundA * detsum;
  if (det > err) return 1;
  if (-det > err) return -1;
  return SM_PRED_UNRESOLVED;
}

SM_HD inline int incircle_filter(const double *a, const double *b, const double *c,
                                 const double *d) {
  double adx = a[0] - d[0], bdx = b[0] - d[0], cdx = c[0] - d[0];
  double ady = a[1] - d[1], bdy = b[1] - d[1], cdy = c[1] - d[1];
  double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
  double alift = adx * adx + ady * ady;
  double cdxady = cdx * ady, adxcdy = adx * cdy;
  double blift = bdx * bdx + bdy * bdy;
  double adxbdy = adx * bdy, bdxady = bdx * ady;
  double clift = cdx * cdx + cdy * cdy;
  double det = alift * (bdxcdy - cdxbdy) + blift * (cdxady - adxcdy) +
               clift * (adxbdy - bdxady);
  double perm = (fabs(bdxcdy) + fabs(cdxbdy)) * alift +
                (fabs(cdxady) + fabs(adxcdy)) * blift +
                (fabs(adxbdy) + fabs(bdxady)) * clift;
  double err = kIccBoundA * perm;
  if (det > err) return 1;
  if (-det > err) return -1;
  return SM_PRED_UNRESOLVED;
}

// ---------------------------------------------------------------- orient3d
// Sign of the volume det[b-a; c-a; d-a] (> 0: a positively oriented
// tetrahedron, the convention of simp_vol), computed as -det[a-d; b-d; c-d].
template <int N = 96>
SM_HD inline int orient3d_exact(const double *a, const double *b, const double *c,
                                const double *d) {
  Exp<2> ad[3], bd[3], cd[3];
  for (int k = 0; k < 3; ++k) {
    from_diff(a[k], d[k], ad[k]);
    from_diff(b[k], d[k], bd[k]);
    from_diff(c[k], d[k], cd[k]);
  }
  // det = ad . (bd x cd)
  Exp<8> t1, t2, nt;
  Exp<16> m0, m1, m2;
  mul(bd[1], cd[2], t1); mul(bd[2], cd[1], t2); negate(t2, nt); add(t1, nt, m0);
  mul(bd[2], cd[0], t1); mul(bd[0], cd[2], t2); negate(t2, nt); add(t1, nt, m1);
  mul(bd[0], cd[1], t1); mul(bd[1], cd[0], t2); negate(t2, nt); add(t1, nt, m2);
  Exp<64> p0, p1, p2;
  mul(ad[0], m0, p0);
  mul(ad[1], m1, p1);
  mul(ad[2], m2, p2);
  Exp<N> s, det;
  add(p0, p1, s);
  add(s, p2, det);
  int sg = sign_of(det);
  return sg == SM_PRED_UNRESOLVED ? sg : -sg;
}

SM_HD inline double orient3d_det(const double *a, const double *b, const double *c,
                                 const double *d, double &perm) {
  double adx = a[0] - d[0], bdx = b[0] - d[0], cdx = c[0] - d[0];
  double ady = a[1] - d[1], bdy = b[1] - d[1], cdy = c[1] - d[1];
  double adz = a[2] - d[2], bdz = b[2] - d[2], cdz = c[2] - d[2];
  double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
  double cdxady = cdx * ady, adxcdy = adx * cdy;
  double adxbdy = adx * bdy, bdxady = bdx * ady;
  double det = adz * (bdxcdy - cdxbdy) + bdz * (cdxady - adxcdy) + cdz * (adxbdy - bdxady);
  perm = (fabs(bdxcdy) + fabs(cdxbdy)) * fabs(adz) +
         (fabs(cdxady) + fabs(adxcdy)) * fabs(bdz) +
         (fabs(adxbdy) + fabs(bdxady)) * fabs(cdz);
  return det;
}

SM_HD inline int orient3d(const double *a, const double *b, const double *c,
                          const double *d) {
  double perm;
  double det = orient3d_det(a, b, c, d, perm);
  double err = kO3dBoundA * perm;
  if (det > err) return -1;
  if (-det > err) return 1;
  return orient3d_exact(a, b, c, d);
}

// ---------------------------------------------------------------- insphere
// > 0 when e lies strictly inside the sphere through a, b, c, d where
// orient3d(a, b, c, d) > 0.
template <int N = 160>
SM_HD inline int insphere_exact(const double *a, const double *b, const double *c,
                                const double *d, const double *e) {
  const double *P[4] = {a, b, c, d};
  Exp<2> v[4][3];
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 3; ++k) from_diff(P[i][k], e[k], v[i][k]);
  Exp<24> lift[4];
  for (int i = 0; i < 4; ++i) {
    Exp<8> sx, sy, sz;
    Exp<16> sxy;
    mul(v[i][0], v[i][0], sx);
    mul(v[i][1], v[i][1], sy);
    mul(v[i][2], v[i][2], sz);
    add(sx, sy, sxy);
    add(sxy, sz, lift[i]);
  }
  // 2x2 minors of the xy columns: m[i][j] = x_i*y_j - x_j*y_i
  Exp<16> m[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = i + 1; j < 4; ++j) {
      Exp<8> t1, t2, nt;
      mul(v[i][0], v[j][1], t1);
      mul(v[j][0], v[i][1], t2);
      negate(t2, nt);
      add(t1, nt, m[i][j]);
    }
  // 3x3 minors with z: c3(i,j,k) = z_i*m[j][k] - z_j*m[i][k] + z_k*m[i][j]
  auto c3 = [&](int i, int j, int k, Exp<N> &out) {
    Exp<64> a1, a2, a3, na2;
    Exp<N> s;
    mul(v[i][2], m[j][k], a1);
    mul(v[j][2], m[i][k], a2);
    mul(v[k][2], m[i][j], a3);
    negate(a2, na2);
    add(a1, na2, s);
    add(s, a3, out);
  };
  Exp<N> c012, c013, c023, c123;
  c3(0, 1, 2, c012);
  c3(0, 1, 3, c013);
  c3(0, 2, 3, c023);
  c3(1, 2, 3, c123);
  // det4 = -l0*c123 + l1*c023 - l2*c013 + l3*c012
  Exp<N> q0, q1, q2, q3, nq, acc, acc2;
  mul(lift[0], c123, q0);
  mul(lift[1], c023, q1);
  mul(lift[2], c013, q2);
  mul(lift[3], c012, q3);
  negate(q0, nq);
  add(nq, q1, acc);
  negate(q2, nq);
  add(acc, nq, acc2);
  add(acc2, q3, acc);
  // sign convention: positive when e is inside for a positively oriented tet
  Exp<N> res;
  negate(acc, res);
  return sign_of(res);
}

// Double-precision stage: the 4x4 lifted determinant expanded along the
// lift column, with a forward error bound of 24 eps times its permanent (the
// evaluation tree is at most 10 roundings deep).
SM_HD inline int insphere_filter(const double *a, const double *b, const double *c,
                                 const double *d, const double *e) {
  const double *P[4] = {a, b, c, d};
  double x[4], y[4], z[4], l[4];
  for (int i = 0; i < 4; ++i) {
    x[i] = P[i][0] - e[0];
    y[i] = P[i][1] - e[1];
    z[i] = P[i][2] - e[2];
    l[i] = x[i] * x[i] + y[i] * y[i] + z[i] * z[i];
  }
  double m[4][4], mp[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = i + 1; j < 4; ++j) {
      double u = x[i] * y[j], v = x[j] * y[i];
      m[i][j] = u - v;
      mp[i][j] = fabs(u) + fabs(v);
    }
  auto c3 = [&](int i, int j, int k, double &perm) {
    perm = fabs(z[i]) * mp[j][k] + fabs(z[j]) * mp[i][k] + fabs(z[k]) * mp[i][j];
    return z[i] * m[j][k] - z[j] * m[i][k] + z[k] * m[i][j];
  };
  double p0, p1, p2, p3;
  double c123 = c3(1, 2, 3, p0), c023 = c3(0, 2, 3, p1);
  double c013 = c3(0, 1, 3, p2), c012 = c3(0, 1, 2, p3);
  double det = (-l[0] * c123 + l[1] * c023) + (-l[2] * c013 + l[3] * c012);
  double perm = (l[0] * p0 + l[1] * p1) + (l[2] * p2 + l[3] * p3);
  double err = 24.0 * kEps * perm;
  // this expansion of the lifted determinant is negative when e lies inside
  // the sphere of a tetrahedron with orient3d(a,b,c,d) > 0
  if (-det > err) return 1;
  if (det > err) return -1;
  return SM_PRED_UNRESOLVED;
}

SM_HD inline int insphere(const double *a, const double *b, const double *c,
                          const double *d, const double *e) {
  int s = insphere_filter(a, b, c, d, e);
  if (s != SM_PRED_UNRESOLVED) return s;
  return insphere_exact(a, b, c, d, e);
}

SM_HD inline int orient3d_filter(const double *a, const double *b, const double *c,
                                 const double *d) {
  double perm;
  double det = orient3d_det(a, b, c, d, perm);
  double err = kO3dBoundA * perm;
  if (det > err) return -1;
  if (-det > err) return 1;
  return SM_PRED_UNRESOLVED;
}

// Second stage without expansions: the plain double evaluation is repeated
// with error-free transformations, and its sign is exact when every
// difference, product and partial sum in it was exact (all error terms
// zero) -- the usual case for degenerate configurations of points on a
// lattice of modest multiples of a power of two, where the filter cannot
// decide.  The last operation needs no check: rounding a sum of exact
// values keeps its sign.  Registers only; SM_PRED_UNRESOLVED otherwise.
struct ExactTrack {
  bool ok = true;
  SM_HD double d(double x, double y) {
    double r, e;
    two_diff(x, y, r, e);
    ok = ok && e == 0.0;
    return r;
  }
  SM_HD double p(double x, double y) {
    double r, e;
    two_prod(x, y, r, e);
    ok = ok && e == 0.0;
    return r;
  }
  SM_HD double s(double x, double y) {
    double r, e;
    two_sum(x, y, r, e);
    ok = ok && e == 0.0;
    return r;
  }
};

SM_HD inline int orient2d_if_exact(const double *a, const double *b, const double *c) {
  ExactTrack k;
  double acx = k.d(a[0], c[0]), acy = k.d(a[1], c[1]);
  double bcx = k.d(b[0], c[0]), bcy = k.d(b[1], c[1]);
  double l = k.p(acx, bcy), r = k.p(acy, bcx);
  if (!k.ok) return SM_PRED_UNRESOLVED;
  double det = l - r;
  return (det > 0.0) - (det < 0.0);
}

SM_HD inline int incircle_if_exact(const double *a, const double *b, const double *c,
                                   const double *d) {
  ExactTrack k;
  double adx = k.d(a[0], d[0]), ady = k.d(a[1], d[1]);
  double bdx = k.d(b[0], d[0]), bdy = k.d(b[1], d[1]);
  double cdx = k.d(c[0], d[0]), cdy = k.d(c[1], d[1]);
  double bc = k.d(k.p(bdx, cdy), k.p(cdx, bdy));
  double ca = k.d(k.p(cdx, ady), k.p(adx, cdy));
  double ab = k.d(k.p(adx, bdy), k.p(bdx, ady));
  double al = k.s(k.p(adx, adx), k.p(ady, ady));
  double bl = k.s(k.p(bdx, bdx), k.p(bdy, bdy));
  double cl = k.s(k.p(cdx, cdx), k.p(cdy, cdy));
  double t1 = k.p(al, bc), t2 = k.p(bl, ca), t3 = k.p(cl, ab);
  double s12 = k.s(t1, t2);
  if (!k.ok) return SM_PRED_UNRESOLVED;
  double det = s12 + t3;
  return (det > 0.0) - (det < 0.0);
}

// ------------------------------------------------------ fixed-point stage
// Exact signs in integer arithmetic, registers only (no expansion arrays).
// A finite double is m * 2^e with an odd integer m (|m| < 2^53) or zero.
// With E the smallest e over the nonzero coordinates involved, every
// coordinate difference is an integer multiple of 2^E; when each quotient
// is below 2^62 in magnitude (nearby points of similar magnitude: every
// small quad of a mesh that does not straddle a coordinate axis at a tiny
// offset) the determinants are integer polynomials in the quotients.
// orient2d's is below 2^125 in magnitude and incircle's below 2^252, so
// 128- and 256-bit two's-complement integers hold them exactly.  Other
// inputs return SM_PRED_UNRESOLVED and the expansion stage decides.
typedef __int128 fix_i128;
typedef unsigned __int128 fix_u128;

struct FixNum {
  uint64_t mag;  // odd, or 0 for a zero
  int e;
  bool neg;
};

SM_HD inline bool fix_split(double x, FixNum &f) {
  uint64_t u;
  __builtin_memcpy(&u, &x, sizeof u);
  const int be = (int)((u >> 52) & 0x7ff);
  uint64_t m = u & ((1ull << 52) - 1);
  if (be == 0x7ff) return false;  // inf or nan
  int e;
  if (be == 0) {
    if (m == 0) {
      f.mag = 0;
      f.e = 0;
      f.neg = false;
      return true;
    }
    e = -1074;
  } else {
    m |= 1ull << 52;
    e = be - 1075;
  }
  const int tz = __builtin_ctzll(m);
  f.mag = m >> tz;
  f.e = e + tz;
  f.neg = (u >> 63) != 0;
  return true;
}

// (a - b) / 2^E exactly, when that is below 2^62 in magnitude (E <= the
// exponent of every nonzero operand)
SM_HD inline bool fix_diff(const FixNum &a, const FixNum &b, int E, int64_t &d) {
  fix_i128 A = 0, B = 0;
  if (a.mag) {
    const int s = a.e - E;
    if (s > 72) return false;  // |m| 2^s < 2^125
    A = (fix_i128)((fix_u128)a.mag << s);
    if (a.neg) A = -A;
  }
  if (b.mag) {
    const int s = b.e - E;
    if (s > 72) return false;
    B = (fix_i128)((fix_u128)b.mag << s);
    if (b.neg) B = -B;
  }
  const fix_i128 D = A - B;
  const fix_i128 lim = (fix_i128)1 << 62;
  if (D >= lim || D <= -lim) return false;
  d = (int64_t)D;
  return true;
}

// the coordinates of np points (np * dim doubles) split; E = smallest
// exponent of a nonzero one (0 when all are zero)
template <int NP, int DIM>
SM_HD inline bool fix_split_all(const double *const *P, FixNum (&f)[NP][DIM], int &E) {
  E = 1 << 30;
  for (int i = 0; i < NP; ++i)
    for (int k = 0; k < DIM; ++k) {
      if (!fix_split(P[i][k], f[i][k])) return false;
      if (f[i][k].mag && f[i][k].e < E) E = f[i][k].e;
    }
  if (E == (1 << 30)) E = 0;
  return true;
}

struct Fix256 {
  uint64_t w[4];  // two's complement, least significant word first
};

// x * |y| for x, |y| < 2^126, negated when y < 0
SM_HD inline Fix256 fix_mul(fix_u128 x, fix_i128 y) {
  const bool neg = y < 0;
  const fix_u128 ym = neg ? (fix_u128)(-y) : (fix_u128)y;
  const uint64_t x0 = (uint64_t)x, x1 = (uint64_t)(x >> 64);
  const uint64_t y0 = (uint64_t)ym, y1 = (uint64_t)(ym >> 64);
  const fix_u128 p00 = (fix_u128)x0 * y0, p01 = (fix_u128)x0 * y1;
  const fix_u128 p10 = (fix_u128)x1 * y0, p11 = (fix_u128)x1 * y1;
  Fix256 r;
  r.w[0] = (uint64_t)p00;
  const fix_u128 mid = (p00 >> 64) + (uint64_t)p01 + (uint64_t)p10;
  r.w[1] = (uint64_t)mid;
  const fix_u128 hi = (mid >> 64) + (p01 >> 64) + (p10 >> 64) + (uint64_t)p11;
  r.w[2] = (uint64_t)hi;
  r.w[3] = (uint64_t)(hi >> 64) + (uint64_t)(p11 >> 64);
  if (neg) {
    uint64_t carry = 1;
    for (int i = 0; i < 4; ++i) {
      const uint64_t t = ~r.w[i] + carry;
      carry = (carry && t == 0) ? 1 : 0;
      r.w[i] = t;
    }
  }
  return r;
}

SM_HD inline void fix_add(Fix256 &a, const Fix256 &b) {
  uint64_t c = 0;
  for (int i = 0; i < 4; ++i) {
    uint64_t s = a.w[i] + c;
    uint64_t c1 = s < c;
    s += b.w[i];
    c1 += s < b.w[i];
    a.w[i] = s;
    c = c1;
  }
}

SM_HD inline int fix_sign(const Fix256 &a) {
  if (a.w[3] >> 63) return -1;
  return (a.w[0] | a.w[1] | a.w[2] | a.w[3]) != 0;
}

SM_HD inline int orient2d_fix(const double *a, const double *b, const double *c) {
  const double *P[3] = {a, b, c};
  FixNum f[3][2];
  int E;
  if (!fix_split_all<3, 2>(P, f, E)) return SM_PRED_UNRESOLVED;
  int64_t acx, acy, bcx, bcy;
  if (!fix_diff(f[0][0], f[2][0], E, acx) || !fix_diff(f[0][1], f[2][1], E, acy) ||
      !fix_diff(f[1][0], f[2][0], E, bcx) || !fix_diff(f[1][1], f[2][1], E, bcy))
    return SM_PRED_UNRESOLVED;
  const fix_i128 det = (fix_i128)acx * bcy - (fix_i128)acy * bcx;
  return (det > 0) - (det < 0);
}

SM_HD inline int incircle_fix(const double *a, const double *b, const double *c,
                              const double *d) {
  const double *P[4] = {a, b, c, d};
  FixNum f[4][2];
  int E;
  if (!fix_split_all<4, 2>(P, f, E)) return SM_PRED_UNRESOLVED;
  int64_t v[3][2];
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 2; ++k)
      if (!fix_diff(f[i][k], f[3][k], E, v[i][k])) return SM_PRED_UNRESOLVED;
  // 2x2 minors and lifts, each below 2^125 in magnitude
  const fix_i128 bc = (fix_i128)v[1][0] * v[2][1] - (fix_i128)v[2][0] * v[1][1];
  const fix_i128 ca = (fix_i128)v[2][0] * v[0][1] - (fix_i128)v[0][0] * v[2][1];
  const fix_i128 ab = (fix_i128)v[0][0] * v[1][1] - (fix_i128)v[1][0] * v[0][1];
  fix_u128 lift[3];
  for (int i = 0; i < 3; ++i)
    lift[i] = (fix_u128)((fix_i128)v[i][0] * v[i][0]) + (fix_u128)((fix_i128)v[i][1] * v[i][1]);
  Fix256 s = fix_mul(lift[0], bc);
  fix_add(s, fix_mul(lift[1], ca));
  fix_add(s, fix_mul(lift[2], ab));
  return fix_sign(s);
}

// 3D, same scheme: the coordinate differences of the points against the
// last one as 62-bit integers; orient3d's determinant is below 2^190 (a
// 256-bit sum of 64 x 128-bit products), insphere's below 2^318 (lifts
// below 2^126 times 3x3 minors below 2^190: 384-bit sums).
struct Fix384 {
  uint64_t w[6];  // two's complement, least significant word first
};

// v negated in place (two's complement over n words)
SM_HD inline void fix_negate(uint64_t *v, int n) {
  uint64_t carry = 1;
  for (int i = 0; i < n; ++i) {
    const uint64_t t = ~v[i] + carry;
    carry = (carry && t == 0) ? 1 : 0;
    v[i] = t;
  }
}

// x * y for x < 2^128 and a signed 256-bit y below 2^255 in magnitude
SM_HD inline Fix384 fix_mul384(fix_u128 x, const Fix256 &y) {
  uint64_t ym[4] = {y.w[0], y.w[1], y.w[2], y.w[3]};
  const bool neg = (y.w[3] >> 63) != 0;
  if (neg) fix_negate(ym, 4);
  const uint64_t xm[2] = {(uint64_t)x, (uint64_t)(x >> 64)};
  Fix384 r = {{0, 0, 0, 0, 0, 0}};
  for (int i = 0; i < 2; ++i) {
    uint64_t carry = 0;
    for (int j = 0; j < 4; ++j) {
      const fix_u128 t = (fix_u128)xm[i] * ym[j] + r.w[i + j] + carry;
      r.w[i + j] = (uint64_t)t;
      carry = (uint64_t)(t >> 64);
    }
    r.w[i + 4] += carry;
  }
  if (neg) fix_negate(r.w, 6);
  return r;
}

template <int NW, class T>
SM_HD inline void fix_addw(T &a, const T &b) {
  uint64_t c = 0;
  for (int i = 0; i < NW; ++i) {
    uint64_t s = a.w[i] + c;
    uint64_t c1 = s < c;
    s += b.w[i];
    c1 += s < b.w[i];
    a.w[i] = s;
    c = c1;
  }
}

template <int NW, class T>
SM_HD inline int fix_signw(const T &a) {
  if (a.w[NW - 1] >> 63) return -1;
  uint64_t o = 0;
  for (int i = 0; i < NW; ++i) o |= a.w[i];
  return o != 0;
}

// z * m for a 62-bit z and a 128-bit m, as a signed 256-bit integer
SM_HD inline Fix256 fix_mul_zm(int64_t z, fix_i128 m) {
  const fix_u128 zm = z < 0 ? (fix_u128)(-(fix_i128)z) : (fix_u128)z;
  return fix_mul(zm, z < 0 ? -m : m);
}

// the 3x3 determinant | x y z | of rows i, j, k (rows of 62-bit integers),
// expanded along z with the 2x2 minors in x, y
SM_HD inline Fix256 fix_det3(const int64_t (&v)[4][3], int i, int j, int k) {
  const fix_i128 mjk = (fix_i128)v[j][0] * v[k][1] - (fix_i128)v[k][0] * v[j][1];
  const fix_i128 mik = (fix_i128)v[i][0] * v[k][1] - (fix_i128)v[k][0] * v[i][1];
  const fix_i128 mij = (fix_i128)v[i][0] * v[j][1] - (fix_i128)v[j][0] * v[i][1];
  Fix256 s = fix_mul_zm(v[i][2], mjk);
  fix_addw<4>(s, fix_mul_zm(v[j][2], -mik));
  fix_addw<4>(s, fix_mul_zm(v[k][2], mij));
  return s;
}

// orient3d's sign convention (> 0: b - a, c - a, d - a positively
// oriented), i.e. the sign of -det[a - d; b - d; c - d]
SM_HD inline int orient3d_fix(const double *a, const double *b, const double *c,
                              const double *d) {
  const double *P[4] = {a, b, c, d};
  FixNum f[4][3];
  int E;
  if (!fix_split_all<4, 3>(P, f, E)) return SM_PRED_UNRESOLVED;
  int64_t v[4][3];
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k)
      if (!fix_diff(f[i][k], f[3][k], E, v[i][k])) return SM_PRED_UNRESOLVED;
  return -fix_signw<4>(fix_det3(v, 0, 1, 2));
}

// insphere's sign convention (> 0: e inside the sphere of a positively
// oriented abcd): the lifted determinant of a - e .. d - e expanded along
// the lift column, -(-l0 D123 + l1 D023 - l2 D013 + l3 D012)
SM_HD inline int insphere_fix(const double *a, const double *b, const double *c,
                              const double *d, const double *e) {
  const double *P[5] = {a, b, c, d, e};
  FixNum f[5][3];
  int E;
  if (!fix_split_all<5, 3>(P, f, E)) return SM_PRED_UNRESOLVED;
  int64_t v[4][3];
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 3; ++k)
      if (!fix_diff(f[i][k], f[4][k], E, v[i][k])) return SM_PRED_UNRESOLVED;
  fix_u128 l[4];
  for (int i = 0; i < 4; ++i)
    l[i] = (fix_u128)((fix_i128)v[i][0] * v[i][0]) + (fix_u128)((fix_i128)v[i][1] * v[i][1]) +
           (fix_u128)((fix_i128)v[i][2] * v[i][2]);
  Fix256 D[4] = {fix_det3(v, 1, 2, 3), fix_det3(v, 0, 2, 3), fix_det3(v, 0, 1, 3),
                 fix_det3(v, 0, 1, 2)};
  // det = -l0 D0 + l1 D1 - l2 D2 + l3 D3
  Fix384 s = fix_mul384(l[1], D[1]);
  fix_addw<6>(s, fix_mul384(l[3], D[3]));
  Fix256 n0 = D[0], n2 = D[2];
  fix_negate(n0.w, 4);
  fix_negate(n2.w, 4);
  fix_addw<6>(s, fix_mul384(l[0], n0));
  fix_addw<6>(s, fix_mul384(l[2], n2));
  return -fix_signw<6>(s);
}

// Host entry points: same filters, exact stage with room for any input a
// double-precision coordinate set can produce in practice.
inline int orient2d_h(const double *a, const double *b, const double *c) {
  return orient2d(a, b, c);
}
inline int incircle_h(const double *a, const double *b, const double *c, const double *d) {
  int s = incircle_filter(a, b, c, d);
  if (s != SM_PRED_UNRESOLVED) return s;
  s = incircle_if_exact(a, b, c, d);
  if (s != SM_PRED_UNRESOLVED) return s;
  s = incircle_fix(a, b, c, d);
  if (s != SM_PRED_UNRESOLVED) return s;
  s = incircle_exact<96>(a, b, c, d);
  if (s != SM_PRED_UNRESOLVED) return s;
  return incircle_exact<2048>(a, b, c, d);
}
inline int orient3d_h(const double *a, const double *b, const double *c, const double *d) {
  int s = orient3d_filter(a, b, c, d);
  if (s != SM_PRED_UNRESOLVED) return s;
  s = orient3d_fix(a, b, c, d);
  if (s != SM_PRED_UNRESOLVED) return s;
  s = orient3d_exact<96>(a, b, c, d);
  if (s != SM_PRED_UNRESOLVED) return s;
  return orient3d_exact<2048>(a, b, c, d);
}
inline int insphere_h(const double *a, const double *b, const double *c, const double *d,
                      const double *e) {
  int s = insphere_filter(a, b, c, d, e);
  if (s != SM_PRED_UNRESOLVED) return s;
  s = insphere_fix(a, b, c, d, e);
  if (s != SM_PRED_UNRESOLVED) return s;
  s = insphere_exact<160>(a, b, c, d, e);
  if (s != SM_PRED_UNRESOLVED) return s;
  return insphere_exact<4096>(a, b, c, d, e);
}

}  // namespace smpred
