// Device evaluation of a lowered SDF program (FP64).
//
// Follows SeismicMesh/geometry/signed_distance_functions.py: primitives
// (:429-623, box SDFs in fast_geometry.cpp:104-185), transforms in the order
// translate, Rx, Ry, Rz, stretch (:89-101), sharp CSG as min/max reductions
// and smooth CSG as left folds (:300-426), Repeat via a floored modulo
// (:292-294).  Every arithmetic step is written in numpy's order so that,
// compiled without contraction, untransformed primitives and CSG reproduce
// the reference's float64 values bit for bit.
#pragma once

#include "common.h"

namespace smk {

constexpr int kSdfStack = 24;
constexpr int kSdfPoints = 4;

__device__ inline double np_mod(double a, double b) {
  double m = fmod(a, b);
  if (m != 0.0) {
    if ((b < 0.0) != (m < 0.0)) m += b;
  } else {
    m = copysign(0.0, b);
  }
  return m;
}

template <int DIM>
__device__ inline void sdf_transform(const sm_node &nd, const double *x, double *q) {
#pragma unroll
  for (int k = 0; k < DIM; ++k) q[k] = x[k];
  const int f = nd.flags;
  if (!f) return;
  if (f & SM_XF_TRANSLATE) {
#pragma unroll
    for (int k = 0; k < DIM; ++k) q[k] = q[k] - nd.xf[k];
  }
  if (DIM == 2) {
    if (f & SM_XF_ROT_X) {
      const double *R = nd.xf + 3;
      double a = R[0] * q[0] + R[1] * q[1];
      double b = R[2] * q[0] + R[3] * q[1];
      q[0] = a;
      q[1] = b;
    }
  } else {
    const int bits[3] = {SM_XF_ROT_X, SM_XF_ROT_Y, SM_XF_ROT_Z};
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      if (f & bits[r]) {
        const double *R = nd.xf + 3 + 9 * r;
        double a = R[0] * q[0] + R[1] * q[1] + R[2] * q[2];
        double b = R[3] * q[0] + R[4] * q[1] + R[5] * q[2];
        double c = R[6] * q[0] + R[7] * q[1] + R[8] * q[2];
        q[0] = a;
        q[1] = b;
        q[2] = c;
      }
    }
  }
  if (f & SM_XF_STRETCH) {
    const double *v = nd.xf + 30;
    const double alpha = nd.xf[33];
    double dot = 0.0;
#pragma unroll
    for (int k = 0; k < DIM; ++k) dot = dot + v[k] * q[k];
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      double vx = dot * v[k];
      q[k] = vx / alpha + (q[k] - vx);
    }
  }
}

template <int DIM>
__device__ inline double sdf_primitive(const sm_node &nd, const double *x) {
  double q[3];
  sdf_transform<DIM>(nd, x, q);
  const double *P = nd.prm;
  switch (nd.op) {
    case SM_OP_DISK: {
      double dx = q[0] - P[0], dy = q[1] - P[1];
      return sqrt(dx * dx + dy * dy) - P[2];
    }
    case SM_OP_RECT: {
      // fast_geometry.cpp:165-185 (std::min keeps the first on ties)
      double s1 = -P[2] + q[1], s2 = P[3] - q[1];
      double s3 = -P[0] + q[0], s4 = P[1] - q[0];
      double t1 = (s2 < s1) ? s2 : s1;
      double t2 = (s3 < t1) ? s3 : t1;
      return -((s4 < t2) ? s4 : t2);
    }
    case SM_OP_BALL: {
      double dx = q[0] - P[0], dy = q[1] - P[1], dz = q[2] - P[2];
      return sqrt(dx * dx + dy * dy + dz * dz) - P[3];
    }
    case SM_OP_CUBE: {
      // fast_geometry.cpp:104-134
      double s1 = -P[4] + q[2], s2 = P[5] - q[2];
      double s3 = -P[2] + q[1], s4 = P[3] - q[1];
      double s5 = -P[0] + q[0], s6 = P[1] - q[0];
      double t = (s2 < s1) ? s2 : s1;
      t = (s3 < t) ? s3 : t;
      t = (s4 < t) ? s4 : t;
      t = (s5 < t) ? s5 : t;
      return -((s6 < t) ? s6 : t);
    }
    case SM_OP_TORUS: {
      double lxz = sqrt(q[0] * q[0] + q[2] * q[2]);
      double a = lxz - P[0], b = q[1];
      return sqrt(a * a + b * b) - P[1];
    }
    case SM_OP_PRISM: {
      double qx = fabs(q[0]), qz = fabs(q[2]);
      double m = np_max(qx * 0.866025 + q[1] * 0.5, -q[1]) - P[0] * 0.5;
      return np_max(qz - P[1], m);
    }
    case SM_OP_CYLINDER: {
      double lxz = sqrt(q[0] * q[0] + q[2] * q[2]);
      double d0 = fabs(lxz) - P[0], d1 = fabs(q[1]) - P[1];
      double m0 = np_max(d0, 0.0), m1 = np_max(d1, 0.0);
      return np_min(np_max(d0, d1), 0.0) + sqrt(m0 * m0 + m1 * m1);
    }
    default:
      return 0.0;
  }
}

__device__ inline double smooth_union(double d1, double d2, double k) {
  double h = np_max(k - fabs(d1 - d2), 0.0);
  return np_min(d1, d2) - (h * h * 0.25) / k;
}
__device__ inline double smooth_intersection(double d1, double d2, double k) {
  double h = np_max(k - fabs(d1 - d2), 0.0);
  return np_max(d1, d2) + h * h * 0.25 / k;
}
__device__ inline double smooth_difference(double d1, double d2, double k) {
  double h = np_max(k - fabs(-d1 - d2), 0.0);
  return np_max(-d1, d2) + (h * h * 0.25) / k;
}

// Register-resident value stack.  Every lane of a wave walks the same
// program, so `sp` is wave-uniform and the switches below are uniform
// branches; no value ever goes to scratch memory.
constexpr int kRegStack = 8;
struct RegStack {
  double s[kRegStack];
  int sp = 0;
  __device__ inline double get(int i) const {
    switch (i) {
      case 0: return s[0];
      case 1: return s[1];
      case 2: return s[2];
      case 3: return s[3];
      case 4: return s[4];
      case 5: return s[5];
      case 6: return s[6];
      default: return s[7];
    }
  }
  __device__ inline void set(int i, double v) {
    switch (i) {
      case 0: s[0] = v; break;
      case 1: s[1] = v; break;
      case 2: s[2] = v; break;
      case 3: s[3] = v; break;
      case 4: s[4] = v; break;
      case 5: s[5] = v; break;
      case 6: s[6] = v; break;
      default: s[7] = v; break;
    }
  }
};

// Interpreter for programs of stack depth <= 8 and at most one pushed
// evaluation point (every program the Python lowering emits for flat or
// moderately nested CSG); sdf_eval_deep handles the rest.
template <int DIM>
__device__ inline double sdf_eval_reg(const sm_node *__restrict__ tape, int ntape,
                                      const double *x0) {
  RegStack st;
  double x[3], q[3];
  bool moved = false;
#pragma unroll
  for (int k = 0; k < DIM; ++k) x[k] = x0[k];
  for (int i = 0; i < ntape; ++i) {
    const sm_node &nd = tape[i];
    const int op = nd.op;
    if (op < SM_OP_UNION) {
      double v = sdf_primitive<DIM>(nd, moved ? q : x);
      if (st.sp < kRegStack) st.set(st.sp++, v);
    } else if (op <= SM_OP_DIFFERENCE) {
      const int n = nd.nargs;
      if (n < 1 || n > st.sp) continue;
      const double k = nd.prm[0];
      const int b = st.sp - n;
      double r;
      // one branch per operation (a chain of selects on the opcode here was
      // miscompiled for gfx950 into the wrong reduction)
      if (k == 0.0 && op == SM_OP_UNION) {
        r = st.get(b);
        for (int j = 1; j < n; ++j) r = np_min(r, st.get(b + j));
      } else if (k == 0.0 && op == SM_OP_INTERSECT) {
        r = st.get(b);
        for (int j = 1; j < n; ++j) r = np_max(r, st.get(b + j));
      } else if (k == 0.0) {
        r = st.get(b);
        for (int j = 1; j < n; ++j) r = np_max(r, -st.get(b + j));
      } else if (op == SM_OP_DIFFERENCE) {
        r = st.get(b + n - 1);
        for (int j = n - 2; j >= 0; --j) r = smooth_difference(r, st.get(b + j), k);
      } else if (op == SM_OP_UNION) {
        r = st.get(b);
        for (int j = 1; j < n; ++j) r = smooth_union(r, st.get(b + j), k);
      } else {
        r = st.get(b);
        for (int j = 1; j < n; ++j) r = smooth_intersection(r, st.get(b + j), k);
      }
      st.sp = b;
      st.set(st.sp++, r);
    } else if (op == SM_OP_SDIFF_STEP) {
      if (st.sp < 2) continue;
      double bv = st.get(--st.sp);
      double av = st.get(st.sp - 1);
      st.set(st.sp - 1, smooth_difference(av, bv, nd.prm[0]));
    } else if (op == SM_OP_PUSH_MOD) {
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        double p = nd.prm[c];
        q[c] = np_mod(x[c] + 0.5 * p, p) - 0.5 * p;
      }
      moved = true;
    } else if (op == SM_OP_POP_POINT) {
      moved = false;
    }
  }
  return st.sp > 0 ? st.get(st.sp - 1) : 0.0;
}

template <int DIM>
__device__ inline double sdf_eval_deep(const sm_node *__restrict__ tape, int ntape,
                                       const double *x0) {
  double st[kSdfStack];
  double pts[kSdfPoints][3];
  int sp = 0, pp = 0;
#pragma unroll
  for (int k = 0; k < DIM; ++k) pts[0][k] = x0[k];
  for (int i = 0; i < ntape; ++i) {
    const sm_node &nd = tape[i];
    const int op = nd.op;
    // malformed programs (stack over/underflow) are clamped, never indexed
    // out of bounds; the host validates programs before launch
    if (op < SM_OP_UNION) {
      double v = sdf_primitive<DIM>(nd, pts[pp]);
      if (sp < kSdfStack) st[sp++] = v;
    } else if (op <= SM_OP_DIFFERENCE) {
      // n-ary reduction in the reference's order (lowering emits n = 2
      // folds to keep the stack shallow)
      const int n = nd.nargs;
      if (n < 1 || n > sp) continue;
      const double k = nd.prm[0];
      double *a = st + sp - n;
      double r;
      if (k == 0.0) {
        if (op == SM_OP_UNION) {
          r = a[0];
          for (int j = 1; j < n; ++j) r = np_min(r, a[j]);
        } else if (op == SM_OP_INTERSECT) {
          r = a[0];
          for (int j = 1; j < n; ++j) r = np_max(r, a[j]);
        } else {
          r = a[0];
          for (int j = 1; j < n; ++j) r = np_max(r, -a[j]);
        }
      } else {
        if (op == SM_OP_UNION) {
          r = a[0];
          for (int j = 1; j < n; ++j) r = smooth_union(r, a[j], k);
        } else if (op == SM_OP_INTERSECT) {
          r = a[0];
          for (int j = 1; j < n; ++j) r = smooth_intersection(r, a[j], k);
        } else {
          // left fold over the reversed list (signed_distance_functions.py:423)
          r = a[n - 1];
          for (int j = n - 2; j >= 0; --j) r = smooth_difference(r, a[j], k);
        }
      }
      sp -= n;
      st[sp++] = r;
    } else if (op == SM_OP_SDIFF_STEP) {
      // one step of the reversed left fold of a smooth Difference
      if (sp < 2) continue;
      double b = st[--sp];
      double a = st[sp - 1];
      st[sp - 1] = smooth_difference(a, b, nd.prm[0]);
    } else if (op == SM_OP_PUSH_MOD) {
      if (pp + 1 >= kSdfPoints) continue;
#pragma unroll
      for (int c = 0; c < DIM; ++c) {
        double p = nd.prm[c];
        pts[pp + 1][c] = np_mod(pts[pp][c] + 0.5 * p, p) - 0.5 * p;
      }
      ++pp;
    } else if (op == SM_OP_POP_POINT) {
      if (pp > 0) --pp;
    }
  }
  return sp > 0 ? st[sp - 1] : 0.0;
}

// Programs are tagged by the host: nd[0].pad[0] = 1 marks a program that
// needs the deep (scratch-stack) interpreter.  DEEP selects the interpreter
// at compile time (hot kernels are instantiated both ways so the common
// variant has no scratch memory at all); DYNAMIC reads the tag.  A lone
// primitive is evaluated directly.
enum SdfMode { kSdfReg = 0, kSdfDeep = 1, kSdfDynamic = 2 };

template <int DIM, int MODE = kSdfDynamic>
__device__ inline double sdf_eval(const sm_node *__restrict__ tape, int ntape,
                                  const double *x0) {
  if (ntape == 1) return sdf_primitive<DIM>(tape[0], x0);
  if (MODE == kSdfDeep) return sdf_eval_deep<DIM>(tape, ntape, x0);
  if (MODE == kSdfDynamic && tape[0].pad[0] != 0.0) return sdf_eval_deep<DIM>(tape, ntape, x0);
  return sdf_eval_reg<DIM>(tape, ntape, x0);
}

}  // namespace smk
