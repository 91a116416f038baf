// One DistMesh position update on the device.
//
// Follows SeismicMesh/generation/mesh_generator.py:
//   _remove_triangles_outside  :734-738  -> k_keep
//   _get_edges / unique_edges  :680-688  -> build_edges (edges.hip)
//   _compute_forces            :691-712  -> k_edge_geom, sums, k_edge_force,
//                                           k_vertex_update (scatter)
//   Ftot[ifix] = 0, p += dt*F  :499-502  -> k_vertex_update
//   _project_points_back_newton :762-784 -> k_vertex_update
//
// Layout in HBM: points are (N, dim) FP64 row-major; edges (E, 2) int32 in
// the reference's sorted order; per-edge L, h and F vectors in the compute
// precision T.  The scatter of generation/utils.py:48-68 (coo_matrix
// (...).toarray(), which adds duplicates in entry order) is a deterministic
// gather: vertex v sums -Fvec over the edges where it is the larger id, in
// increasing edge order, then +Fvec over the edges where it is the smaller
// id -- exactly the reference's order of additions.  With T = double and the
// L0 sums taken in numpy's summation order (pairwise blocks of 128 inside
// sequential chunks of 8192) an iteration reproduces the reference's float64
// arithmetic bit for bit; the one transcendental scalar, (sum L^d / sum h^d)
// ** (1/d), is taken on the host with the C library's pow as numpy does.
#include <cmath>
#include <vector>

#include "common.h"
#include "edges.h"
#include "mesher.h"
#include "sdf_eval.h"
#include "sizing_eval.h"

namespace smk {

namespace {

// A cell is real when all its vertex ids lie in [0, N): ghost cells carry
// -1 and the far vertices of a far square sit at ids N.. (their cells are
// never kept, whatever the SDF or a surrogate says there).
template <int DIM, int MODE>
__global__ void k_keep(int64_t T, const int32_t *tri, const double *p, const sm_node *tape,
                       int ntape, double geps, int32_t N, uint8_t *keep) {
  int64_t t = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (t >= T) return;
  const int32_t *c = tri + (DIM + 1) * t;
  bool real = true;
#pragma unroll
  for (int k = 0; k <= DIM; ++k) real &= (uint32_t)c[k] < (uint32_t)N;
  if (!real) {
    keep[t] = 0;
    return;
  }
  double x[3];
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    double s = p[(int64_t)c[0] * DIM + d];
#pragma unroll
    for (int k = 1; k <= DIM; ++k) s = s + p[(int64_t)c[k] * DIM + d];
    x[d] = s / (double)(DIM + 1);
  }
  keep[t] = sdf_eval<DIM, MODE>(tape, ntape, x) < -geps;
}

// Batch handles: a triangle belongs to the domain of its vertices (the
// triangulations of different domains never share a triangle); its
// centroid is tested against that domain's program and geps.
template <int MODE>
__global__ void k_keep_batch(int64_t T, const int32_t *tri, const double *p, BatchView bv,
                             int32_t N, uint8_t *keep) {
  int64_t t = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (t >= T) return;
  const int32_t *c = tri + 3 * t;
  if ((uint32_t)c[0] >= (uint32_t)N || (uint32_t)c[1] >= (uint32_t)N ||
      (uint32_t)c[2] >= (uint32_t)N) {
    keep[t] = 0;
    return;
  }
  const int dom = bv.vdom[c[0]];
  double x[2];
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    double s = p[(int64_t)c[0] * 2 + d];
    s = s + p[(int64_t)c[1] * 2 + d];
    s = s + p[(int64_t)c[2] * 2 + d];
    x[d] = s / 3.0;
  }
  const double geps = bv.geps_mult * bv.h[dom];
  keep[t] = sdf_eval<2, MODE>(bv.tapes + bv.tape_off[dom], bv.tape_len[dom], x) < -geps;
}

// second-endpoint lists: count, fill, order
__global__ void k_count_second(const int32_t *Ep, const int32_t *edges, int32_t *cnt) {
  const int64_t E = *Ep;
  int64_t e = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (e < E) atomicAdd(&cnt[edges[2 * e + 1]], 1);
}
__global__ void k_fill_second(const int32_t *Ep, const int32_t *edges, const int32_t *ptr,
                              int32_t *cur, int32_t *list) {
  const int64_t E = *Ep;
  int64_t e = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (e >= E) return;
  int b = edges[2 * e + 1];
  list[ptr[b] + atomicAdd(&cur[b], 1)] = (int32_t)e;
}
__global__ void k_sort_second(int64_t N, const int32_t *ptr, int32_t *list) {
  int64_t v = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (v >= N) return;
  int32_t s = ptr[v], e = ptr[v + 1];
  for (int32_t k = s + 1; k < e; ++k) {
    int32_t x = list[k];
    int32_t j = k - 1;
    while (j >= s && list[j] > x) {
      list[j + 1] = list[j];
      --j;
    }
    list[j + 1] = x;
  }
}

template <typename T> __device__ inline T zero_len();
template <> __device__ inline double zero_len<double>() { return 2.220446049250313e-16; }
template <> __device__ inline float zero_len<float>() { return 2.220446049250313e-16f; }

// x^3 rounded once: x*x kept as a double-double (fma remainder), times x
// with the product's remainder folded back before the final rounding.
__device__ inline double cube_rn(double x) {
  double p = x * x;
  double pe = fma(x, x, -p);
  double hi = p * x;
  double lo = fma(p, x, -hi) + pe * x;
  return hi + lo;
}

// Per-edge bar length and sizing; per-block partial sums of L^d and h^d
// (FP64) for the fast path, or the per-edge powers for the exact path.
template <int DIM, typename T, bool EXACT>
__global__ void k_edge_geom(const int32_t *Ep, const int32_t *edges, const double *p,
                            GridView grid, const double *h_ext, BatchView bv, T *L_out, T *h_out,
                            double *Lp, double *hp, double *part) {
  const int64_t E = *Ep;
  const unsigned blk = xcd_block();  // the partial sums are stored by this index
  int64_t e = (int64_t)blk * blockDim.x + threadIdx.x;
  double sl = 0.0, sh = 0.0;
  if (e < E) {
    const int a = edges[2 * e], b = edges[2 * e + 1];
    T L2 = 0;
    double mid[3];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      double pa = p[(int64_t)a * DIM + d], pb = p[(int64_t)b * DIM + d];
      T bar = (T)(pa - pb);
      L2 = L2 + bar * bar;
      mid[d] = (pa + pb) / 2.0;
    }
    T L = sqrt(L2);
    if (L == (T)0) L = zero_len<T>();
    T h;
    if (bv.vdom) h = (T)bv.h[bv.vdom[a]];
    else if (h_ext) h = (T)h_ext[e];
    else h = (T)grid_eval<DIM>(grid, mid);
    L_out[e] = L;
    h_out[e] = h;
    if (EXACT) {
      if (DIM == 2) {
        Lp[e] = (double)L * (double)L;
        hp[e] = (double)h * (double)h;
      } else {
        Lp[e] = cube_rn((double)L);
        hp[e] = cube_rn((double)h);
      }
    } else {
      double Ld = L, hd = h;
      sl = DIM == 2 ? Ld * Ld : Ld * Ld * Ld;
      sh = DIM == 2 ? hd * hd : hd * hd * hd;
    }
  }
  if (!EXACT) {
    // fixed-order block reduction (deterministic)
    __shared__ double rl[256], rh[256];
    rl[threadIdx.x] = sl;
    rh[threadIdx.x] = sh;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) {
        rl[threadIdx.x] += rl[threadIdx.x + w];
        rh[threadIdx.x] += rh[threadIdx.x + w];
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      part[2 * blk] = rl[0];
      part[2 * blk + 1] = rh[0];
    }
  }
}

// sums the per-block partials in a fixed order; sums[2] = scale (fast path)
__global__ void k_sum_partials(int64_t nb, const double *part, double *sums, int dim) {
  __shared__ double rl[256], rh[256];
  double sl = 0.0, sh = 0.0;
  for (int64_t i = threadIdx.x; i < nb; i += blockDim.x) {
    sl += part[2 * i];
    sh += part[2 * i + 1];
  }
  rl[threadIdx.x] = sl;
  rh[threadIdx.x] = sh;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      rl[threadIdx.x] += rl[threadIdx.x + w];
      rh[threadIdx.x] += rh[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    sums[0] = rl[0];
    sums[1] = rh[0];
    double r = rl[0] / rh[0];
    sums[2] = dim == 2 ? sqrt(r) : cbrt(r);
  }
}

// numpy's float64 add.reduce over a contiguous array: sequential over chunks
// of 8192 elements; inside a chunk, pairwise with leaves of <= 128 elements
// summed by 8 interleaved accumulators.  One workgroup per chunk computes
// the leaves in parallel and combines them in the recursion's order.
constexpr int kNpChunk = 8192;
constexpr int kNpMaxLeaves = 256;

__device__ double np_leaf(const double *a, int n) {
  if (n < 8) {
    double r = 0.0;
    for (int i = 0; i < n; ++i) r += a[i];
    return r;
  }
  double r[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) r[j] = a[j];
  int i;
  for (i = 8; i < n - (n % 8); i += 8) {
#pragma unroll
    for (int j = 0; j < 8; ++j) r[j] += a[i + j];
  }
  double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
  for (; i < n; ++i) res += a[i];
  return res;
}

__global__ void k_np_chunks(const double *a, int64_t n, double *chunk_sums) {
  __shared__ int loff[kNpMaxLeaves], llen[kNpMaxLeaves];
  __shared__ double lsum[kNpMaxLeaves];
  __shared__ int nleaves;
  const int64_t c0 = (int64_t)blockIdx.x * kNpChunk;
  const int m = (int)((n - c0) < kNpChunk ? (n - c0) : kNpChunk);
  if (threadIdx.x == 0) {
    // enumerate leaves left to right with an explicit stack
    int so[32], sn[32], sp = 0, L = 0;
    so[sp] = 0;
    sn[sp] = m;
    ++sp;
    while (sp > 0) {
      --sp;
      int o = so[sp], k = sn[sp];
      if (k <= 128) {
        loff[L] = o;
        llen[L] = k;
        ++L;
      } else {
        int k2 = k / 2;
        k2 -= k2 % 8;
        // push right first so the left half is processed first
        so[sp] = o + k2; sn[sp] = k - k2; ++sp;
        so[sp] = o; sn[sp] = k2; ++sp;
      }
    }
    nleaves = L;
  }
  __syncthreads();
  for (int l = threadIdx.x; l < nleaves; l += blockDim.x)
    lsum[l] = np_leaf(a + c0 + loff[l], llen[l]);
  __syncthreads();
  if (threadIdx.x == 0) {
    // combine: recursion pw(o,k) = pw(left) + pw(right), leaves consumed in
    // order; iterative post-order with a value stack
    int so[32], sn[32], st[32], sp = 0;
    double vs[32];
    int vp = 0, next = 0;
    so[0] = 0; sn[0] = m; st[0] = 0; sp = 1;
    while (sp > 0) {
      int top = sp - 1;
      int k = sn[top];
      if (k <= 128) {
        vs[vp++] = lsum[next++];
        --sp;
        continue;
      }
      int k2 = k / 2;
      k2 -= k2 % 8;
      if (st[top] == 0) {
        st[top] = 1;
        so[sp] = so[top]; sn[sp] = k2; st[sp] = 0; ++sp;
      } else if (st[top] == 1) {
        st[top] = 2;
        so[sp] = so[top] + k2; sn[sp] = k - k2; st[sp] = 0; ++sp;
      } else {
        double r = vs[--vp];
        double l = vs[--vp];
        vs[vp++] = l + r;
        --sp;
      }
    }
    chunk_sums[blockIdx.x] = vs[0];
  }
}

// Batch handles: one workgroup per domain sums L^d and h^d over the
// domain's edges (a contiguous range of the sorted edge list: the rows of
// its vertex range) in a fixed order, then scale = (sum L^d / sum h^d)^(1/d).
template <typename T>
__global__ void k_domain_sums(const int32_t *eptr, BatchView bv, const T *L, const T *h,
                              int dim) {
  __shared__ double rl[256], rh[256];
  const int dom = blockIdx.x;
  const int32_t es = eptr[bv.voff[dom]], ee = eptr[bv.voff[dom + 1]];
  double sl = 0.0, sh = 0.0;
  for (int32_t e = es + (int32_t)threadIdx.x; e < ee; e += (int32_t)blockDim.x) {
    const double Le = L[e], he = h[e];
    sl += dim == 2 ? Le * Le : Le * Le * Le;
    sh += dim == 2 ? he * he : he * he * he;
  }
  rl[threadIdx.x] = sl;
  rh[threadIdx.x] = sh;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      rl[threadIdx.x] += rl[threadIdx.x + w];
      rh[threadIdx.x] += rh[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double r = rh[0] > 0.0 ? rl[0] / rh[0] : 0.0;
    bv.scale[dom] = dim == 2 ? sqrt(r) : cbrt(r);
  }
}

__global__ void k_np_total(const double *chunk_sums, int64_t nchunks, double *out) {
  double r = chunk_sums[0];
  for (int64_t i = 1; i < nchunks; ++i) r += chunk_sums[i];
  *out = r;
}

template <int DIM, typename T>
// maxF_clear (optional): the vertex update's movement maximum, cleared here
// (this launch always precedes the update and has at least one workgroup)
__global__ void k_edge_force(const int32_t *Ep, const int32_t *edges, const double *p,
                             const T *L, const T *h, const double *scale_ptr, BatchView bv,
                             double L0mult, T *F, uint64_t *maxF_clear) {
  if (maxF_clear && blockIdx.x == 0 && threadIdx.x == 0) *maxF_clear = 0ull;
  const int64_t E = *Ep;
  int64_t e = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const int a = edges[2 * e], b = edges[2 * e + 1];
  const T scale = bv.vdom ? (T)bv.scale[bv.vdom[a]] : (T)scale_ptr[2];
  const T Le = L[e];
  T L0 = h[e] * (T)L0mult * scale;
  T f = L0 - Le;
  if (f < (T)0) f = (T)0;
  T r = f / Le;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    T bar = (T)(p[(int64_t)a * DIM + d] - p[(int64_t)b * DIM + d]);
    F[e * DIM + d] = r * bar;
  }
}

template <int DIM, int MODE>
__device__ inline double project_level(double *x, const sm_node *tape, int ntape, double deps,
                                       double hmin, int level) {
  double d = sdf_eval<DIM, MODE>(tape, ntape, x);
  bool in = level == 0 ? (d > 0.0) : (d > 0.0 && d < hmin / 1.5);
  if (!in) return d;
  double g[DIM];
#pragma unroll
  for (int i = 0; i < DIM; ++i) {
    double y[3];
#pragma unroll
    for (int k = 0; k < DIM; ++k) y[k] = x[k] + (k == i ? deps : 0.0);
    g[i] = (sdf_eval<DIM, MODE>(tape, ntape, y) - d) / deps;
  }
  double g2 = g[0] * g[0];
#pragma unroll
  for (int i = 1; i < DIM; ++i) g2 = g2 + g[i] * g[i];
  if (g2 < deps) g2 = deps;
#pragma unroll
  for (int i = 0; i < DIM; ++i) x[i] = x[i] - (d * g[i]) / g2;
  return d;
}

// atomicMax of the 64-bit patterns over a whole wave with one atomic: the
// wave's maximum is taken by shuffles first (every lane of the wave must
// call this).  For non-negative doubles the bit patterns order as the
// values; a NaN pattern exceeds every finite one, as with a per-lane atomic.
__device__ inline void wave_max_bits(uint64_t *dst, uint64_t bits) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const uint64_t o = (uint64_t)__shfl_xor((unsigned long long)bits, off);
    bits = o > bits ? o : bits;
  }
  if (__lane_id() == 0) atomicMax((unsigned long long *)dst, (unsigned long long)bits);
}

template <int DIM, typename T, int MODE>
__global__ void k_vertex_update(int64_t N, double *p, const int32_t *first_ptr,
                                const int32_t *second_ptr, const int32_t *second, const T *F,
                                int nfix, double dt, const sm_node *tapes, const int *tape_off,
                                const int *tape_len, int nlevels, double deps, double hmin,
                                BatchView bv, uint64_t *maxF, int project) {
  const int64_t v = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  const bool live = v < N;
  T f[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) f[d] = (T)0;
  int dom = 0;
  uint64_t sbits = 0;
  if (live) {
    for (int32_t k = second_ptr[v], ke = second_ptr[v + 1]; k < ke; ++k) {
      int64_t e = second[k];
#pragma unroll
      for (int d = 0; d < DIM; ++d) f[d] = f[d] + (-F[e * DIM + d]);
    }
    for (int32_t e = first_ptr[v], ee = first_ptr[v + 1]; e < ee; ++e) {
#pragma unroll
      for (int d = 0; d < DIM; ++d) f[d] = f[d] + F[(int64_t)e * DIM + d];
    }
    bool fixed = v < nfix;
    if (bv.vdom) {
      dom = bv.vdom[v];
      fixed = v - bv.voff[dom] < bv.nfix[dom];
    }
    if (fixed) {
#pragma unroll
      for (int d = 0; d < DIM; ++d) f[d] = (T)0;
    }
    double s2 = (double)f[0] * (double)f[0];
#pragma unroll
    for (int d = 1; d < DIM; ++d) s2 = s2 + (double)f[d] * (double)f[d];
    double s = sqrt(s2);
    // batches track the movement relative to each domain's h0
    if (bv.vdom) s = s / bv.h[dom];
    sbits = (uint64_t)__double_as_longlong(s);
  }
  // every lane of the wave reaches this point (lanes past the end add 0)
  wave_max_bits(maxF, sbits);
  if (!live) return;
  double x[3];
#pragma unroll
  for (int d = 0; d < DIM; ++d) x[d] = p[v * DIM + d] + dt * (double)f[d];
  if (!project) {
    // the surrogate path projects in k_project_surrogate
  } else if (bv.vdom) {
    const double h0 = bv.h[dom];
    project_level<DIM, MODE>(x, bv.tapes + bv.tape_off[dom], bv.tape_len[dom],
                             bv.deps_mult * h0, h0, 0);
  } else {
    for (int l = 0; l < nlevels; ++l)
      project_level<DIM, MODE>(x, tapes + tape_off[l], tape_len[l], deps, hmin, l);
  }
#pragma unroll
  for (int d = 0; d < DIM; ++d) p[v * DIM + d] = x[d];
}

// counters of the surrogate: one atomic per wave
__device__ inline void wave_count(uint64_t *ctr, bool pred) {
  const uint64_t act = __ballot(1);
  const uint64_t hit = __ballot(pred);
  if ((int)__lane_id() == __ffsll((long long)act) - 1 && hit)
    atomicAdd((unsigned long long *)ctr, (unsigned long long)__popcll(hit));
}

__device__ inline void record_error(uint64_t *ctr, double dn, double d) {
  const double err = fabs(dn - d);
  if (err == err) atomicMax((unsigned long long *)ctr, (unsigned long long)__double_as_longlong(err));
}

// Centroid filter with the surrogate: outside the band |d_nn| < band(x)
// the network's sign decides (a centroid with d_nn <= -band is kept, one
// with d_nn >= band dropped); inside the band, or where d_nn is not finite,
// the exact program decides as in k_keep.  The decisions agree with k_keep
// wherever |d_nn - d| < band(x) - geps.
template <int DIM, int MODE>
__global__ void k_keep_surrogate(int64_t T, const int32_t *tri, const double *p,
                                 const sm_node *tape, int ntape, double geps, const float *dnn,
                                 BandView bv, int32_t N, uint8_t *keep, uint64_t *sg) {
  const int64_t t = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (t >= T) return;
  const int32_t *c = tri + (DIM + 1) * t;
  bool real = true;
#pragma unroll
  for (int k = 0; k <= DIM; ++k) real &= (uint32_t)c[k] < (uint32_t)N;
  const double dn = real ? (double)dnn[t] : 0.0;
  double x[3] = {0.0, 0.0, 0.0};
  // the centroid is needed for a per-region band, or for the exact test
  const bool need_x = real && (bv.v != nullptr || !(fabs(dn) >= bv.band));
  if (need_x) {
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      double s = p[(int64_t)c[0] * DIM + d];
#pragma unroll
      for (int k = 1; k <= DIM; ++k) s = s + p[(int64_t)c[k] * DIM + d];
      x[d] = s / (double)(DIM + 1);
    }
  }
  const bool exact = real && !(fabs(dn) >= (bv.v ? band_at<DIM>(bv, x) : bv.band));
  wave_count(sg + 0, real && !exact);
  wave_count(sg + 1, exact);
  if (!real) {
    keep[t] = 0;
  } else if (!exact) {
    keep[t] = dn < 0.0;
  } else {
    const double d = sdf_eval<DIM, MODE>(tape, ntape, x);
    record_error(sg + 2, dn, d);
    keep[t] = d < -geps;
  }
}

// Projection with the surrogate: points with d_nn <= -band(x) are inside
// (no projection); every other point is tested and projected with the
// exact program, as in k_vertex_update.
template <int DIM, int MODE>
__global__ void k_project_surrogate(int64_t N, double *p, const float *dnn, BandView bv,
                                    const sm_node *tape, int ntape, double deps, double hmin,
                                    uint64_t *sg) {
  const int64_t v = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (v >= N) return;
  const double dn = (double)dnn[v];
  double band = bv.band;
  if (bv.v) {
    double y[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int d = 0; d < DIM; ++d) y[d] = p[v * DIM + d];
    band = band_at<DIM>(bv, y);
  }
  const bool exact = !(dn <= -band);
  wave_count(sg + 0, !exact);
  wave_count(sg + 1, exact);
  if (!exact) return;
  double x[3];
#pragma unroll
  for (int d = 0; d < DIM; ++d) x[d] = p[v * DIM + d];
  const double d = project_level<DIM, MODE>(x, tape, ntape, deps, hmin, 0);
  record_error(sg + 2, dn, d);
#pragma unroll
  for (int k = 0; k < DIM; ++k) p[v * DIM + k] = x[k];
}

}  // namespace

template <int DIM, int MODE>
static void launch_keep(sm_mesher *m, int64_t off, int64_t T) {
  const int B = m->blk[kBlkKeep];
  hipLaunchKernelGGL((k_keep<DIM, MODE>), dim3(blocks_for(T, B)), dim3(B), 0, m->stream, T,
                     m->d_tri + (DIM + 1) * off, m->d_p, m->d_tapes[0], m->tape_len[0],
                     m->prm.geps, (int32_t)m->N, m->d_keep + off);
}

template <int DIM, int MODE>
static void launch_keep_surrogate(sm_mesher *m, int64_t off, int64_t T) {
  const int B = m->blk[kBlkKeep];
  hipLaunchKernelGGL((k_keep_surrogate<DIM, MODE>), dim3(blocks_for(T, B)), dim3(B), 0,
                     m->stream, T, m->d_tri + (DIM + 1) * off, m->d_p, m->d_tapes[0],
                     m->tape_len[0], m->prm.geps, m->d_dnn + off, band_view(m), (int32_t)m->N,
                     m->d_keep + off, m->d_sg);
}

template <int DIM, int MODE>
static int project_surrogate(sm_mesher *m) {
  int rc = mlp_eval_points(&m->net, m->d_p, m->N, m->d_dnn, m->stream);
  if (rc) return rc;
  const int B = 256;
  hipLaunchKernelGGL((k_project_surrogate<DIM, MODE>), dim3(blocks_for(m->N, B)), dim3(B), 0,
                     m->stream, m->N, m->d_p, m->d_dnn, band_view(m), m->d_tapes[0],
                     m->tape_len[0], m->prm.deps, m->prm.h0, m->d_sg);
  SM_LAUNCH_CHECK();
  return SM_OK;
}

// Domains given as host functions (sm_mesher_set_host_sdf).  The centroid
// filter (mesh_generator.py:734-738): centroids summed in the cell's vertex
// order and divided by dim + 1, as k_keep does; keep d < -geps.
static int host_keep(sm_mesher *m) {
  const int dim = m->dim, nv = dim + 1;
  const int64_t T = m->T, N = m->N;
  std::vector<int32_t> tri((size_t)T * nv);
  std::vector<double> p((size_t)N * dim);
  SM_HIP(hipMemcpyAsync(tri.data(), m->d_tri, tri.size() * sizeof(int32_t),
                        hipMemcpyDeviceToHost, m->stream));
  SM_HIP(hipMemcpyAsync(p.data(), m->d_p, p.size() * sizeof(double), hipMemcpyDeviceToHost,
                        m->stream));
  SM_HIP(hipStreamSynchronize(m->stream));
  std::vector<int64_t> ids;
  std::vector<double> x;
  ids.reserve((size_t)T);
  x.reserve((size_t)T * dim);
  for (int64_t t = 0; t < T; ++t) {
    const int32_t *c = tri.data() + nv * t;
    bool real = true;
    for (int k = 0; k < nv; ++k) real &= (uint32_t)c[k] < (uint32_t)N;
    if (!real) continue;
    ids.push_back(t);
    for (int d = 0; d < dim; ++d) {
      double acc = p[(size_t)c[0] * dim + d];
      for (int k = 1; k < nv; ++k) acc = acc + p[(size_t)c[k] * dim + d];
      x.push_back(acc / (double)nv);
    }
  }
  std::vector<double> dv(ids.size());
  if (!ids.empty()) {
    const int rc = m->host_sdf(m->host_sdf_ctx, 0, x.data(), (int64_t)ids.size(), dv.data());
    SM_CHECK(rc == 0, SM_ERR_ARG, "host SDF callback failed (%d)", rc);
  }
  std::vector<uint8_t> keep((size_t)T, 0);
  for (size_t k = 0; k < ids.size(); ++k) keep[(size_t)ids[k]] = dv[k] < -m->prm.geps;
  SM_HIP(hipMemcpyAsync(m->d_keep, keep.data(), (size_t)T, hipMemcpyHostToDevice, m->stream));
  SM_HIP(hipStreamSynchronize(m->stream));
  return SM_OK;
}

// _project_points_back_newton (mesh_generator.py:762-784) for each level in
// turn, with the host function: level 0 projects d > 0, a subdomain level
// 0 < d < h0 / 1.5; forward-difference gradient with step deps, |g|^2
// clamped to deps, x -= (d g_i) / |g|^2 (the update kernel's arithmetic).
static int host_project(sm_mesher *m) {
  const int dim = m->dim;
  const int64_t N = m->N;
  const double deps = m->prm.deps, hmin = m->prm.h0;
  std::vector<double> p((size_t)N * dim);
  SM_HIP(hipMemcpyAsync(p.data(), m->d_p, p.size() * sizeof(double), hipMemcpyDeviceToHost,
                        m->stream));
  SM_HIP(hipStreamSynchronize(m->stream));
  std::vector<double> d((size_t)N), y, dy, g;
  std::vector<int64_t> ix;
  for (int l = 0; l < m->prm.nlevels; ++l) {
    int rc = m->host_sdf(m->host_sdf_ctx, l, p.data(), N, d.data());
    SM_CHECK(rc == 0, SM_ERR_ARG, "host SDF callback failed (%d)", rc);
    ix.clear();
    for (int64_t v = 0; v < N; ++v) {
      const double dv = d[(size_t)v];
      if (l == 0 ? dv > 0.0 : (dv > 0.0 && dv < hmin / 1.5)) ix.push_back(v);
    }
    const int64_t n = (int64_t)ix.size();
    if (n == 0) continue;
    y.assign((size_t)n * dim, 0.0);
    dy.assign((size_t)n, 0.0);
    g.assign((size_t)n * dim, 0.0);
    for (int i = 0; i < dim; ++i) {
      for (int64_t k = 0; k < n; ++k)
        for (int c = 0; c < dim; ++c)
          y[(size_t)k * dim + c] = p[(size_t)ix[k] * dim + c] + (c == i ? deps : 0.0);
      rc = m->host_sdf(m->host_sdf_ctx, l, y.data(), n, dy.data());
      SM_CHECK(rc == 0, SM_ERR_ARG, "host SDF callback failed (%d)", rc);
      for (int64_t k = 0; k < n; ++k)
        g[(size_t)k * dim + i] = (dy[(size_t)k] - d[(size_t)ix[k]]) / deps;
    }
    for (int64_t k = 0; k < n; ++k) {
      const double *gk = g.data() + (size_t)k * dim;
      double g2 = gk[0] * gk[0];
      for (int i = 1; i < dim; ++i) g2 = g2 + gk[i] * gk[i];
      if (g2 < deps) g2 = deps;
      const double dk = d[(size_t)ix[k]];
      for (int i = 0; i < dim; ++i) {
        double &pv = p[(size_t)ix[k] * dim + i];
        pv = pv - (dk * gk[i]) / g2;
      }
    }
  }
  SM_HIP(hipMemcpyAsync(m->d_p, p.data(), p.size() * sizeof(double), hipMemcpyHostToDevice,
                        m->stream));
  SM_HIP(hipStreamSynchronize(m->stream));
  return SM_OK;
}

int centroid_keep(sm_mesher *m) {
  if (m->T == 0) return SM_OK;
  if (m->host_sdf) return host_keep(m);
  if (m->ndom > 0) {
    const int B = m->blk[kBlkKeep];
    if (m->deep_sdf)
      hipLaunchKernelGGL((k_keep_batch<kSdfDeep>), dim3(blocks_for(m->T, B)), dim3(B), 0,
                         m->stream, m->T, m->d_tri, m->d_p, batch_view(m), (int32_t)m->N,
                         m->d_keep);
    else
      hipLaunchKernelGGL((k_keep_batch<kSdfReg>), dim3(blocks_for(m->T, B)), dim3(B), 0,
                         m->stream, m->T, m->d_tri, m->d_p, batch_view(m), (int32_t)m->N,
                         m->d_keep);
    SM_LAUNCH_CHECK();
    return SM_OK;
  }
  bool done = false;
  int rc = centroid_keep_range(m, 0, m->T, &done);
  return rc;
}

int centroid_keep_range(sm_mesher *m, int64_t off, int64_t T, bool *done) {
  *done = false;
  if (m->host_sdf || m->ndom > 0) return SM_OK;
  *done = true;
  if (T <= 0) return SM_OK;
  if (m->surrogate) {
    const int pb = prof_begin(m);
    int rc = mlp_eval_cells(&m->net, m->d_tri + (m->dim + 1) * off, T, m->d_p, m->d_dnn + off,
                            m->stream);
    prof_end(m, 8, pb);
    if (rc) return rc;
    if (m->dim == 2) {
      if (m->deep_sdf) launch_keep_surrogate<2, kSdfDeep>(m, off, T);
      else launch_keep_surrogate<2, kSdfReg>(m, off, T);
    } else {
      if (m->deep_sdf) launch_keep_surrogate<3, kSdfDeep>(m, off, T);
      else launch_keep_surrogate<3, kSdfReg>(m, off, T);
    }
    SM_LAUNCH_CHECK();
    return SM_OK;
  }
  if (m->dim == 2) {
    if (m->deep_sdf) launch_keep<2, kSdfDeep>(m, off, T);
    else launch_keep<2, kSdfReg>(m, off, T);
  } else {
    if (m->deep_sdf) launch_keep<3, kSdfDeep>(m, off, T);
    else launch_keep<3, kSdfReg>(m, off, T);
  }
  SM_LAUNCH_CHECK();
  return SM_OK;
}

// first-endpoint offsets live in the unique-edge workspace
static int32_t *eptr_of(sm_mesher *m) {
  return unique_edges_offsets(m->d_ews, m->capT * (m->dim == 2 ? 3 : 6), m->N, m->dim == 2);
}

// Launch bound for per-edge kernels: a triangulation of N points has at
// most 3N - 6 edges, so kernels sized for 3N read the actual count on the
// device and no host round trip is needed between the edge build and the
// forces.
static int64_t edge_launch_bound(const sm_mesher *m) {
  // 3D: no planar bound; the 3D loop reads the edge count back (it
  // synchronises for the host tetrahedralisation anyway)
  return m->dim == 2 ? 3 * m->N + 8 : std::max<int64_t>(m->E, 1);
}

int build_edges(sm_mesher *m, bool sync_edges) {
  hipStream_t s = m->stream;
  const int dim = m->dim;
  int64_t np = m->capT * (dim == 2 ? 3 : 6);
  int32_t *Ep = m->d_ctr + kCtrE;
  // 2D triangulations carry consistent neighbour links here (device flips
  // end with the link fix; every host rebuild writes both arrays); the
  // builder then also fills the second-endpoint lists
  const bool links = dim == 2;
  const SecondLists second{m->d_sptr, m->d_scur, m->d_slist};
  int rc = unique_edges_run(nullptr, m->d_tri, m->d_keep, links ? m->d_adj : nullptr,
                            links ? &second : nullptr, dim, m->T, np, m->N, m->d_edges, Ep,
                            nullptr, m->d_ews, m->ews_bytes, s);
  if (rc) return rc;
  if (sync_edges) {
    SM_HIP(hipMemcpyAsync(m->h_ctr + kCtrE, Ep, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    SM_HIP(hipStreamSynchronize(s));
    m->E = m->h_ctr[kCtrE];
  }
  const int B = 256;
  const int64_t N = m->N, Eb = edge_launch_bound(m);
  if (!links) {
    FillList fl{{m->d_sptr, m->d_scur, nullptr, nullptr}, {N + 1, N + 1, 0, 0}};
    if ((rc = fill_list(fl, 2, 0, s))) return rc;
    hipLaunchKernelGGL(k_count_second, dim3(blocks_for(Eb, B)), dim3(B), 0, s,
                       (const int32_t *)Ep, m->d_edges, m->d_sptr);
    SM_LAUNCH_CHECK();
    if ((rc = scan_exclusive(m->d_sptr, m->d_sptr, N + 1, nullptr, m->d_scan_ws, s))) return rc;
    hipLaunchKernelGGL(k_fill_second, dim3(blocks_for(Eb, B)), dim3(B), 0, s,
                       (const int32_t *)Ep, m->d_edges, m->d_sptr, m->d_scur, m->d_slist);
    SM_LAUNCH_CHECK();
  }
  // either way each vertex's list is then put in increasing edge order
  hipLaunchKernelGGL(k_sort_second, dim3(blocks_for(N, B)), dim3(B), 0, s, N, m->d_sptr,
                     m->d_slist);
  SM_LAUNCH_CHECK();
  return SM_OK;
}

template <int DIM, typename T>
static void launch_edge_force(sm_mesher *m, int B) {
  const unsigned g = blocks_for(edge_launch_bound(m), B);
  hipLaunchKernelGGL((k_edge_force<DIM, T>), dim3(g), dim3(B), 0, m->stream,
                     (const int32_t *)(m->d_ctr + kCtrE), m->d_edges, m->d_p, (const T *)m->d_L,
                     (const T *)m->d_h, m->d_sums, batch_view(m), m->prm.L0mult, (T *)m->d_F,
                     m->d_maxF);
}

template <int DIM, typename T>
static void launch_vertex_update(sm_mesher *m, int B) {
  const int64_t N = m->N;
  const BatchView bv = batch_view(m);
  if (m->deep_sdf)
    hipLaunchKernelGGL((k_vertex_update<DIM, T, kSdfDeep>), dim3(blocks_for(N, B)), dim3(B), 0,
                       m->stream, N, m->d_p, eptr_of(m), m->d_sptr, m->d_slist,
                       (const T *)m->d_F, m->prm.nfix, m->prm.delta_t, m->d_tape_all,
                       m->d_tape_off_dev, m->d_tape_len_dev, m->prm.nlevels, m->prm.deps,
                       m->prm.h0, bv, m->d_maxF, (m->surrogate || m->host_sdf) ? 0 : 1);
  else
    hipLaunchKernelGGL((k_vertex_update<DIM, T, kSdfReg>), dim3(blocks_for(N, B)), dim3(B), 0,
                       m->stream, N, m->d_p, eptr_of(m), m->d_sptr, m->d_slist,
                       (const T *)m->d_F, m->prm.nfix, m->prm.delta_t, m->d_tape_all,
                       m->d_tape_off_dev, m->d_tape_len_dev, m->prm.nlevels, m->prm.deps,
                       m->prm.h0, bv, m->d_maxF, (m->surrogate || m->host_sdf) ? 0 : 1);
}

template <int DIM, typename T>
static int forces_t(sm_mesher *m) {
  hipStream_t s = m->stream;
  const int B = 256;  // k_edge_geom's block reduction assumes 256 threads
  const int64_t E = m->E;
  const bool exact = m->prm.precision == 64;
  const unsigned ge = blocks_for(edge_launch_bound(m), B);
  const int32_t *Ep = m->d_ctr + kCtrE;
  GridView g = m->grid;
  const BatchView bv = batch_view(m);
  int pb = prof_begin(m);
  if (exact)
    hipLaunchKernelGGL((k_edge_geom<DIM, T, true>), dim3(ge), dim3(B), 0, s, Ep, m->d_edges,
                       m->d_p, g, m->d_hext, bv, (T *)m->d_L, (T *)m->d_h, m->d_Lp, m->d_hp,
                       m->d_part);
  else
    hipLaunchKernelGGL((k_edge_geom<DIM, T, false>), dim3(ge), dim3(B), 0, s, Ep, m->d_edges,
                       m->d_p, g, m->d_hext, bv, (T *)m->d_L, (T *)m->d_h, m->d_Lp, m->d_hp,
                       m->d_part);
  SM_LAUNCH_CHECK();
  prof_end(m, 3, pb);
  pb = prof_begin(m);
  if (m->ndom > 0) {
    hipLaunchKernelGGL((k_domain_sums<T>), dim3(m->ndom), dim3(256), 0, s,
                       (const int32_t *)eptr_of(m), bv, (const T *)m->d_L, (const T *)m->d_h,
                       DIM);
    SM_LAUNCH_CHECK();
  } else if (exact) {
    int64_t nch = (E + kNpChunk - 1) / kNpChunk;
    if (nch < 1) nch = 1;
    if (E > 0) {
      hipLaunchKernelGGL(k_np_chunks, dim3((unsigned)nch), dim3(64), 0, s, m->d_Lp, E,
                         m->d_part);
      SM_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_np_total, dim3(1), dim3(1), 0, s, m->d_part, nch, m->d_sums);
      SM_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_np_chunks, dim3((unsigned)nch), dim3(64), 0, s, m->d_hp, E,
                         m->d_part);
      SM_LAUNCH_CHECK();
      hipLaunchKernelGGL(k_np_total, dim3(1), dim3(1), 0, s, m->d_part, nch, m->d_sums + 1);
      SM_LAUNCH_CHECK();
    }
    // (sum L^d / sum h^d) ** (1/d) with the C library's pow, as numpy does
    SM_HIP(hipMemcpyAsync(m->h_sums, m->d_sums, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    SM_HIP(hipStreamSynchronize(s));
    m->h_sums[2] = std::pow(m->h_sums[0] / m->h_sums[1], 1.0 / DIM);
    SM_HIP(hipMemcpyAsync(m->d_sums + 2, m->h_sums + 2, sizeof(double), hipMemcpyHostToDevice, s));
  } else {
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, (int64_t)ge, m->d_part,
                       m->d_sums, DIM);
    SM_LAUNCH_CHECK();
  }
  prof_end(m, 4, pb);
  pb = prof_begin(m);
  launch_edge_force<DIM, T>(m, m->blk[kBlkForce]);
  SM_LAUNCH_CHECK();
  prof_end(m, 5, pb);
  // k_edge_force cleared the movement maximum
  int rc = SM_OK;
  pb = prof_begin(m);
  launch_vertex_update<DIM, T>(m, m->blk[kBlkUpdate]);
  SM_LAUNCH_CHECK();
  if (m->surrogate) {
    rc = m->deep_sdf ? project_surrogate<DIM, kSdfDeep>(m) : project_surrogate<DIM, kSdfReg>(m);
    if (rc) return rc;
  }
  if (m->host_sdf && (rc = host_project(m))) return rc;
  prof_end(m, 6, pb);
  return SM_OK;
}

int forces_and_update(sm_mesher *m) {
  const bool f64 = m->prm.precision == 64;
  if (m->dim == 2) return f64 ? forces_t<2, double>(m) : forces_t<2, float>(m);
  return f64 ? forces_t<3, double>(m) : forces_t<3, float>(m);
}

template <int DIM, typename T>
static void launch_tunable(sm_mesher *m, int which, int B) {
  if (which == kBlkForce) {
    launch_edge_force<DIM, T>(m, B);
  } else if (which == kBlkUpdate) {
    launch_vertex_update<DIM, T>(m, B);
  } else if (m->ndom > 0) {
    if (m->deep_sdf)
      hipLaunchKernelGGL((k_keep_batch<kSdfDeep>), dim3(blocks_for(m->T, B)), dim3(B), 0,
                         m->stream, m->T, m->d_tri, m->d_p, batch_view(m), (int32_t)m->N,
                         m->d_keep);
    else
      hipLaunchKernelGGL((k_keep_batch<kSdfReg>), dim3(blocks_for(m->T, B)), dim3(B), 0,
                         m->stream, m->T, m->d_tri, m->d_p, batch_view(m), (int32_t)m->N,
                         m->d_keep);
  } else {
    if (m->deep_sdf)
      hipLaunchKernelGGL((k_keep<DIM, kSdfDeep>), dim3(blocks_for(m->T, B)), dim3(B), 0,
                         m->stream, m->T, m->d_tri, m->d_p, m->d_tapes[0], m->tape_len[0],
                         m->prm.geps, (int32_t)m->N, m->d_keep);
    else
      hipLaunchKernelGGL((k_keep<DIM, kSdfReg>), dim3(blocks_for(m->T, B)), dim3(B), 0,
                         m->stream, m->T, m->d_tri, m->d_p, m->d_tapes[0], m->tape_len[0],
                         m->prm.geps, (int32_t)m->N, m->d_keep);
  }
}

int time_force_tunable(sm_mesher *m, int which, int B, int reps, double *ms) {
  SM_CHECK(which == kBlkKeep || which == kBlkForce || which == kBlkUpdate, SM_ERR_ARG,
           "time_force_tunable: kernel %d", which);
  SM_CHECK(B >= 64 && B <= 1024 && B % 64 == 0 && reps >= 1, SM_ERR_ARG,
           "time_force_tunable: bad block size or repetition count");
  hipStream_t s = m->stream;
  const bool f64 = m->prm.precision == 64;
  auto launch = [&]() {
    if (m->dim == 2) {
      if (f64) launch_tunable<2, double>(m, which, B);
      else launch_tunable<2, float>(m, which, B);
    } else {
      if (f64) launch_tunable<3, double>(m, which, B);
      else launch_tunable<3, float>(m, which, B);
    }
  };
  // the update moves the points: time it on a copy of them and restore
  const size_t pbytes = (size_t)m->NA * m->dim * sizeof(double);
  if (which == kBlkUpdate)
    SM_HIP(hipMemcpyAsync(m->d_pstage, m->d_p, pbytes, hipMemcpyDeviceToDevice, s));
  hipEvent_t e0, e1;
  SM_HIP(hipEventCreate(&e0));
  SM_HIP(hipEventCreate(&e1));
  launch();
  SM_HIP(hipEventRecord(e0, s));
  for (int i = 0; i < reps; ++i) launch();
  SM_HIP(hipEventRecord(e1, s));
  SM_LAUNCH_CHECK();
  SM_HIP(hipEventSynchronize(e1));
  float el = 0.0f;
  SM_HIP(hipEventElapsedTime(&el, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (which == kBlkUpdate) {
    SM_HIP(hipMemcpyAsync(m->d_p, m->d_pstage, pbytes, hipMemcpyDeviceToDevice, s));
    SM_HIP(hipStreamSynchronize(s));
  }
  *ms = (double)el / reps;
  return SM_OK;
}

}  // namespace smk
