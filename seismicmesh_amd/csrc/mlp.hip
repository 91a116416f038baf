// Neural signed-distance surrogate: a multilayer perceptron (coordinates
// -> distance) with SiLU activations.
//
// Inference (k_mlp_eval_lds) is one fused kernel per batch of points: the
// input layer in FP32 on the vector ALUs, each width x width hidden layer as
// BF16 matrix-core products (v_mfma_f32_32x32x16_bf16, FP32 accumulation) on
// activations held in LDS, bias + SiLU in the epilogue of every layer, the
// output layer as a dot product.  Each wave owns 32 points (the M of one
// 32x32x16 product); a hidden layer is 8 column blocks x 16 k-steps, whose
// weight fragments the workgroup's 8 waves share through LDS.
//
// Training uses FP32 kernels: a tiled GEMM with strided operands and an
// optional deterministic split over k (forward X W^T, backward dZ W,
// weight gradients dZ^T X), bias + SiLU, its derivative, the loss with a
// central-difference eikonal term, and Adam.
//
// The reference has no counterpart: it evaluates the CSG functions of
// signed_distance_functions.py:283-623 directly.
#include <cmath>

#include "common.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

namespace smk {
namespace {

constexpr int kW = 256;       // hidden width of the inference kernel
constexpr int kRow = kW + 8;  // LDS row stride in bf16 (+16 B: row r starts 4 banks after r-1)
constexpr int kMaxHidden = 16;

__device__ inline float silu(float z) { return z / (1.0f + __expf(-z)); }

// Point i of the batch: a row of x, or the centroid of cell i (FP64 sum
// over the cell's vertices, divided by the vertex count).  Returns false
// for a cell with a negative vertex id.
template <int IN, bool CELLS>
__device__ inline bool mlp_input(const double *x, const int32_t *cells, int64_t i, double *q) {
  if (!CELLS) {
#pragma unroll
    for (int c = 0; c < IN; ++c) q[c] = x[i * IN + c];
    return true;
  }
  const int32_t *v = cells + (IN + 1) * i;
  bool real = true;
#pragma unroll
  for (int k = 0; k <= IN; ++k) real &= v[k] >= 0;
  if (!real) return false;
#pragma unroll
  for (int c = 0; c < IN; ++c) {
    double s = x[(int64_t)v[0] * IN + c];
#pragma unroll
    for (int k = 1; k <= IN; ++k) s = s + x[(int64_t)v[k] * IN + c];
    q[c] = s / (double)(IN + 1);
  }
  return true;
}

// SiLU with the hardware reciprocal (1 ulp) for the inference kernel, whose
// activations are rounded to BF16 (8 bits) anyway.
__device__ inline float silu_fast(float z) {
  return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}

typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

// The same network with the hidden layers' weights shared through LDS: a
// workgroup of WAVES waves stages each k-step's eight weight fragments
// (8 KiB) in LDS once -- three stages deep, each filled from a global load
// issued two k-steps earlier -- instead of every wave streaming the whole
// layer from L2.  Each wave reads the next k-step's fragments into a second
// register set before it issues this k-step's eight products, so no product
// waits on an LDS read.  The epilogue's biases and output weights are
// staged in LDS as well.  The launch uses 8 waves of one 32-point tile each
// (132 KiB of activations + 24 KiB of fragments + 3 KiB of epilogue
// operands per workgroup, two waves per SIMD).  Measured on MI355X at 4M
// points: 640 TFLOP/s (612 with two stages and the biases read from global
// memory in the epilogue, 511 with every wave streaming the weights from
// L2).  Counters put the limit in the vector ALU: the bias + SiLU epilogue
// of 512 activations per lane and the FP32 input layer take ~23 k VALU
// cycles per wave against ~12 k of matrix-core cycles, and both waves of a
// SIMD reach their epilogues together (the k-step barriers keep them in
// step), so the two do not overlap.
constexpr int kEvalWaves = 8;
constexpr int kEvalPT = 1;

template <int IN, bool CELLS, int PT, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_mlp_eval_lds(sm_mlp net, const double *x,
                                                             const int32_t *cells, int64_t n,
                                                             float *out) {
  __shared__ __align__(16) __bf16 hs[WAVES][32 * PT][kRow];
  __shared__ __align__(16) bf16x8 wsh[3][8 * 64];
  // the epilogue's operands: each layer's biases (double-buffered: a wave
  // may start the next layer while another still finishes this one) and
  // the output weights, so that it never waits on a global load
  __shared__ __align__(16) float bsh[2][kW];
  __shared__ __align__(16) float osh[kW];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int r = lane % 32, half = lane / 32;
  const int64_t p0 = ((int64_t)blockIdx.x * WAVES + wave) * (32 * PT);
  bool valid[PT];
#pragma unroll
  for (int pt = 0; pt < PT; ++pt) {
    const int64_t pi = p0 + 32 * pt + r;
    float xin[IN];
    {
      double q[3] = {0.0, 0.0, 0.0};
      // both halves of the wave take point r; for cells only the lower half
      // gathers the vertices and hands the scaled input to the upper one
      valid[pt] = pi < n;
      if (valid[pt] && (!CELLS || half == 0)) valid[pt] = mlp_input<IN, CELLS>(x, cells, pi, q);
      if (!valid[pt] && pi < n && half == 0) out[pi] = INFINITY;
#pragma unroll
      for (int c = 0; c < IN; ++c)
        xin[c] = valid[pt] ? (float)((q[c] - net.shift[c]) * net.scale[c]) : 0.0f;
      if (CELLS) {
#pragma unroll
        for (int c = 0; c < IN; ++c) xin[c] = __shfl(xin[c], r);
        valid[pt] = __shfl((int)valid[pt], r) != 0;
      }
    }
    for (int j0 = 0; j0 < kW / 2; j0 += 8) {
      bf16x8 v;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int f = half * (kW / 2) + j0 + j;
        float z = net.b_in[f];
#pragma unroll
        for (int c = 0; c < IN; ++c) z = fmaf(net.w_in[f * IN + c], xin[c], z);
        v[j] = (__bf16)silu_fast(z);
      }
      *reinterpret_cast<bf16x8 *>(&hs[wave][32 * pt + r][half * (kW / 2) + j0]) = v;
    }
  }
  const bf16x8 *whid = reinterpret_cast<const bf16x8 *>(net.w_hid);
  constexpr int kStep = 8 * 64;  // fragments of one k-step
  constexpr int kSteps = kW / 16;
  static_assert(kStep == 64 * WAVES, "one fragment per thread and k-step");
  float s[PT];
#pragma unroll
  for (int pt = 0; pt < PT; ++pt) s[pt] = 0.0f;
  for (int i = threadIdx.x; i < kW; i += 64 * WAVES) osh[i] = net.w_out[i];
  for (int l = 0; l < net.hidden; ++l) {
    const bf16x8 *W = whid + (size_t)l * (kW * kW / 8);
    for (int i = threadIdx.x; i < kW; i += 64 * WAVES) bsh[l & 1][i] = net.b_hid[(size_t)l * kW + i];
    // Software pipeline over the 16 k-steps.  Step ks multiplies the
    // fragments read into registers during step ks - 1, while the reads of
    // step ks + 1 are in flight; LDS holds three k-steps of weights (read
    // ks + 1, written ks + 2 from a global load issued two steps before).
    bf16x8 gq[2];
    gq[0] = W[0 * kStep + threadIdx.x];
    wsh[0][threadIdx.x] = gq[0];
    gq[1] = W[1 * kStep + threadIdx.x];
    wsh[1][threadIdx.x] = gq[1];
    gq[0] = W[2 * kStep + threadIdx.x];
    gq[1] = W[3 * kStep + threadIdx.x];
    __syncthreads();  // stages 0 and 1 and the layer's input activations are in place
    floatx16 acc[PT][8];
#pragma unroll
    for (int pt = 0; pt < PT; ++pt)
#pragma unroll
      for (int nb = 0; nb < 8; ++nb)
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[pt][nb][i] = 0.0f;
    bf16x8 wc[8], ac[PT];
#pragma unroll
    for (int nb = 0; nb < 8; ++nb) wc[nb] = wsh[0][nb * 64 + lane];
#pragma unroll
    for (int pt = 0; pt < PT; ++pt)
      ac[pt] = *reinterpret_cast<const bf16x8 *>(&hs[wave][32 * pt + r][8 * half]);
#pragma unroll
    for (int ks = 0; ks < kSteps; ++ks) {
      bf16x8 wn[8], an[PT];
      if (ks + 1 < kSteps) {
        const bf16x8 *buf = wsh[(ks + 1) % 3];
#pragma unroll
        for (int nb = 0; nb < 8; ++nb) wn[nb] = buf[nb * 64 + lane];
#pragma unroll
        for (int pt = 0; pt < PT; ++pt)
          an[pt] = *reinterpret_cast<const bf16x8 *>(
              &hs[wave][32 * pt + r][16 * (ks + 1) + 8 * half]);
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int nb = 0; nb < 8; ++nb)
#pragma unroll
        for (int pt = 0; pt < PT; ++pt)
          acc[pt][nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wc[nb], ac[pt], acc[pt][nb], 0, 0, 0);
      __builtin_amdgcn_sched_barrier(0);
      if (ks + 2 < kSteps) {
        // stage (ks + 2) % 3 last held k-step ks - 1, whose reads (step
        // ks - 2's prefetch, or the prologue's) every wave completed
        // before a barrier this wave has passed
        wsh[(ks + 2) % 3][threadIdx.x] = gq[ks & 1];
        if (ks + 4 < kSteps) gq[ks & 1] = W[(ks + 4) * kStep + threadIdx.x];
      }
      __syncthreads();
      if (ks + 1 < kSteps) {
#pragma unroll
        for (int nb = 0; nb < 8; ++nb) wc[nb] = wn[nb];
#pragma unroll
        for (int pt = 0; pt < PT; ++pt) ac[pt] = an[pt];
      }
    }
    const float *bias = bsh[l & 1];
    const bool last = l + 1 == net.hidden;
    // bias + SiLU of every accumulator; straight-line code in each branch
    // (the branch is uniform), so the LDS reads of the operands batch
    const auto act = [&](int pt, int nb, int g4, int f0) {
      const float4 b4 = *reinterpret_cast<const float4 *>(bias + f0);
      const float bb[4] = {b4.x, b4.y, b4.z, b4.w};
      bf16x4 v;
#pragma unroll
      for (int e = 0; e < 4; ++e) v[e] = (__bf16)silu_fast(acc[pt][nb][4 * g4 + e] + bb[e]);
      return v;
    };
    if (!last) {
      // every wave has finished reading this layer's input (the barrier of
      // the last k-step), so the output may overwrite it
#pragma unroll
      for (int nb = 0; nb < 8; ++nb)
#pragma unroll
        for (int g4 = 0; g4 < 4; ++g4)
#pragma unroll
          for (int pt = 0; pt < PT; ++pt) {
            const int f0 = 32 * nb + 8 * g4 + 4 * half;
            *reinterpret_cast<bf16x4 *>(&hs[wave][32 * pt + r][f0]) = act(pt, nb, g4, f0);
          }
    } else {
#pragma unroll
      for (int nb = 0; nb < 8; ++nb)
#pragma unroll
        for (int g4 = 0; g4 < 4; ++g4)
#pragma unroll
          for (int pt = 0; pt < PT; ++pt) {
            const int f0 = 32 * nb + 8 * g4 + 4 * half;
            const bf16x4 v = act(pt, nb, g4, f0);
            const float4 o4 = *reinterpret_cast<const float4 *>(&osh[f0]);
            const float oo[4] = {o4.x, o4.y, o4.z, o4.w};
#pragma unroll
            for (int e = 0; e < 4; ++e) s[pt] = fmaf((float)v[e], oo[e], s[pt]);
          }
    }
  }
  if (net.hidden == 0) {
    __syncthreads();
#pragma unroll
    for (int pt = 0; pt < PT; ++pt)
      for (int j = 0; j < kW / 2; ++j) {
        const int f = half * (kW / 2) + j;
        s[pt] = fmaf((float)hs[wave][32 * pt + r][f], net.w_out[f], s[pt]);
      }
  }
#pragma unroll
  for (int pt = 0; pt < PT; ++pt) {
    float t = s[pt] + __shfl_xor(s[pt], 32);
    const int64_t pi = p0 + 32 * pt + r;
    if (half == 0 && valid[pt]) out[pi] = (float)(net.out_scale * (double)(t + net.b_out));
  }
}

// 128 x 128 output tile per workgroup, k in slices of 8 staged in LDS,
// 8 x 8 outputs per thread.
constexpr int kGT = 128, kGK = 8, kGS = kGT + 4;

__global__ __launch_bounds__(256) void k_gemm_f32(int M, int N, int K, int kchunk, float alpha,
                                                  const float *A, int64_t sam, int64_t sak,
                                                  const float *B, int64_t sbk, int64_t sbn,
                                                  float beta, float *C, int64_t ldc,
                                                  float *work) {
  __shared__ __align__(16) float As[kGK][kGS];
  __shared__ __align__(16) float Bs[kGK][kGS];
  const int tid = threadIdx.x, tx = tid % 16, ty = tid / 16;
  const int m0 = blockIdx.y * kGT, n0 = blockIdx.x * kGT;
  const int kb = blockIdx.z * kchunk, ke = min(K, kb + kchunk);
  float acc[8][8];
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[i][j] = 0.0f;
  for (int k0 = kb; k0 < ke; k0 += kGK) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int e = tid + 256 * j;
      int kk, mm;
      if (sak == 1) {
        kk = e % kGK;
        mm = e / kGK;
      } else {
        kk = e / kGT;
        mm = e % kGT;
      }
      const int gm = m0 + mm, gk = k0 + kk;
      As[kk][mm] = (gm < M && gk < ke) ? A[gm * sam + gk * sak] : 0.0f;
      int kn, nn;
      if (sbn == 1) {
        kn = e / kGT;
        nn = e % kGT;
      } else {
        kn = e % kGK;
        nn = e / kGK;
      }
      const int gn = n0 + nn, gk2 = k0 + kn;
      Bs[kn][nn] = (gn < N && gk2 < ke) ? B[gk2 * sbk + gn * sbn] : 0.0f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kGK; ++kk) {
      const float4 a0 = *reinterpret_cast<const float4 *>(&As[kk][ty * 8]);
      const float4 a1 = *reinterpret_cast<const float4 *>(&As[kk][ty * 8 + 4]);
      const float4 b0 = *reinterpret_cast<const float4 *>(&Bs[kk][tx * 8]);
      const float4 b1 = *reinterpret_cast<const float4 *>(&Bs[kk][tx * 8 + 4]);
      const float a[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
      const float b[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
#pragma unroll
      for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int gm = m0 + ty * 8 + i, gn = n0 + tx * 8 + j;
      if (gm < M && gn < N) {
        if (work) {
          work[((int64_t)blockIdx.z * M + gm) * N + gn] = acc[i][j];
        } else {
          float *c = C + gm * ldc + gn;
          *c = alpha * acc[i][j] + (beta != 0.0f ? beta * *c : 0.0f);
        }
      }
    }
}

__global__ void k_split_reduce(int M, int N, int nsplit, const float *work, float alpha,
                               float beta, float *C, int64_t ldc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)M * N) return;
  float s = 0.0f;
  for (int z = 0; z < nsplit; ++z) s += work[(int64_t)z * M * N + i];
  float *c = C + (i / N) * ldc + i % N;
  *c = alpha * s + (beta != 0.0f ? beta * *c : 0.0f);
}

__global__ void k_bias_silu(int64_t rows, int cols, float *z, const float *b, float *a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * cols) return;
  const float v = z[i] + b[i % cols];
  z[i] = v;
  if (a) a[i] = silu(v);
}

// silu'(z) = s (1 + z (1 - s)), s = sigmoid(z)
__global__ void k_silu_backward(int64_t count, const float *z, const float *da, float *dz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float s = 1.0f / (1.0f + __expf(-z[i]));
  dz[i] = da[i] * s * (1.0f + z[i] * (1.0f - s));
}

__global__ void k_make_inputs(int64_t n, int dim, const double *x, double3 shift, double3 scale,
                              float eik_h, float *X) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double sh[3] = {shift.x, shift.y, shift.z}, sc[3] = {scale.x, scale.y, scale.z};
  float q[3] = {0.0f, 0.0f, 0.0f};
  for (int c = 0; c < dim; ++c) q[c] = (float)((x[i * dim + c] - sh[c]) * sc[c]);
  const int nb = eik_h > 0.0f ? 2 * dim + 1 : 1;
  for (int b = 0; b < nb; ++b) {
    float *row = X + ((int64_t)b * n + i) * dim;
    for (int c = 0; c < dim; ++c) {
      float v = q[c];
      if (b > 0 && c == (b - 1) / 2) v += (b % 2 == 1) ? eik_h : -eik_h;
      row[c] = v;
    }
  }
}

__global__ void k_sdf_loss(int64_t n, int dim, const float *y, const float *t, float eik_h,
                           float lam, float trunc, float *grad, float *part) {
  __shared__ float red[256];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float e = 0.0f;
  if (i < n) {
    const float inv_n = 1.0f / (float)n;
    const float dd = y[i] - t[i];
    grad[i] = 2.0f * dd * inv_n;
    e = dd * dd * inv_n;
    // a truncated target (|t| = trunc) has no unit gradient to match
    const bool eik = lam > 0.0f && !(trunc > 0.0f && fabsf(t[i]) >= trunc);
    if (lam > 0.0f && !eik) {
      for (int k = 0; k < dim; ++k) {
        grad[(2 * k + 1) * n + i] = 0.0f;
        grad[(2 * k + 2) * n + i] = 0.0f;
      }
    }
    if (eik) {
      float g[3], g2 = 0.0f;
      for (int k = 0; k < dim; ++k) {
        g[k] = (y[(2 * k + 1) * n + i] - y[(2 * k + 2) * n + i]) / (2.0f * eik_h);
        g2 += g[k] * g[k];
      }
      const float gn = sqrtf(g2), rr = gn - 1.0f;
      e += lam * rr * rr * inv_n;
      const float c = gn > 0.0f ? lam * 2.0f * rr / gn * inv_n / (2.0f * eik_h) : 0.0f;
      for (int k = 0; k < dim; ++k) {
        grad[(2 * k + 1) * n + i] = c * g[k];
        grad[(2 * k + 2) * n + i] = -c * g[k];
      }
    }
  }
  red[threadIdx.x] = e;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_adam(int64_t count, float *w, const float *g, float *m1, float *m2, float lr,
                       float b1, float b2, float eps, float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.0f - b1) * gi;
  const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  w[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// Hidden weights W[l][out][in] (FP32) -> BF16 fragments in the order the
// inference kernel reads them: index ((l * 16 + ks) * 8 + nb) * 64 + lane
// holds W[l][32 nb + lane % 32][16 ks + 8 (lane / 32) + j], j = 0..7.
__global__ void k_pack_hidden(int hidden, const float *w, bf16x8 *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per_layer = kW * kW / 8;
  if (i >= hidden * per_layer) return;
  const int l = (int)(i / per_layer);
  const int rem = (int)(i % per_layer);
  const int lane = rem % 64, nb = (rem / 64) % 8, ks = rem / 512;
  const float *src = w + (size_t)l * kW * kW + (size_t)(32 * nb + lane % 32) * kW + 16 * ks +
                     8 * (lane / 32);
  bf16x8 v;
#pragma unroll
  for (int j = 0; j < 8; ++j) v[j] = (__bf16)src[j];
  out[i] = v;
}

__global__ void k_to_bf16(int64_t count, const float *src, __bf16 *dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) dst[i] = (__bf16)src[i];
}

template <bool CELLS>
int launch_eval(const sm_mlp *net, const double *x, const int32_t *cells, int64_t n, float *d,
                hipStream_t s) {
  if (n == 0) return SM_OK;
  constexpr int P = kEvalPT, Wv = kEvalWaves;
  const unsigned g = blocks_for(n, 32 * P * Wv);
  if (net->in_dim == 2)
    hipLaunchKernelGGL((k_mlp_eval_lds<2, CELLS, P, Wv>), dim3(g), dim3(64 * Wv), 0, s, *net, x,
                       cells, n, d);
  else
    hipLaunchKernelGGL((k_mlp_eval_lds<3, CELLS, P, Wv>), dim3(g), dim3(64 * Wv), 0, s, *net, x,
                       cells, n, d);
  SM_LAUNCH_CHECK();
  return SM_OK;
}

}  // namespace

int mlp_check(const sm_mlp *net) {
  SM_CHECK(net, SM_ERR_ARG, "mlp: null network");
  SM_CHECK(net->width == kW, SM_ERR_ARG, "mlp: width must be %d (got %d)", kW, net->width);
  SM_CHECK(net->in_dim == 2 || net->in_dim == 3, SM_ERR_ARG, "mlp: in_dim must be 2 or 3");
  SM_CHECK(net->hidden >= 0 && net->hidden <= kMaxHidden, SM_ERR_ARG,
           "mlp: hidden layer count must be in [0, %d]", kMaxHidden);
  SM_CHECK(net->w_in && net->b_in && net->w_out &&
               (net->hidden == 0 || (net->w_hid && net->b_hid)),
           SM_ERR_ARG, "mlp: null weight pointer");
  SM_CHECK(((uintptr_t)net->w_hid & 15) == 0 && ((uintptr_t)net->b_hid & 15) == 0, SM_ERR_ARG,
           "mlp: w_hid and b_hid must be 16-byte aligned");
  return SM_OK;
}

int mlp_eval_points(const sm_mlp *net, const double *x, int64_t n, float *d, hipStream_t s) {
  return launch_eval<false>(net, x, nullptr, n, d, s);
}

int mlp_eval_cells(const sm_mlp *net, const int32_t *cells, int64_t ncells, const double *p,
                   float *d, hipStream_t s) {
  return launch_eval<true>(net, p, cells, ncells, d, s);
}

}  // namespace smk

using namespace smk;

extern "C" int sm_mlp_eval(const sm_mlp *net, const double *x, int64_t n, float *d,
                           void *stream) {
  int rc = mlp_check(net);
  if (rc) return rc;
  SM_CHECK(x && d && n >= 0, SM_ERR_ARG, "mlp_eval: bad argument");
  return mlp_eval_points(net, x, n, d, stream_of(stream));
}

extern "C" int sm_mlp_eval_centroids(const sm_mlp *net, const int32_t *cells, int64_t ncells,
                                     const double *p, float *d, void *stream) {
  int rc = mlp_check(net);
  if (rc) return rc;
  SM_CHECK(cells && p && d && ncells >= 0, SM_ERR_ARG, "mlp_eval_centroids: bad argument");
  return mlp_eval_cells(net, cells, ncells, p, d, stream_of(stream));
}

extern "C" int sm_gemm_f32(int M, int N, int K, float alpha, const float *A, int64_t sam,
                           int64_t sak, const float *B, int64_t sbk, int64_t sbn, float beta,
                           float *C, int64_t ldc, int nsplit, float *work, void *stream) {
  SM_CHECK(M >= 0 && N >= 0 && K >= 0 && A && B && C && ldc >= N, SM_ERR_ARG,
           "gemm_f32: bad argument");
  SM_CHECK(nsplit >= 1 && (nsplit == 1 || work), SM_ERR_ARG,
           "gemm_f32: nsplit > 1 needs a work buffer");
  if (M == 0 || N == 0) return SM_OK;
  int kchunk = (K + nsplit - 1) / nsplit;
  kchunk = ((kchunk + kGK - 1) / kGK) * kGK;
  if (kchunk == 0) kchunk = kGK;
  // slices actually used (<= nsplit); the caller's work buffer holds nsplit
  const int ns = nsplit == 1 ? 1 : (K + kchunk - 1) / kchunk;
  const int nz = ns > 0 ? ns : 1;
  hipStream_t s = stream_of(stream);
  dim3 g((N + kGT - 1) / kGT, (M + kGT - 1) / kGT, nz);
  hipLaunchKernelGGL(k_gemm_f32, g, dim3(256), 0, s, M, N, K, kchunk, alpha, A, sam, sak, B, sbk,
                     sbn, beta, C, ldc, nsplit == 1 ? nullptr : work);
  SM_LAUNCH_CHECK();
  if (nsplit > 1) {
    hipLaunchKernelGGL(k_split_reduce, dim3(blocks_for((int64_t)M * N, 256)), dim3(256), 0, s, M,
                       N, nz, work, alpha, beta, C, ldc);
    SM_LAUNCH_CHECK();
  }
  return SM_OK;
}

extern "C" int sm_mlp_bias_silu(int64_t rows, int cols, float *z, const float *b, float *a,
                                void *stream) {
  SM_CHECK(rows >= 0 && cols > 0 && z && b, SM_ERR_ARG, "mlp_bias_silu: bad argument");
  const int64_t cnt = rows * cols;
  if (cnt == 0) return SM_OK;
  hipLaunchKernelGGL(k_bias_silu, dim3(blocks_for(cnt, 256)), dim3(256), 0, stream_of(stream),
                     rows, cols, z, b, a);
  SM_LAUNCH_CHECK();
  return SM_OK;
}

extern "C" int sm_mlp_silu_backward(int64_t count, const float *z, const float *da, float *dz,
                                    void *stream) {
  SM_CHECK(count >= 0 && z && da && dz, SM_ERR_ARG, "mlp_silu_backward: bad argument");
  if (count == 0) return SM_OK;
  hipLaunchKernelGGL(k_silu_backward, dim3(blocks_for(count, 256)), dim3(256), 0,
                     stream_of(stream), count, z, da, dz);
  SM_LAUNCH_CHECK();
  return SM_OK;
}

extern "C" int sm_mlp_make_inputs(int64_t n, int dim, const double *x, const double *shift,
                                  const double *scale, float eik_h, float *X, void *stream) {
  SM_CHECK(n >= 0 && (dim == 2 || dim == 3) && x && shift && scale && X && eik_h >= 0.0f,
           SM_ERR_ARG, "mlp_make_inputs: bad argument");
  if (n == 0) return SM_OK;
  const double3 sh = {shift[0], shift[1], dim == 3 ? shift[2] : 0.0};
  const double3 sc = {scale[0], scale[1], dim == 3 ? scale[2] : 0.0};
  hipLaunchKernelGGL(k_make_inputs, dim3(blocks_for(n, 256)), dim3(256), 0, stream_of(stream), n,
                     dim, x, sh, sc, eik_h, X);
  SM_LAUNCH_CHECK();
  return SM_OK;
}

extern "C" int sm_mlp_sdf_loss(int64_t n, int dim, const float *y, const float *t, float eik_h,
                               float lam, float trunc, float *grad, float *part, void *stream) {
  SM_CHECK(n > 0 && (dim == 2 || dim == 3) && y && t && grad && part, SM_ERR_ARG,
           "mlp_sdf_loss: bad argument");
  SM_CHECK(lam == 0.0f || eik_h > 0.0f, SM_ERR_ARG, "mlp_sdf_loss: eikonal term needs eik_h > 0");
  SM_CHECK(trunc >= 0.0f, SM_ERR_ARG, "mlp_sdf_loss: negative truncation");
  hipLaunchKernelGGL(k_sdf_loss, dim3(blocks_for(n, 256)), dim3(256), 0, stream_of(stream), n,
                     dim, y, t, eik_h, lam, trunc, grad, part);
  SM_LAUNCH_CHECK();
  return SM_OK;
}

extern "C" int sm_mlp_adam(int64_t count, float *w, const float *g, float *m1, float *m2,
                           float lr, float b1, float b2, float eps, int step, void *stream) {
  SM_CHECK(count >= 0 && w && g && m1 && m2 && step >= 1, SM_ERR_ARG, "mlp_adam: bad argument");
  if (count == 0) return SM_OK;
  const float c1 = 1.0f - std::pow(b1, (float)step), c2 = 1.0f - std::pow(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(blocks_for(count, 256)), dim3(256), 0, stream_of(stream),
                     count, w, g, m1, m2, lr, b1, b2, eps, c1, c2);
  SM_LAUNCH_CHECK();
  return SM_OK;
}

extern "C" int sm_mlp_pack_hidden(int hidden, int width, const float *w, void *packed,
                                  void *stream) {
  SM_CHECK(width == kW, SM_ERR_ARG, "mlp_pack_hidden: width must be %d (got %d)", kW, width);
  SM_CHECK(hidden >= 0 && hidden <= kMaxHidden && (hidden == 0 || (w && packed)), SM_ERR_ARG,
           "mlp_pack_hidden: bad argument");
  SM_CHECK(((uintptr_t)packed & 15) == 0, SM_ERR_ARG, "mlp_pack_hidden: output not 16-byte aligned");
  const int64_t n = (int64_t)hidden * kW * kW / 8;
  if (n == 0) return SM_OK;
  hipLaunchKernelGGL(k_pack_hidden, dim3(blocks_for(n, 256)), dim3(256), 0, stream_of(stream),
                     hidden, w, reinterpret_cast<bf16x8 *>(packed));
  SM_LAUNCH_CHECK();
  return SM_OK;
}

extern "C" int sm_f32_to_bf16(int64_t count, const float *src, void *dst, void *stream) {
  SM_CHECK(count >= 0 && src && dst, SM_ERR_ARG, "f32_to_bf16: bad argument");
  if (count == 0) return SM_OK;
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks_for(count, 256)), dim3(256), 0, stream_of(stream),
                     count, src, reinterpret_cast<__bf16 *>(dst));
  SM_LAUNCH_CHECK();
  return SM_OK;
}
