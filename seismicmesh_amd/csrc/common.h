// Shared device helpers for libsmgpu (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "smgpu_internal.h"

#define SM_HIP(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return sm_set_error(SM_ERR_HIP, "%s:%d %s -> %s", __FILE__, __LINE__, #call,   \
                          hipGetErrorString(e_));                                     \
  } while (0)

#define SM_CHECK(cond, code, ...)                   \
  do {                                              \
    if (!(cond)) return sm_set_error(code, __VA_ARGS__); \
  } while (0)

#define SM_LAUNCH_CHECK() SM_HIP(hipGetLastError())

namespace smk {

constexpr int kWave = 64;

inline hipStream_t stream_of(void *s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned blocks_for(int64_t n, int block) {
  int64_t b = (n + block - 1) / block;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  xcd_block() renumbers them so that XCD x processes the x-th
// contiguous share of the blocks: a kernel that gathers from neighbouring
// elements then finds them in its own XCD's L2 instead of every L2 fetching
// the whole array.  A bijection of [0, gridDim.x) for any grid size.
constexpr unsigned kNumXcd = 8;
__device__ inline unsigned xcd_block() {
  const unsigned b = blockIdx.x, nb = gridDim.x;
  const unsigned q = nb / kNumXcd, r = nb % kNumXcd, x = b % kNumXcd;
  return x * q + (x < r ? x : r) + b / kNumXcd;
}

// Fill kernel used instead of hipMemsetAsync so the sequence stays safe to
// capture in a graph.
template <typename T>
__global__ void k_fill(T *p, int64_t n, T v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// Several int32 arrays set to one value in a single launch (one graph node
// instead of one per array).
struct FillList {
  int32_t *p[6];
  int64_t n[6];
};
template <typename L>
__global__ void k_fill_list(L f, int32_t v) {
  const int k = blockIdx.y;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < f.n[k]; i += stride) f.p[k][i] = v;
}
inline int fill_list(const FillList &f, int count, int32_t v, hipStream_t s) {
  int64_t nmax = 0;
  for (int k = 0; k < count; ++k) nmax = f.n[k] > nmax ? f.n[k] : nmax;
  if (count <= 0 || nmax <= 0) return SM_OK;
  unsigned g = blocks_for(nmax, 256);
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(k_fill_list<FillList>, dim3(g, (unsigned)count), dim3(256), 0, s, f, v);
  SM_LAUNCH_CHECK();
  return SM_OK;
}

template <typename T>
inline int fill(T *p, int64_t n, T v, hipStream_t s) {
  if (n <= 0) return SM_OK;
  unsigned g = blocks_for(n, 256);
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(k_fill<T>, dim3(g), dim3(256), 0, s, p, n, v);
  SM_LAUNCH_CHECK();
  return SM_OK;
}

// numpy's minimum / maximum (np.minimum(a, b) returns b on ties)
__device__ inline double np_min(double a, double b) { return (a < b || a != a) ? a : b; }
__device__ inline double np_max(double a, double b) { return (a > b || a != a) ? a : b; }

// Exclusive scan of int32 counts; out may alias in.  total (device, may be
// null) receives the sum.  Workspace bytes from scan_workspace(n).
size_t scan_workspace(int64_t n);
int scan_exclusive(const int32_t *in, int32_t *out, int64_t n, int32_t *total, void *ws,
                   hipStream_t s);
// Two independent scans of the same length in the same launches (each as
// scan_exclusive); workspace bytes 2 * scan_workspace(n).
int scan_exclusive2(const int32_t *in0, int32_t *out0, int32_t *total0, const int32_t *in1,
                    int32_t *out1, int32_t *total1, int64_t n, void *ws, hipStream_t s);

}  // namespace smk
