// Exclusive prefix sums of int32 counts (reduce / scan-of-sums / apply).
// A tile is 1024 counts: 256 threads x 4, wave64 shuffles inside a wave and
// one LDS round across the 4 waves of the block.
#include "common.h"

namespace smk {

namespace {

constexpr int kThreads = 256;
constexpr int kItems = 4;
constexpr int kTile = kThreads * kItems;

__device__ inline int32_t wave_incl_scan(int32_t v) {
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    int32_t u = __shfl_up(v, off, kWave);
    if (lane >= off) v += u;
  }
  return v;
}

// Exclusive block scan of one value per thread; returns the block total.
__device__ inline int32_t block_excl_scan(int32_t v, int32_t &excl) {
  __shared__ int32_t wsum[kThreads / kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  int32_t incl = wave_incl_scan(v);
  if (lane == kWave - 1) wsum[wid] = incl;
  __syncthreads();
  int32_t base = 0, total = 0;
#pragma unroll
  for (int w = 0; w < kThreads / kWave; ++w) {
    if (w < wid) base += wsum[w];
    total += wsum[w];
  }
  excl = base + incl - v;
  __syncthreads();
  return total;
}

__global__ void k_reduce(const int32_t *in, int64_t n, int32_t *bsum) {
  int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)threadIdx.x * kItems;
  int32_t s = 0;
#pragma unroll
  for (int k = 0; k < kItems; ++k)
    if (base + k < n) s += in[base + k];
  int32_t excl;
  int32_t tot = block_excl_scan(s, excl);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// scans the tile and adds offs[block] (or 0); optionally writes the total
__global__ void k_apply(const int32_t *in, int32_t *out, int64_t n, const int32_t *offs,
                        int32_t *total) {
  int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)threadIdx.x * kItems;
  int32_t v[kItems];
  int32_t s = 0;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0;
    s += v[k];
  }
  int32_t excl;
  int32_t tot = block_excl_scan(s, excl);
  int32_t off = offs ? offs[blockIdx.x] : 0;
  int32_t run = off + excl;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (total && threadIdx.x == 0 && gridDim.x == 1) *total = off + tot;
}

}  // namespace

size_t scan_workspace(int64_t n) {
  size_t bytes = 0;
  while (n > kTile) {
    int64_t nb = (n + kTile - 1) / kTile;
    bytes += ((size_t)nb * sizeof(int32_t) + 255) & ~(size_t)255;
    n = nb;
  }
  return bytes + 256;
}

int scan_exclusive(const int32_t *in, int32_t *out, int64_t n, int32_t *total, void *ws,
                   hipStream_t s) {
  if (n <= 0) {
    if (total) return fill<int32_t>(total, 1, 0, s);
    return SM_OK;
  }
  if (n <= kTile) {
    hipLaunchKernelGGL(k_apply, dim3(1), dim3(kThreads), 0, s, in, out, n,
                       (const int32_t *)nullptr, total);
    SM_LAUNCH_CHECK();
    return SM_OK;
  }
  int64_t nb = (n + kTile - 1) / kTile;
  int32_t *bsum = reinterpret_cast<int32_t *>(ws);
  char *next = reinterpret_cast<char *>(ws) + (((size_t)nb * sizeof(int32_t) + 255) & ~(size_t)255);
  hipLaunchKernelGGL(k_reduce, dim3((unsigned)nb), dim3(kThreads), 0, s, in, n, bsum);
  SM_LAUNCH_CHECK();
  int rc = scan_exclusive(bsum, bsum, nb, total, next, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_apply, dim3((unsigned)nb), dim3(kThreads), 0, s, in, out, n,
                     (const int32_t *)bsum, (int32_t *)nullptr);
  SM_LAUNCH_CHECK();
  return SM_OK;
}

}  // namespace smk
