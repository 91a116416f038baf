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

// Up to two independent arrays of the same length scanned by the same
// launches: blockIdx.y selects the array; per-block sums are stored array
// after array ([y][block]).
struct ScanJob {
  const int32_t *in[2];
  int32_t *out[2];
  int32_t *total[2];
};

__global__ void k_reduce(ScanJob j, int64_t n, int32_t *bsum) {
  const int y = blockIdx.y;
  const int32_t *in = j.in[y];
  int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)threadIdx.x * kItems;
  int32_t s = 0;
#pragma unroll
  for (int k = 0; k < kItems; ++k)
    if (base + k < n) s += in[base + k];
  int32_t excl;
  int32_t tot = block_excl_scan(s, excl);
  if (threadIdx.x == 0) bsum[(int64_t)y * gridDim.x + blockIdx.x] = tot;
}

// scans the tile and adds offs[y][block] (or 0); a single-block launch
// writes each array's total
__global__ void k_apply(ScanJob j, int64_t n, const int32_t *offs) {
  const int y = blockIdx.y;
  const int32_t *in = j.in[y];
  int32_t *out = j.out[y];
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
  int32_t off = offs ? offs[(int64_t)y * gridDim.x + blockIdx.x] : 0;
  int32_t run = off + excl;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (j.total[y] && threadIdx.x == 0 && gridDim.x == 1) *j.total[y] = off + tot;
}

// The tile scan with the block's offset taken as the sum of the per-block
// sums of the blocks before it, for grids of at most kTile blocks (no
// separate scan of the sums); the last block writes each array's total.
__global__ void k_apply_sums(ScanJob j, int64_t n, const int32_t *bsum) {
  const int y = blockIdx.y;
  const int32_t *bs = bsum + (int64_t)y * gridDim.x;
  int32_t part = 0;
  for (int k = threadIdx.x; k < (int)blockIdx.x; k += blockDim.x) part += bs[k];
  int32_t unused;
  const int32_t off = block_excl_scan(part, unused);
  const int32_t *in = j.in[y];
  int32_t *out = j.out[y];
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
  int32_t run = off + excl;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (j.total[y] && threadIdx.x == 0 && blockIdx.x == gridDim.x - 1) *j.total[y] = off + tot;
}

inline size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

int scan_run(const ScanJob &j, int count, int64_t n, void *ws, hipStream_t s) {
  if (n <= 0) {
    for (int y = 0; y < count; ++y)
      if (j.total[y]) {
        int rc = fill<int32_t>(j.total[y], 1, 0, s);
        if (rc) return rc;
      }
    return SM_OK;
  }
  if (n <= kTile) {
    hipLaunchKernelGGL(k_apply, dim3(1, (unsigned)count), dim3(kThreads), 0, s, j, n,
                       (const int32_t *)nullptr);
    SM_LAUNCH_CHECK();
    return SM_OK;
  }
  const int64_t nb = (n + kTile - 1) / kTile;
  int32_t *bsum = reinterpret_cast<int32_t *>(ws);
  char *next = reinterpret_cast<char *>(ws) + al256((size_t)count * nb * sizeof(int32_t));
  hipLaunchKernelGGL(k_reduce, dim3((unsigned)nb, (unsigned)count), dim3(kThreads), 0, s, j, n,
                     bsum);
  SM_LAUNCH_CHECK();
  if (nb <= kTile) {
    hipLaunchKernelGGL(k_apply_sums, dim3((unsigned)nb, (unsigned)count), dim3(kThreads), 0, s, j,
                       n, (const int32_t *)bsum);
    SM_LAUNCH_CHECK();
    return SM_OK;
  }
  ScanJob jb{};
  for (int y = 0; y < count; ++y) {
    jb.in[y] = bsum + y * nb;
    jb.out[y] = bsum + y * nb;
    jb.total[y] = j.total[y];
  }
  int rc = scan_run(jb, count, nb, next, s);
  if (rc) return rc;
  ScanJob ja = j;
  for (int y = 0; y < count; ++y) ja.total[y] = nullptr;
  hipLaunchKernelGGL(k_apply, dim3((unsigned)nb, (unsigned)count), dim3(kThreads), 0, s, ja, n,
                     (const int32_t *)bsum);
  SM_LAUNCH_CHECK();
  return SM_OK;
}

}  // namespace

size_t scan_workspace(int64_t n) {
  size_t bytes = 0;
  while (n > kTile) {
    int64_t nb = (n + kTile - 1) / kTile;
    bytes += al256((size_t)nb * sizeof(int32_t));
    n = nb;
  }
  return bytes + 256;
}

int scan_exclusive(const int32_t *in, int32_t *out, int64_t n, int32_t *total, void *ws,
                   hipStream_t s) {
  const ScanJob j{{in, nullptr}, {out, nullptr}, {total, nullptr}};
  return scan_run(j, 1, n, ws, s);
}

int scan_exclusive2(const int32_t *in0, int32_t *out0, int32_t *total0, const int32_t *in1,
                    int32_t *out1, int32_t *total1, int64_t n, void *ws, hipStream_t s) {
  const ScanJob j{{in0, in1}, {out0, out1}, {total0, total1}};
  return scan_run(j, 2, n, ws, s);
}

}  // namespace smk
