// Allocation-free variant of sm_unique_edges, for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).  Its workspace is a library-owned device
// array of SM_NOALLOC_SCRATCH_BYTES, one copy per GPU, held until the
// process ends.  In host code `scratch` only names the symbol: its address
// on each GPU comes from hipGetSymbolAddress, looked up once per device
// with that device current and kept.  Every call on one GPU shares that
// GPU's array: calls that overlap on one GPU race.  The edge kernels clear
// the counters and offsets they use before using them, so what a previous
// call left in the array does not matter.
#include <mutex>

#include "common.h"
#include "edges.h"

namespace {

__device__ __align__(256) char scratch[SM_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// the device address of `scratch` on the current GPU
int scratch_ptr(void **out) {
  int dev = -1;
  SM_HIP(hipGetDevice(&dev));
  SM_CHECK(dev >= 0 && dev < kMaxDevices, SM_ERR_ARG, "unique_edges_noalloc: device %d", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    SM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    SM_CHECK(p != nullptr, SM_ERR_HIP, "unique_edges_noalloc: no address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return SM_OK;
}

}  // namespace

extern "C" int sm_unique_edges_noalloc(const int32_t *pairs, int64_t npairs, int64_t nvert,
                                       int32_t *edges, int64_t *count, void *stream) {
  SM_CHECK(npairs >= 0 && nvert >= 0, SM_ERR_ARG, "unique_edges_noalloc: negative size");
  SM_CHECK(npairs == 0 || (pairs && edges), SM_ERR_ARG, "unique_edges_noalloc: null pointer");
  SM_CHECK(count, SM_ERR_ARG, "unique_edges_noalloc: null count");
  const size_t need = smk::unique_edges_workspace(npairs, nvert);
  SM_CHECK(need <= (size_t)SM_NOALLOC_SCRATCH_BYTES, SM_ERR_CAPACITY,
           "unique_edges_noalloc: %lld pairs over %lld vertices need %zu bytes of scratch, "
           "more than the %lld it has",
           (long long)npairs, (long long)nvert, need, (long long)SM_NOALLOC_SCRATCH_BYTES);
  void *ws = nullptr;
  const int rc = scratch_ptr(&ws);
  if (rc) return rc;
  return smk::unique_edges_run(pairs, nullptr, nullptr, nullptr, nullptr, 2, npairs, npairs,
                               nvert, edges, nullptr, count, ws, need, smk::stream_of(stream));
}
