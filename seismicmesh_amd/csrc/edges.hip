// Sorted unique edges and the per-vertex edge lists the force scatter uses.
//
// Replaces SeismicMesh/geometry/cpp/fast_geometry.cpp:49-77 (unique_edges:
// (min,max) per pair, std::sort, std::unique) as called by
// generation/mesh_generator.py:680-688 (_get_edges).
//
// Method: the sort key of an edge is (min, max) with min a vertex id, so the
// lexicographic sort is a counting sort on min (one bucket per vertex: a
// histogram, an exclusive scan and a scatter) followed by a sort of each
// bucket's max ids.  Buckets hold a vertex's higher neighbours (a handful in
// a 2D mesh: sorted in one thread's registers; a few dozen in 3D, where an
// edge comes once per incident cell: one wave, one id per lane); long
// buckets (arbitrary user input) are sorted by a whole workgroup in LDS
// or, beyond that, in global memory.
// Duplicates are dropped inside each sorted bucket; a second scan turns the
// per-bucket unique counts into output offsets.  For a 2D triangulation with
// neighbour links each shared edge is produced once in the first place (by
// the lower-numbered of its two cells) and the second scan is skipped.  The
// output is exactly the reference's sorted unique edge list.
#include "common.h"
#include "edges.h"

namespace smk {

namespace {

__device__ inline void pair_of(const int32_t *pairs, int64_t i, int32_t &a, int32_t &b) {
  int32_t u = pairs[2 * i], v = pairs[2 * i + 1];
  a = u < v ? u : v;
  b = u < v ? v : u;
}

// local edge j of a cell in the reference's order (mesh_generator.py:683-687)
__device__ inline void cell_edge(const int32_t *cells, int64_t c, int dim, int j, int32_t &a,
                                 int32_t &b) {
  const int32_t *t = cells + c * (dim + 1);
  int u, v;
  switch (j) {
    case 0: u = t[0]; v = t[1]; break;
    case 1: u = t[1]; v = t[2]; break;
    case 2: u = t[2]; v = t[0]; break;
    case 3: u = t[0]; v = t[3]; break;
    case 4: u = t[1]; v = t[3]; break;
    default: u = t[2]; v = t[3]; break;
  }
  a = u < v ? u : v;
  b = u < v ? v : u;
}

struct PairSrc {
  const int32_t *pairs;
  const int32_t *cells;
  const uint8_t *keep;  // optional cell mask
  // optional (2D cells only): adj[3 c + k] is the cell across the edge
  // opposite local vertex k, or < 0.  An edge shared by two kept cells is
  // then produced once, by the lower-numbered one, so every bucket is free
  // of duplicates and its offsets are the output offsets.
  const int32_t *adj;
  int dim;
  int64_t n;            // pairs, or cells
  __device__ inline int per() const { return pairs ? 1 : (dim == 2 ? 3 : 6); }
  __device__ inline bool get(int64_t i, int j, int32_t &a, int32_t &b) const {
    if (pairs) {
      pair_of(pairs, i, a, b);
      return true;
    }
    if (keep && !keep[i]) return false;
    if (adj) {
      // local edge j joins vertices j and (j + 1) % 3: opposite vertex (j + 2) % 3
      const int32_t o = adj[3 * i + (j + 2) % 3];
      if (o >= 0 && o < i && (!keep || keep[o])) return false;
    }
    cell_edge(cells, i, dim, j, a, b);
    return true;
  }
};

// Tetrahedra: with the ids sorted, s0 < s1 < s2 < s3, the six edges are
// three in bucket s0 (partners s1, s2, s3), two in bucket s1 (s2, s3) and
// one in bucket s2 (s3): three contiguous runs instead of six pairs.  The
// host tetrahedralisation hands out cells sorted by their lowest id, so
// consecutive lanes share s0: the lanes of a run of equal s0 (consecutive,
// all taking part) reserve their slots with one atomic by the run's first
// lane.  Returns false for a lane without a cell.
struct TetRun {
  int32_t s[4];
  bool ok, lead;
  int len;    // lanes in the run (leader only)
  int first;  // the run's first lane
};

__device__ inline TetRun tet_run(const PairSrc &src, int64_t i) {
  TetRun r;
  r.ok = i < src.n && (!src.keep || src.keep[i]);
  if (r.ok) {
    const int32_t *t = src.cells + 4 * i;
#pragma unroll
    for (int k = 0; k < 4; ++k) r.s[k] = t[k];
    // sorting network on four values
#define SM_CSWAP(a, b)                      \
  {                                         \
    const int32_t x = r.s[a], y = r.s[b];   \
    r.s[a] = x < y ? x : y;                 \
    r.s[b] = x < y ? y : x;                 \
  }
    SM_CSWAP(0, 1) SM_CSWAP(2, 3) SM_CSWAP(0, 2) SM_CSWAP(1, 3) SM_CSWAP(1, 2)
#undef SM_CSWAP
  } else {
    r.s[0] = r.s[1] = r.s[2] = r.s[3] = -1;
  }
  const int lane = threadIdx.x % 64;
  const int32_t prev = __shfl_up(r.s[0], 1);
  const uint64_t okm = __ballot(r.ok);
  const bool prev_ok = lane > 0 && ((okm >> (lane - 1)) & 1);
  r.lead = r.ok && !(prev_ok && prev == r.s[0]);
  const uint64_t brk = __ballot(!r.ok || r.lead);
  const uint64_t after = lane == 63 ? 0 : brk >> (lane + 1);
  r.len = after ? 1 + __builtin_ctzll(after) : 64 - lane;
  const uint64_t starts = __ballot(r.lead) & (lane == 63 ? ~0ull : ((2ull << lane) - 1));
  r.first = starts ? 63 - __builtin_clzll(starts) : lane;
  return r;
}

__global__ void k_count_tets(PairSrc src, int32_t *cnt) {
  const int64_t i = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  const TetRun r = tet_run(src, i);
  if (r.lead) atomicAdd(&cnt[r.s[0]], 3 * r.len);
  if (r.ok) {
    atomicAdd(&cnt[r.s[1]], 2);
    atomicAdd(&cnt[r.s[2]], 1);
  }
}

__global__ void k_scatter_tets(PairSrc src, const int32_t *rowptr, int32_t *cur,
                               int32_t *col) {
  const int64_t i = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  const TetRun r = tet_run(src, i);
  int32_t base = 0;
  if (r.lead) base = rowptr[r.s[0]] + atomicAdd(&cur[r.s[0]], 3 * r.len);
  base = __shfl(base, r.first);
  if (!r.ok) return;
  const int lane = threadIdx.x % 64;
  const int32_t p0 = base + 3 * (lane - r.first);
  col[p0] = r.s[1];
  col[p0 + 1] = r.s[2];
  col[p0 + 2] = r.s[3];
  const int32_t p1 = rowptr[r.s[1]] + atomicAdd(&cur[r.s[1]], 2);
  col[p1] = r.s[2];
  col[p1 + 1] = r.s[3];
  col[rowptr[r.s[2]] + atomicAdd(&cur[r.s[2]], 1)] = r.s[3];
}

// cnt2 (optional): the same edges counted by their second endpoint
__global__ void k_count(PairSrc src, int32_t *cnt, int32_t *cnt2) {
  int64_t i = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (i >= src.n) return;
  const int per = src.per();
  for (int j = 0; j < per; ++j) {
    int32_t a, b;
    if (src.get(i, j, a, b)) {
      atomicAdd(&cnt[a], 1);
      if (cnt2) atomicAdd(&cnt2[b], 1);
    }
  }
}

__global__ void k_scatter(PairSrc src, const int32_t *rowptr, int32_t *cur, int32_t *col) {
  int64_t i = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (i >= src.n) return;
  const int per = src.per();
  for (int j = 0; j < per; ++j) {
    int32_t a, b;
    if (src.get(i, j, a, b)) {
      int32_t pos = rowptr[a] + atomicAdd(&cur[a], 1);
      col[pos] = b;
    }
  }
}

constexpr int kRegSort = 16;    // buckets up to this size: per-thread sort
constexpr int kWaveSort = 64;   // up to this size: one wave, one id per lane
constexpr int kLdsSort = 4096;  // up to this size: one workgroup, LDS bitonic

// Sorts 16 register values with a bitonic network: every index is a
// compile-time constant after unrolling, so the values never leave VGPRs.
__device__ inline void sort16(int32_t (&v)[16]) {
#pragma unroll
  for (int size = 2; size <= 16; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const int j = k ^ stride;
        if (j > k) {
          const bool up = (k & size) == 0;
          int32_t a = v[k], b = v[j];
          int32_t lo = a < b ? a : b, hi = a < b ? b : a;
          v[k] = up ? lo : hi;
          v[j] = up ? hi : lo;
        }
      }
    }
  }
}

// Appends `want` lanes' values to list (count at *n) with one atomic per
// wave: the first wanting lane reserves the wave's slots, every lane takes
// the slot of its rank among the wanting lanes.
__device__ inline void wave_append(bool want, int32_t value, int32_t *list, int32_t *n) {
  const uint64_t m = __ballot(want);
  if (!m) return;
  const int lane = threadIdx.x % 64;
  const int leader = __builtin_ctzll(m);
  int32_t base = 0;
  if (lane == leader) base = atomicAdd(n, (int32_t)__popcll(m));
  base = __shfl(base, leader);
  if (want) list[base + (int32_t)__popcll(m & ((1ull << lane) - 1))] = value;
}

// Per-bucket sort + unique count for short buckets (one thread each);
// medium buckets (a 3D mesh's: each edge comes once per incident cell) are
// listed for the one-wave pass, long ones for the workgroup pass (one
// atomic per wave and list: in 3D nearly every row is listed, and one
// atomic per row on the same two counters serialises).
__global__ void k_row_sort(int64_t nrows, const int32_t *rowptr, int32_t *col, int32_t *ucnt,
                           int32_t *long_rows, int32_t *nlong, int32_t *mid_rows,
                           int32_t *nmid) {
  int64_t r = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (r == 0) ucnt[nrows] = 0;  // the scan's last input (its total slot)
  const bool row = r < nrows;
  const int32_t s = row ? rowptr[r] : 0, e = row ? rowptr[r + 1] : 0;
  const int32_t len = e - s;
  wave_append(row && len > kRegSort && len <= kWaveSort, (int32_t)r, mid_rows, nmid);
  wave_append(row && len > kWaveSort, (int32_t)r, long_rows, nlong);
  if (!row) return;
  if (len > kRegSort) {
    ucnt[r] = 0;  // filled by the wave or workgroup pass
    return;
  }
  int32_t v[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) v[k] = k < len ? col[s + k] : INT32_MAX;
  sort16(v);
  int32_t u = 0;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    if (k < len) {
      col[s + k] = v[k];
      u += (k == 0 || v[k] != v[k - 1]);
    }
  }
  ucnt[r] = u;
}

// Medium buckets (kRegSort < len <= 64), one wave each: lane k holds id k
// (INT32_MAX past the end), a bitonic network over the 64 lanes exchanges
// partners with lane shuffles, and a ballot counts the distinct ids.
// (block b of nb blocks given to this pass)
__device__ inline void wave_sort_rows(unsigned b, unsigned nb, const int32_t *mid_rows,
                                      const int32_t *nmid, const int32_t *rowptr, int32_t *col,
                                      int32_t *ucnt) {
  const int lane = threadIdx.x % 64;
  const int64_t nw = (int64_t)nb * (blockDim.x / 64);
  const int32_t n = *nmid;
  for (int64_t wv = ((int64_t)b * blockDim.x + threadIdx.x) / 64; wv < n; wv += nw) {
    const int32_t r = mid_rows[wv];
    const int32_t s = rowptr[r], len = rowptr[r + 1] - s;
    int32_t v = lane < len ? col[s + lane] : INT32_MAX;
#pragma unroll
    for (int size = 2; size <= 64; size <<= 1) {
#pragma unroll
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        const int32_t o = __shfl_xor(v, stride);
        const bool up = (lane & size) == 0;
        const bool low = (lane & stride) == 0;
        // the lower lane of a pair keeps the smaller value in an ascending
        // block, the larger in a descending one
        v = (low == up) ? (v < o ? v : o) : (v < o ? o : v);
      }
    }
    const int32_t prev = __shfl_up(v, 1);
    const bool first = lane < len && (lane == 0 || v != prev);
    const uint64_t m = __ballot(first);
    if (lane < len) col[s + lane] = v;
    if (lane == 0) ucnt[r] = (int32_t)__popcll(m);
  }
}

// Long buckets, one workgroup each: a bitonic sort in LDS (padded with
// INT32_MAX) for buckets up to kLdsSort ids (pass `big` = false, many
// workgroups), or in the global staging area for longer ones (pass `big` =
// true, launched with a single workgroup so the staging area is private).
__device__ inline void long_sort_rows(unsigned b, unsigned nb, const int32_t *long_rows,
                                      const int32_t *nlong, const int32_t *rowptr, int32_t *col,
                                      int32_t *ucnt, int32_t *stage, int big, int32_t *sh,
                                      int32_t *red) {
  const int nl = *nlong;
  for (int li = (int)b; li < nl; li += (int)nb) {
    const int32_t r = long_rows[li];
    const int32_t s = rowptr[r], e = rowptr[r + 1];
    const int32_t len = e - s;
    int32_t p2 = 1;
    while (p2 < len) p2 <<= 1;
    const bool lds = p2 <= kLdsSort;
    if (lds == (big != 0)) continue;
    int32_t *buf = lds ? sh : stage;
    for (int32_t k = threadIdx.x; k < p2; k += blockDim.x) buf[k] = k < len ? col[s + k] : INT32_MAX;
    __threadfence_block();
    __syncthreads();
    for (int32_t size = 2; size <= p2; size <<= 1) {
      for (int32_t stride = size >> 1; stride > 0; stride >>= 1) {
        for (int32_t k = threadIdx.x; k < p2; k += blockDim.x) {
          int32_t j = k ^ stride;
          if (j > k) {
            bool up = (k & size) == 0;
            int32_t x = buf[k], y = buf[j];
            if ((x > y) == up) {
              buf[k] = y;
              buf[j] = x;
            }
          }
        }
        __threadfence_block();
        __syncthreads();
      }
    }
    int32_t u = 0;
    for (int32_t k = threadIdx.x; k < len; k += blockDim.x) {
      col[s + k] = buf[k];
      u += (k == 0 || buf[k] != buf[k - 1]);
    }
    red[threadIdx.x] = u;
    __syncthreads();
    if (threadIdx.x == 0) {
      int32_t t = 0;
      for (int k = 0; k < (int)blockDim.x; ++k) t += red[k];
      ucnt[r] = t;
    }
    __syncthreads();
  }
}

// workgroups for the LDS-sized long buckets: a graded 3D mesh has
// thousands of rows over 64 ids (vertices where the size changes), and
// 64 workgroups working through them one bitonic sort at a time took
// 150-400 us per launch on the EAGE-shaped workload; idle ones exit at once
constexpr unsigned kLongBlocks = 1024;

// The bucket sorts after the per-thread pass, in one launch of 256-thread
// workgroups (the passes touch disjoint buckets): blocks [0, nwave) take
// the one-wave buckets, the next kLongBlocks the LDS-sized long buckets,
// the last one the buckets sorted in the global staging area (one
// workgroup, so the staging area is private to it).
__global__ void k_bucket_sorts(const int32_t *mid_rows, const int32_t *nmid,
                               const int32_t *long_rows, const int32_t *nlong,
                               const int32_t *rowptr, int32_t *col, int32_t *ucnt, int32_t *stage,
                               unsigned nwave) {
  __shared__ int32_t sh[kLdsSort];
  __shared__ int32_t red[256];
  const unsigned b = blockIdx.x;
  if (b < nwave)
    wave_sort_rows(b, nwave, mid_rows, nmid, rowptr, col, ucnt);
  else if (b < nwave + kLongBlocks)
    long_sort_rows(b - nwave, kLongBlocks, long_rows, nlong, rowptr, col, ucnt, stage, 0, sh, red);
  else
    long_sort_rows(0, 1, long_rows, nlong, rowptr, col, ucnt, stage, 1, sh, red);
}

__global__ void k_emit(int64_t nrows, const int32_t *rowptr, const int32_t *col,
                       const int32_t *eptr, int32_t *edges, SecondLists second) {
  int64_t r = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  int32_t s = rowptr[r], e = rowptr[r + 1];
  int32_t o = eptr[r];
  for (int32_t k = s; k < e; ++k) {
    if (k == s || col[k] != col[k - 1]) {
      const int32_t b = col[k];
      edges[2 * (int64_t)o] = (int32_t)r;
      edges[2 * (int64_t)o + 1] = b;
      if (second.list) second.list[second.ptr[b] + atomicAdd(&second.cur[b], 1)] = o;
      ++o;
    }
  }
}

// The same for the buckets of tetrahedra (dozens of sorted ids with
// duplicates): one wave per row, 64 ids at a time, each distinct id's
// output slot from a ballot of the first occurrences (grid-stride over the
// rows).
__global__ void k_emit_wave(int64_t nrows, const int32_t *rowptr, const int32_t *col,
                            const int32_t *eptr, int32_t *edges, SecondLists second) {
  const int lane = threadIdx.x % 64;
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x / 64);
  for (int64_t r = ((int64_t)xcd_block() * blockDim.x + threadIdx.x) / 64; r < nrows; r += nw) {
    const int32_t s = rowptr[r], e = rowptr[r + 1];
    int32_t o = eptr[r];
    for (int32_t k0 = s; k0 < e; k0 += 64) {
      const int32_t k = k0 + lane;
      const bool in = k < e;
      const int32_t b = in ? col[k] : 0;
      const bool first = in && (k == s || col[k - 1] != b);
      const uint64_t m = __ballot(first);
      if (first) {
        const int32_t oo = o + (int32_t)__popcll(m & ((1ull << lane) - 1));
        edges[2 * (int64_t)oo] = (int32_t)r;
        edges[2 * (int64_t)oo + 1] = b;
        if (second.list) second.list[second.ptr[b] + atomicAdd(&second.cur[b], 1)] = oo;
      }
      o += (int32_t)__popcll(m);
    }
  }
}

__global__ void k_count_to_i64(const int32_t *c, int64_t *out) { *out = (int64_t)*c; }

inline size_t al(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

size_t unique_edges_workspace(int64_t npairs, int64_t nvert) {
  size_t b = 0;
  b += al((nvert + 1) * 4);  // cnt -> rowptr
  b += al((nvert + 1) * 4);  // cur
  b += al(npairs * 4);       // col
  b += al((nvert + 1) * 4);  // ucnt -> eptr
  b += al((nvert + 1) * 4);  // long rows
  b += al((nvert + 1) * 4);  // medium rows
  b += al(8 * 4);            // counters
  b += al((size_t)npairs * 2 * 4 + 4096 * 4);  // staging for very long buckets
  b += 2 * scan_workspace(nvert + 1);
  return b;
}

int32_t *unique_edges_offsets(void *ws, int64_t npairs, int64_t nvert, bool adj) {
  // workspace order: rowptr, cur, col, eptr (see unique_edges_run); with
  // neighbour links the bucket offsets are the output offsets
  char *w = reinterpret_cast<char *>(ws);
  if (adj) return reinterpret_cast<int32_t *>(w);
  w += al((nvert + 1) * 4);
  w += al((nvert + 1) * 4);
  w += al(npairs * 4);
  return reinterpret_cast<int32_t *>(w);
}

int unique_edges_run(const int32_t *pairs, const int32_t *cells, const uint8_t *keep,
                     const int32_t *adj, const SecondLists *second, int dim, int64_t n,
                     int64_t npairs, int64_t nvert, int32_t *edges, int32_t *count32,
                     int64_t *count64, void *ws, size_t ws_bytes, hipStream_t s) {
  SM_CHECK(!adj || (cells && dim == 2), SM_ERR_ARG, "unique edges: adjacency needs 2D cells");
  SM_CHECK(!second || (adj && second->ptr && second->cur && second->list), SM_ERR_ARG,
           "unique edges: second-endpoint lists need neighbour links");
  const SecondLists sl = second ? *second : SecondLists{nullptr, nullptr, nullptr};
  SM_CHECK(ws_bytes >= unique_edges_workspace(npairs, nvert), SM_ERR_CAPACITY,
           "unique edges: workspace %zu < %zu bytes", ws_bytes,
           unique_edges_workspace(npairs, nvert));
  char *w = reinterpret_cast<char *>(ws);
  int32_t *rowptr = reinterpret_cast<int32_t *>(w); w += al((nvert + 1) * 4);
  int32_t *cur = reinterpret_cast<int32_t *>(w); w += al((nvert + 1) * 4);
  int32_t *col = reinterpret_cast<int32_t *>(w); w += al(npairs * 4);
  int32_t *eptr = reinterpret_cast<int32_t *>(w); w += al((nvert + 1) * 4);
  int32_t *longr = reinterpret_cast<int32_t *>(w); w += al((nvert + 1) * 4);
  int32_t *midr = reinterpret_cast<int32_t *>(w); w += al((nvert + 1) * 4);
  int32_t *ctr = reinterpret_cast<int32_t *>(w); w += al(8 * 4);
  int32_t *stage = reinterpret_cast<int32_t *>(w);
  w += al((size_t)npairs * 2 * 4 + 4096 * 4);
  void *sws = w;

  int rc;
  {
    FillList fl{{rowptr, cur, ctr, sl.ptr, sl.cur}, {nvert + 1, nvert + 1, 8, nvert + 1, nvert + 1}};
    if ((rc = fill_list(fl, second ? 5 : 3, 0, s))) return rc;
  }
  PairSrc src{pairs, cells, keep, adj, dim, n};
  const int B = 256;
  int32_t *total = count32 ? count32 : ctr + 3;
  // tetrahedra without neighbour links: three runs per cell
  const bool tets = cells && dim == 3 && !adj && !second;
  if (n > 0) {
    if (tets)
      hipLaunchKernelGGL(k_count_tets, dim3(blocks_for(n, B)), dim3(B), 0, s, src, rowptr);
    else
      hipLaunchKernelGGL(k_count, dim3(blocks_for(n, B)), dim3(B), 0, s, src, rowptr, sl.ptr);
    SM_LAUNCH_CHECK();
  }
  // with adjacency the buckets hold no duplicates: their offsets and total
  // are the output's
  if (second)
    rc = scan_exclusive2(rowptr, rowptr, total, sl.ptr, sl.ptr, nullptr, nvert + 1, sws, s);
  else
    rc = scan_exclusive(rowptr, rowptr, nvert + 1, adj ? total : ctr + 1, sws, s);
  if (rc) return rc;
  if (n > 0) {
    if (tets)
      hipLaunchKernelGGL(k_scatter_tets, dim3(blocks_for(n, B)), dim3(B), 0, s, src, rowptr, cur,
                         col);
    else
      hipLaunchKernelGGL(k_scatter, dim3(blocks_for(n, B)), dim3(B), 0, s, src, rowptr, cur, col);
    SM_LAUNCH_CHECK();
  }
  if (nvert > 0) {
    hipLaunchKernelGGL(k_row_sort, dim3(blocks_for(nvert, B)), dim3(B), 0, s, nvert, rowptr,
                       col, eptr, longr, ctr + 2, midr, ctr + 4);
    SM_LAUNCH_CHECK();
    // one wave per medium bucket (grid-stride over at most nvert of them),
    // then the long buckets, in one launch
    unsigned gw = blocks_for(nvert * 64, 256);
    if (gw > 8192) gw = 8192;
    hipLaunchKernelGGL(k_bucket_sorts, dim3(gw + kLongBlocks + 1), dim3(256), 0, s, midr, ctr + 4,
                       longr, ctr + 2, rowptr, col, eptr, stage, gw);
    SM_LAUNCH_CHECK();
  }
  if (!adj) {
    if (nvert == 0 && (rc = fill<int32_t>(eptr, 1, 0, s))) return rc;
    if ((rc = scan_exclusive(eptr, eptr, nvert + 1, total, sws, s))) return rc;
  }
  if (nvert > 0 && tets) {
    unsigned gw = blocks_for(nvert * 64, 256);
    if (gw > 16384) gw = 16384;
    hipLaunchKernelGGL(k_emit_wave, dim3(gw), dim3(256), 0, s, nvert, rowptr, col,
                       (const int32_t *)eptr, edges, sl);
    SM_LAUNCH_CHECK();
  } else if (nvert > 0) {
    hipLaunchKernelGGL(k_emit, dim3(blocks_for(nvert, B)), dim3(B), 0, s, nvert, rowptr, col,
                       adj ? (const int32_t *)rowptr : (const int32_t *)eptr, edges, sl);
    SM_LAUNCH_CHECK();
  }
  if (count64) {
    hipLaunchKernelGGL(k_count_to_i64, dim3(1), dim3(1), 0, s, total, count64);
    SM_LAUNCH_CHECK();
  }
  return SM_OK;
}

}  // namespace smk

extern "C" size_t sm_unique_edges_workspace(int64_t npairs, int64_t nvert) {
  return smk::unique_edges_workspace(npairs, nvert);
}

extern "C" int sm_unique_edges(const int32_t *pairs, int64_t npairs, int64_t nvert,
                               int32_t *edges, int64_t *count, void *ws, size_t ws_bytes,
                               void *stream) {
  SM_CHECK(npairs >= 0 && nvert >= 0, SM_ERR_ARG, "unique_edges: negative size");
  SM_CHECK(npairs == 0 || (pairs && edges), SM_ERR_ARG, "unique_edges: null pointer");
  SM_CHECK(count && ws, SM_ERR_ARG, "unique_edges: null count/workspace");
  return smk::unique_edges_run(pairs, nullptr, nullptr, nullptr, nullptr, 2, npairs, npairs, nvert,
                               edges, nullptr, count, ws, ws_bytes, smk::stream_of(stream));
}

extern "C" size_t sm_cell_edges_workspace(int64_t ncells, int dim, int64_t nvert) {
  return smk::unique_edges_workspace(ncells * (dim == 2 ? 3 : 6), nvert);
}

extern "C" int sm_cell_edges(const int32_t *cells, int64_t ncells, int dim, int64_t nvert,
                             int32_t *edges, int64_t *count, void *ws, size_t ws_bytes,
                             void *stream) {
  SM_CHECK(dim == 2 || dim == 3, SM_ERR_ARG, "cell_edges: dim must be 2 or 3");
  SM_CHECK(ncells >= 0 && nvert >= 0, SM_ERR_ARG, "cell_edges: negative size");
  SM_CHECK(ncells == 0 || (cells && edges), SM_ERR_ARG, "cell_edges: null pointer");
  SM_CHECK(count && ws, SM_ERR_ARG, "cell_edges: null count/workspace");
  int64_t np = ncells * (dim == 2 ? 3 : 6);
  return smk::unique_edges_run(nullptr, cells, nullptr, nullptr, nullptr, dim, ncells, np, nvert,
                               edges, nullptr, count, ws, ws_bytes, smk::stream_of(stream));
}
