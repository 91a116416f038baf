// Unique-edge builder shared by the C entry points and the force loop.
#pragma once
#include "common.h"

namespace smk {

size_t unique_edges_workspace(int64_t npairs, int64_t nvert);

// Either `pairs` (n pairs) or `cells` (n cells of dim+1 ids, optionally masked
// by `keep`) is given; npairs bounds the number of pairs.  `adj` (optional,
// 2D cells only) holds each cell's neighbour across the edge opposite local
// vertex k at 3 c + k (or < 0); the links must be mutual.  Writes the sorted
// unique edges, the int32 count (device, optional) and the int64 count
// (device, optional).
// Optional output of a run with neighbour links: for each vertex v the ids
// of the edges whose second endpoint is v, as offsets `ptr` (nvert + 1) into
// `list` (one entry per edge), in no particular order within a vertex;
// `cur` is nvert + 1 counters of scratch.
struct SecondLists {
  int32_t *ptr, *cur, *list;
};

int unique_edges_run(const int32_t *pairs, const int32_t *cells, const uint8_t *keep,
                     const int32_t *adj, const SecondLists *second, int dim, int64_t n,
                     int64_t npairs, int64_t nvert, int32_t *edges, int32_t *count32,
                     int64_t *count64, void *ws, size_t ws_bytes, hipStream_t s);

// After unique_edges_run with the same npairs and nvert: the device array
// (nvert + 1 entries) of each first endpoint's offset in the edge list, in
// the workspace `ws`; `adj` says whether the run was given neighbour links.
int32_t *unique_edges_offsets(void *ws, int64_t npairs, int64_t nvert, bool adj);

}  // namespace smk
