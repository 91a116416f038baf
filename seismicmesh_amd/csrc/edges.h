// Unique-edge builder shared by the C entry points and the force loop.
#pragma once
#include "common.h"

namespace smk {

size_t unique_edges_workspace(int64_t npairs, int64_t nvert);

// Either `pairs` (n pairs) or `cells` (n cells of dim+1 ids, optionally masked
// by `keep`) is given; npairs bounds the number of pairs.  Writes the sorted
// unique edges, the int32 count (device, optional) and the int64 count
// (device, optional).
int unique_edges_run(const int32_t *pairs, const int32_t *cells, const uint8_t *keep, int dim,
                     int64_t n, int64_t npairs, int64_t nvert, int32_t *edges, int32_t *count32,
                     int64_t *count64, void *ws, size_t ws_bytes, hipStream_t s);

}  // namespace smk
