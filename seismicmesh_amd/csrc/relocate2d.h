// Host repair of 2D triangulations that edge flips cannot untangle.
#pragma once

#include <cstdint>

// For every inverted real triangle of the slot range [t0, t0 + nt) (tri and
// adj point at slot t0; ids are global), removes one of its vertices (the
// one that moved most since pold) from the triangulation and re-inserts it.
// Returns the number of relocations; *failed counts inverted triangles none
// of whose vertices could be relocated (left as they were; the range stays
// a consistent triangulation of the same slots).
int64_t sm_relocate2d(const double *p, const double *pold, int32_t *tri, int32_t *adj,
                      int64_t t0, int64_t nt, int64_t *failed);
