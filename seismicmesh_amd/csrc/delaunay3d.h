// Host 3D Delaunay tetrahedralisation: serial, or split into slabs built on
// worker threads (delaunay3d.cpp).
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <vector>

namespace smdt {

struct Builder3;

// insphere with the symbolic perturbation of the lifting map: > 0 inside,
// < 0 outside, never 0 for distinct points
int insphere_perturbed(const double *a, const double *b, const double *c, const double *d,
                       const double *e);
// lowest id first, second-lowest second, orientation kept
void canonical_tet(int32_t v[4]);

// Slab-parallel builder; keeps its per-thread scratch between calls.
struct Slabs3 {
  // Called on the calling thread: once per slab in slab order with that
  // slab's certified tetrahedra as soon as it is built, then once with
  // slab = -1 and every uncertified tetrahedron.  The pointer is valid only
  // during the call.
  using Emit = std::function<void(int slab, const int32_t *tets, int64_t count)>;
  struct Stats {
    int64_t certified = 0, uncertified = 0, local_points = 0;
  };

  Slabs3();
  ~Slabs3();
  int run(const double *p, int64_t n, int axis, int nslabs, double halo, int nthreads,
          const Emit &emit);

  std::vector<uint8_t> inserted;  // (n) 0 for skipped duplicates
  Stats stats;

 private:
  std::vector<std::unique_ptr<Builder3>> builders;
  std::vector<std::vector<int32_t>> members, certified, uncertified;
  std::vector<int32_t> unc_all;
  std::vector<int> core;
  std::vector<double> win_lo, win_hi;
};

}  // namespace smdt
