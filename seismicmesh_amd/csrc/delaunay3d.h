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

// The tetrahedralisation of a point set, built on one or more threads
// (delaunay3d.cpp); keeps its scratch, and the worker threads of its
// parallel builds (idle between calls, joined on destruction), between
// calls.  One build at a time per object.
struct DT3 {
  // Called on the calling thread with the finite tetrahedra (canonical, in
  // increasing order) in nslabs consecutive pieces (piece index first),
  // then once with -1 and no cells.  The pointer is valid during the call.
  using Emit = std::function<void(int piece, const int32_t *tets, int64_t count)>;
  struct Stats {
    int64_t parallel = 0;  // 1 when the build ran on several threads
    int64_t retried = 0;   // points left by the threads to the final pass
    std::vector<int64_t> slab_points;
  };

  DT3();
  ~DT3();
  // nthreads threads insert the points of nslabs slabs along `axis` into
  // one shared tetrahedralisation (one thread: the serial build).  The
  // sorted cells are written to `out` (room for out_cap cells) when it is
  // given and large enough, else to an internal buffer; the pieces point
  // into it.
  int run(const double *p, int64_t n, int axis, int nslabs, int nthreads, const Emit &emit,
          int32_t *out = nullptr, int64_t out_cap = 0);
  // Checks cells (count, 4; positively oriented) against every one of the n
  // points: 1 when no point lies inside any cell's sphere (perturbed test),
  // 0 when one does or a sphere is too wide to check here (more than
  // max_bins grid bins).  Uses nthreads threads.
  int verify(const double *p, int64_t n, const int32_t *cells, int64_t count, int nthreads,
             int64_t max_bins = 1 << 15);

  std::vector<uint8_t> inserted;  // (n) 0 for skipped duplicates
  Stats stats;

 private:
  void bin_points(const double *p, int64_t n);
  int64_t bin_of(double x, int k) const;
  int sphere_empty(const double *p, const int32_t *v, int axis, double skip_lo, double skip_hi,
                   int64_t max_bins) const;

  std::unique_ptr<Builder3> B;
  std::vector<int64_t> start_;
  std::vector<int32_t> sorted_;
  std::vector<std::vector<int32_t>> part_;  // (threads^2) cells per thread and range
  // points binned on a uniform grid (verify)
  double blo_[3] = {0, 0, 0}, bhi_[3] = {0, 0, 0};
  int64_t nb_[3] = {1, 1, 1};
  std::vector<int32_t> bin_start, bin_pts;
};

}  // namespace smdt
