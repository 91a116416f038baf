// Host 3D Delaunay tetrahedralisation (incremental Bowyer-Watson, exact
// predicates), serial or split into slabs built on worker threads.
//
// Replaces the role of CGAL's Delaunay_triangulation_3 behind
// SeismicMesh/generation/cpp/delaunay_class3.cpp:116-129 (insert) and
// :276-344 (get_finite_cells / get_finite_vertices), which the 3D loop and
// sliver_removal call every iteration (mesh_generator.py:202-215, :465-469).
//
// Uniqueness.  Cospherical points (the initial lattices are full of them)
// admit several Delaunay tetrahedralisations.  The in-sphere test here is
// made total by a symbolic perturbation of the lifting map (Devillers and
// Teillaud, "Perturbations for Delaunay and weighted Delaunay 3D
// triangulations", Comput. Geom. 44, 2011 -- the scheme CGAL's
// Delaunay_triangulation_3 uses): each point's lifted coordinate |p|^2 is
// raised by an infinitesimal whose order follows the lexicographic (x, y, z)
// order of the points, the largest point carrying the dominant term.  The
// lifted determinant is linear in each lifted coordinate, so when it
// vanishes its perturbed sign is that of the first non-zero cofactor taken
// in that order, and each cofactor is an orientation of four of the five
// points.  The tetrahedralisation is then a function of the point set alone,
// whatever the insertion order, which is what lets slabs be built apart.
//
// Convex hull.  As in CGAL, each facet of the convex hull carries an
// infinite tetrahedron with a symbolic vertex at infinity: a point beyond
// the facet conflicts with it, a point in the facet's plane does when it
// lies inside the facet's circumcircle (decided, ties included, by the
// perturbed in-sphere test of the finite tetrahedron across the facet,
// whose sphere meets that plane in the same circle).  The cells are the
// Delaunay tetrahedralisation of the points' convex hull, flat hull slivers
// included.
//
// Slabs (the decomposition of mesh_generator.py:808-852 / decomp/blocker.py:
// equal-width slabs along one axis).  Slab k builds the triangulation of its
// own points, every point whose axis coordinate lies in the window
// [lo_k - halo, hi_k + halo], and the `nglobal` first points (the fixed
// points, which every slab takes).  It reports the tetrahedra it owns --
// those whose other vertices have their midpoint along the axis in slab k
// (of all four when all are fixed) -- and certifies each one whose
// circumsphere spans, along the axis, less
// than the window: every point inside or on that sphere is then in the
// window, so the tetrahedron is Delaunay for the whole point set.  By the
// uniqueness above, every tetrahedron of the whole set whose non-fixed
// vertices span at most 2 halo along the axis lies in its owner's window
// and so is found there.  Owned tetrahedra that cannot be certified are
// reported after all certified ones; the caller checks those that survive
// its centroid filter against every point (Slabs3::verify) and rebuilds
// serially if one of them is not Delaunay.  What this leaves open: a
// tetrahedron whose non-fixed vertices span more than 2 halo along the axis
// is seen by no slab; the force loop's kept cells span about one or two
// edge lengths, and the default halo is five.
//
// Output tetrahedra are positively oriented and canonical: the lowest vertex
// id first, the second-lowest next, the last two in the order that keeps
// the orientation.  Exact duplicates are skipped: of several equal points
// only the lowest id is used.
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <functional>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "delaunay3d.h"
#include "predicates.h"
#include "smgpu_internal.h"

namespace {

struct Tet {
  int v[4];
  int n[4];
};

// Growable array of tetrahedra whose new elements are left uninitialised
// (the concurrent build sizes it for its whole capacity up front and
// writes each slot before anything reads it).
class TetArray {
 public:
  size_t size() const { return n_; }
  Tet &operator[](size_t i) { return buf_[i]; }
  const Tet &operator[](size_t i) const { return buf_[i]; }
  void clear() { n_ = 0; }
  void resize(size_t m) {
    reserve(m);
    n_ = m;
  }
  void push_back(const Tet &t) {
    if (n_ == cap_) reserve(std::max<size_t>(64, 2 * cap_));
    buf_[n_++] = t;
  }
  void reserve(size_t c) {
    if (c <= cap_) return;
    std::unique_ptr<Tet[]> nb(new Tet[c]);
    if (n_) std::memcpy(nb.get(), buf_.get(), n_ * sizeof(Tet));
    buf_ = std::move(nb);
    cap_ = c;
  }

 private:
  std::unique_ptr<Tet[]> buf_;
  size_t n_ = 0, cap_ = 0;
};

// 3D Hilbert index of a point of the 1024^3 grid, one octant (3 bits) per
// level from a 24-state table: kHilbertDigit[s][o] is the curve position
// of octant o = (x bit, y bit, z bit) of the cell in state s, and
// kHilbertNext[s][o] the state of that octant's sub-curve.  The tables are
// Skilling's transpose-form Hilbert curve ("Programming the Hilbert curve",
// AIP Conf. Proc. 707, 2004) written as an automaton: each octant's
// sub-curve is the parent curve under one of the 48 signed axis
// permutations, found by matching the two-level curve, and the states are
// the compositions reached from the identity (24 of them).  The order only
// steers insertion, not the result.
static const uint8_t kHilbertNext[24][8] = {
    {1, 2, 3, 0, 4, 5, 6, 0},
    {7, 8, 9, 10, 11, 2, 1, 1},
    {6, 0, 12, 13, 14, 2, 1, 2},
    {15, 16, 3, 3, 9, 10, 17, 0},
    {18, 5, 4, 4, 15, 16, 9, 10},
    {19, 5, 4, 5, 3, 0, 20, 13},
    {9, 10, 17, 0, 7, 8, 6, 6},
    {0, 21, 13, 9, 6, 7, 12, 7},
    {22, 17, 10, 23, 8, 6, 8, 12},
    {2, 15, 1, 9, 5, 7, 4, 9},
    {16, 11, 10, 1, 8, 18, 10, 4},
    {17, 6, 23, 12, 11, 14, 11, 1},
    {23, 13, 21, 22, 12, 12, 7, 8},
    {20, 13, 14, 2, 12, 13, 19, 5},
    {21, 22, 7, 8, 14, 14, 11, 2},
    {3, 15, 20, 15, 0, 21, 13, 9},
    {16, 3, 16, 20, 22, 17, 10, 23},
    {11, 1, 17, 3, 18, 4, 17, 6},
    {18, 19, 18, 4, 17, 3, 23, 20},
    {19, 19, 18, 5, 21, 22, 15, 16},
    {20, 20, 15, 16, 23, 13, 21, 22},
    {14, 21, 2, 15, 19, 21, 5, 7},
    {22, 14, 16, 11, 22, 19, 8, 18},
    {23, 20, 11, 14, 23, 12, 18, 19},
};
static const uint8_t kHilbertDigit[24][8] = {
    {0, 1, 3, 2, 7, 6, 4, 5},
    {0, 7, 1, 6, 3, 4, 2, 5},
    {0, 1, 7, 6, 3, 2, 4, 5},
    {6, 1, 5, 2, 7, 0, 4, 3},
    {4, 3, 5, 2, 7, 0, 6, 1},
    {4, 5, 3, 2, 7, 6, 0, 1},
    {0, 7, 3, 4, 1, 6, 2, 5},
    {0, 3, 7, 4, 1, 2, 6, 5},
    {4, 7, 3, 0, 5, 6, 2, 1},
    {0, 3, 1, 2, 7, 4, 6, 5},
    {4, 7, 5, 6, 3, 0, 2, 1},
    {6, 7, 1, 0, 5, 4, 2, 3},
    {4, 3, 7, 0, 5, 2, 6, 1},
    {4, 5, 7, 6, 3, 2, 0, 1},
    {6, 1, 7, 0, 5, 2, 4, 3},
    {6, 5, 1, 2, 7, 4, 0, 3},
    {2, 1, 5, 6, 3, 0, 4, 7},
    {6, 7, 5, 4, 1, 0, 2, 3},
    {2, 3, 5, 4, 1, 0, 6, 7},
    {2, 5, 3, 4, 1, 6, 0, 7},
    {2, 5, 1, 6, 3, 4, 0, 7},
    {6, 5, 7, 4, 1, 2, 0, 3},
    {2, 1, 3, 0, 5, 6, 4, 7},
    {2, 3, 1, 0, 5, 4, 6, 7},
};

inline uint64_t hilbert3(uint32_t x, uint32_t y, uint32_t z) {
  uint64_t h = 0;
  int s = 0;
  for (int b = 9; b >= 0; --b) {
    const int o = (int)(((x >> b) & 1u) << 2 | ((y >> b) & 1u) << 1 | ((z >> b) & 1u));
    h = h << 3 | kHilbertDigit[s][o];
    s = kHilbertNext[s][o];
  }
  return h;
}

// Worker threads kept between parallel sections: run(nt, fn) calls fn(0)
// on the calling thread and fn(1 .. nt-1) on workers 1 .. nt-1 (started
// on first use), and returns when all have finished.  Idle workers sleep
// on a condition variable.  Sections do not nest.
class WorkerPool {
 public:
  WorkerPool() = default;
  WorkerPool(const WorkerPool &) = delete;
  WorkerPool &operator=(const WorkerPool &) = delete;
  ~WorkerPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    wake_.notify_all();
    for (auto &t : th_) t.join();
  }
  void run(int nt, const std::function<void(int)> &fn) {
    if (nt <= 1) {
      fn(0);
      return;
    }
    while ((int)th_.size() < nt - 1) {
      const int w = (int)th_.size() + 1;
      th_.emplace_back([this, w] { loop(w); });
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      fn_ = &fn;
      nt_ = nt;
      pending_ = nt - 1;
      ++gen_;
    }
    wake_.notify_all();
    fn(0);
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [this] { return pending_ == 0; });
    fn_ = nullptr;
  }

 private:
  void loop(int w) {
    uint64_t seen = 0;
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
      wake_.wait(lk, [&] { return stop_ || gen_ != seen; });
      if (stop_) return;
      seen = gen_;
      if (w >= nt_) continue;
      const std::function<void(int)> *f = fn_;
      lk.unlock();
      (*f)(w);
      lk.lock();
      if (--pending_ == 0) done_.notify_one();
    }
  }
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable wake_, done_;
  const std::function<void(int)> *fn_ = nullptr;
  int nt_ = 0, pending_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

// kOther[a][b]: the two slots of a tetrahedron other than a and b (a != b)
constexpr int kOther[4][4][2] = {
    {{0, 0}, {2, 3}, {1, 3}, {1, 2}},
    {{2, 3}, {0, 0}, {0, 3}, {0, 2}},
    {{1, 3}, {0, 3}, {0, 0}, {0, 1}},
    {{1, 2}, {0, 2}, {0, 1}, {0, 0}},
};

inline bool lex_greater(const double *a, const double *b) {
  if (a[0] != b[0]) return a[0] > b[0];
  if (a[1] != b[1]) return a[1] > b[1];
  return a[2] > b[2];
}

// SMGPU_TRACE_DT3=1 prints each build's phase times (stderr)
bool trace_dt3() {
  static const bool on = [] {
    const char *e = std::getenv("SMGPU_TRACE_DT3");
    return e && *e && *e != '0';
  }();
  return on;
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

}  // namespace

namespace smdt {

// > 0: e inside the sphere of the positively oriented tetrahedron abcd,
// < 0: outside; never 0 for distinct points (perturbed lifting, see above).
// d(det)/d(lift of vertex i) is the orientation of abcd with vertex i
// replaced by e; d(det)/d(lift of e) is -orient(abcd) < 0.
int insphere_perturbed(const double *a, const double *b, const double *c, const double *d,
                       const double *e) {
  int s = smpred::insphere_h(a, b, c, d, e);
  if (s != 0) return s;
  const double *pt[5] = {a, b, c, d, e};
  int ord[5] = {0, 1, 2, 3, 4};
  // decreasing lexicographic order (insertion sort of five)
  for (int i = 1; i < 5; ++i)
    for (int j = i; j > 0 && lex_greater(pt[ord[j]], pt[ord[j - 1]]); --j)
      std::swap(ord[j], ord[j - 1]);
  for (int k = 0; k < 5; ++k) {
    const int i = ord[k];
    if (i == 4) return -1;
    const double *w[4] = {a, b, c, d};
    w[i] = e;
    const int o = smpred::orient3d_h(w[0], w[1], w[2], w[3]);
    if (o != 0) return o;
  }
  return -1;
}

// Canonical vertex order of a positively oriented tetrahedron: lowest id
// first, second-lowest second, the rest by an even permutation.
void canonical_tet(int32_t v[4]) {
  int i = 0;
  for (int k = 1; k < 4; ++k)
    if (v[k] < v[i]) i = k;
  if (i != 0) {
    // the double transposition (0 i)(j k) is even
    int j = -1, k = -1;
    for (int s = 1; s < 4; ++s)
      if (s != i) (j < 0 ? j : k) = s;
    std::swap(v[0], v[i]);
    std::swap(v[j], v[k]);
  }
  // rotate positions 1..3 (a 3-cycle, even) until the smallest is at 1
  while (v[1] > v[2] || v[1] > v[3]) {
    const int t = v[1];
    v[1] = v[2];
    v[2] = v[3];
    v[3] = t;
  }
}

// Incremental Bowyer-Watson on a local point array (reused between builds).
// The triangulation covers the convex hull: every hull facet carries an
// infinite tetrahedron whose fourth vertex is the symbolic vertex kInf (the
// facet seen from outside), so that the result is the Delaunay
// tetrahedralisation of the points themselves, hull slivers included.
// A tetrahedron with kInf in slot i is positively oriented when putting a
// point beyond its hull facet into slot i gives orient3d > 0.
struct Builder3 {
  static constexpr int kInf = -1;
  std::vector<double> P;          // local points (n, 3)
  std::vector<int64_t> gid;       // local -> global id
  std::vector<double> P2;         // scratch of the renumbering
  std::vector<int64_t> gid2;
  std::vector<int64_t> dropped;   // global ids of skipped duplicates
  TetArray T;
  std::vector<int> free_slots;
  std::vector<char> mark;
  std::vector<std::pair<uint64_t, int>> key, tmpkey;
  std::vector<int> deferred;
  int last = 0;
  int nloc = 0;

  const double *pt(int v) const { return P.data() + 3 * (size_t)v; }

  static int inf_slot(const Tet &r) {
    for (int i = 0; i < 4; ++i)
      if (r.v[i] == kInf) return i;
    return -1;
  }

  // orient3d of tetrahedron r with slot i replaced by point q
  int orient_with(const Tet &r, int i, const double *q) const {
    const double *w[4];
    for (int k = 0; k < 4; ++k) w[k] = k == i ? q : pt(r.v[k]);
    return smpred::orient3d_h(w[0], w[1], w[2], w[3]);
  }

  bool conflict(int t, const double *q) const {
    const Tet &r = T[t];
    const int i = inf_slot(r);
    if (i < 0) return insphere_perturbed(pt(r.v[0]), pt(r.v[1]), pt(r.v[2]), pt(r.v[3]), q) > 0;
    // beyond the hull facet: in conflict; behind it: not; in its plane: as
    // for the finite tetrahedron across the facet (whose sphere meets the
    // plane in the facet's circumcircle), perturbation included
    const int o = orient_with(r, i, q);
    if (o != 0) return o > 0;
    const Tet &f = T[r.n[i]];
    return insphere_perturbed(pt(f.v[0]), pt(f.v[1]), pt(f.v[2]), pt(f.v[3]), q) > 0;
  }

  int new_tet() {
    if (!free_slots.empty()) {
      int t = free_slots.back();
      free_slots.pop_back();
      return t;
    }
    T.push_back(Tet{});
    return (int)T.size() - 1;
  }

  // visibility walk: a tetrahedron in conflict with q (the finite one
  // containing it, or an infinite one whose facet q lies beyond); -1 (with
  // *dup set) when q duplicates a vertex, -2 on failure
  int locate(int qi, int *dup) {
    const double *q = pt(qi);
    int t = last;
    {
      const int i = inf_slot(T[t]);
      if (i >= 0) t = T[t].n[i];
    }
    for (size_t steps = 0; steps < 8 * T.size() + 64; ++steps) {
      const Tet &r = T[t];
      int moved = -1;
      for (int i = 0; i < 4 && moved < 0; ++i)
        if (orient_with(r, i, q) < 0) moved = i;
      if (moved < 0) {
        for (int i = 0; i < 4; ++i) {
          const double *V = pt(r.v[i]);
          if (V[0] == q[0] && V[1] == q[1] && V[2] == q[2]) {
            *dup = r.v[i];
            return -1;
          }
        }
        return t;
      }
      const int nb = r.n[moved];
      if (nb < 0) return -2;
      if (inf_slot(T[nb]) >= 0) return nb;  // q lies beyond this hull facet
      t = nb;
    }
    return -2;
  }

  // per-insertion scratch, reused across insertions
  struct BF {
    int t, i, out, back;  // cavity tet, its face, the outside tet, its face
  };
  std::vector<int> cav, stack, ids, outside;
  std::vector<BF> bnd;
  std::vector<Tet> nt;
  // open-addressing table of the new tetrahedra's faces through q, keyed by
  // the face's edge opposite q; slots hold (key, tet, local face)
  struct Slot {
    uint64_t key;
    int t, j;
  };
  std::vector<Slot> table;
  std::vector<int> used;

  // 1 inserted, 0 duplicate, < 0 failure
  int insert(int qi) {
    int dup = -1;
    int t0 = locate(qi, &dup);
    if (t0 == -1) {
      // keep the lowest global id on the vertex that is already in place
      const int64_t keep = std::min(gid[dup], gid[qi]), drop = std::max(gid[dup], gid[qi]);
      gid[dup] = keep;
      dropped.push_back(drop);
      return 0;
    }
    if (t0 < 0) return SM_ERR_NUMERIC;
    const double *q = pt(qi);
    cav.clear();
    bnd.clear();
    outside.clear();
    stack.assign(1, t0);
    if ((size_t)T.size() > mark.size()) mark.resize(T.size() * 2, 0);
    // mark: 1 in the cavity, 2 tested and outside it (a tetrahedron next to
    // several cavity tetrahedra is tested once)
    mark[t0] = 1;
    while (!stack.empty()) {
      int t = stack.back();
      stack.pop_back();
      cav.push_back(t);
      for (int i = 0; i < 4; ++i) {
        int nb = T[t].n[i];
        if (nb >= 0 && mark[nb] == 1) continue;
        bool in = false;
        if (nb >= 0 && mark[nb] == 0) {
          in = conflict(nb, q);
          mark[nb] = in ? 1 : 2;
          if (in) stack.push_back(nb);
          else outside.push_back(nb);
        }
        if (!in) {
          // the outside tet's face back to t, found before any relinking
          int back = -1;
          if (nb >= 0)
            for (int j = 0; j < 4; ++j)
              if (T[nb].n[j] == t) back = j;
          bnd.push_back({t, i, nb, back});
        }
      }
    }
    // new tetrahedra: the cavity tetrahedron with the boundary face's
    // opposite vertex replaced by q (same orientation)
    const size_t nb = bnd.size(), nc = cav.size();
    nt.resize(nb);
    for (size_t k = 0; k < nb; ++k) {
      Tet r = T[bnd[k].t];
      r.v[bnd[k].i] = qi;
      for (int j = 0; j < 4; ++j) r.n[j] = -1;
      r.n[bnd[k].i] = bnd[k].out;
      nt[k] = r;
    }
    for (int t : cav) mark[t] = 0;
    for (int t : outside) mark[t] = 0;
    ids.resize(nb);
    for (size_t k = 0; k < nb && k < nc; ++k) ids[k] = cav[k];
    for (size_t k = nc; k < nb; ++k) ids[k] = new_tet();
    for (size_t k = nb; k < nc; ++k) free_slots.push_back(cav[k]);
    if ((size_t)T.size() > mark.size()) mark.resize(T.size() * 2, 0);
    for (size_t k = 0; k < nb; ++k) T[ids[k]] = nt[k];
    // outside neighbours point back to the new tetrahedra
    for (size_t k = 0; k < nb; ++k)
      if (bnd[k].out >= 0) T[bnd[k].out].n[bnd[k].back] = ids[k];
    link_faces_through(qi, nb, [&](size_t k) { return bnd[k].i; });
    last = ids[0];
    if ((size_t)qi < vtet.size()) vtet[qi] = ids[0];
    return 1;
  }

  // Links the new tetrahedra ids[0..nb) to each other through their faces
  // containing vertex x (in slot xs(k) of tetrahedron k): the face opposite
  // slot j != xs(k) is {x} + the edge of the other two slots, and each such
  // edge is shared by exactly two of the new tetrahedra.
  template <class SlotOf>
  void link_faces_through(int x, size_t nb, SlotOf xs) {
    (void)x;
    size_t cap = 64;
    while (cap < 4 * 3 * nb) cap <<= 1;
    if (table.size() < cap) table.assign(cap, Slot{~0ull, -1, -1});
    const size_t mask = cap - 1;
    used.clear();
    for (size_t k = 0; k < nb; ++k) {
      const Tet &r = T[ids[k]];
      const int qs = xs(k);
      for (int j = 0; j < 4; ++j) {
        if (j == qs) continue;
        const int s0 = kOther[qs][j][0], s1 = kOther[qs][j][1];
        const uint32_t e0 = (uint32_t)r.v[s0], e1 = (uint32_t)r.v[s1];
        const uint64_t key = ((uint64_t)std::min(e0, e1) << 32) | std::max(e0, e1);
        size_t h = (size_t)((key * 0x9E3779B97F4A7C15ull) >> 32) & mask;
        while (table[h].key != ~0ull && table[h].key != key) h = (h + 1) & mask;
        if (table[h].key == key) {
          const int ot = table[h].t, oj = table[h].j;
          T[ids[k]].n[j] = ot;
          T[ot].n[oj] = ids[k];
        } else {
          table[h] = Slot{key, ids[k], j};
          used.push_back((int)h);
        }
      }
    }
    for (int h : used) table[h].key = ~0ull;
  }

  // The first tetrahedron: four non-coplanar points of the insertion order
  // and the four infinite tetrahedra on its faces.  Points passed over
  // (duplicates of the first, collinear or coplanar with the first ones)
  // are inserted afterwards.  Returns the number of order entries consumed,
  // or 0 when every point lies in one plane.
  size_t start(const std::vector<std::pair<uint64_t, int>> &order) {
    const size_t n = order.size();
    deferred.clear();
    if (n < 4) return 0;
    const int a = order[0].second;
    int b = -1, c = -1, d = -1;
    size_t k = 1;
    for (; k < n && d < 0; ++k) {
      const int v = order[k].second;
      const double *V = pt(v), *A = pt(a);
      if (b < 0) {
        if (V[0] != A[0] || V[1] != A[1] || V[2] != A[2]) b = v;
        else deferred.push_back(v);
      } else if (c < 0) {
        // not collinear with a, b: some 2D projection turns
        const double *B = pt(b);
        const double xy[3][2] = {{A[0], A[1]}, {B[0], B[1]}, {V[0], V[1]}};
        const double yz[3][2] = {{A[1], A[2]}, {B[1], B[2]}, {V[1], V[2]}};
        const double zx[3][2] = {{A[2], A[0]}, {B[2], B[0]}, {V[2], V[0]}};
        if (smpred::orient2d(xy[0], xy[1], xy[2]) != 0 ||
            smpred::orient2d(yz[0], yz[1], yz[2]) != 0 ||
            smpred::orient2d(zx[0], zx[1], zx[2]) != 0)
          c = v;
        else
          deferred.push_back(v);
      } else {
        if (smpred::orient3d_h(A, pt(b), pt(c), V) != 0) d = v;
        else deferred.push_back(v);
      }
    }
    if (d < 0) return 0;
    Tet f{{a, b, c, d}, {-1, -1, -1, -1}};
    if (smpred::orient3d_h(pt(a), pt(b), pt(c), pt(d)) < 0) std::swap(f.v[2], f.v[3]);
    T.push_back(f);
    // infinite tetrahedron on the face opposite slot i: slot i becomes kInf
    // and two other slots swap, so that the finite vertex there (inside)
    // gives a negative orientation
    for (int i = 0; i < 4; ++i) {
      Tet u = f;
      u.v[i] = kInf;
      const int s0 = kOther[i][(i + 1) % 4][0], s1 = kOther[i][(i + 1) % 4][1];
      std::swap(u.v[s0], u.v[s1]);
      for (int j = 0; j < 4; ++j) u.n[j] = -1;
      u.n[i] = 0;
      T.push_back(u);
      T[0].n[i] = 1 + i;
    }
    ids.assign({1, 2, 3, 4});
    link_faces_through(kInf, 4, [&](size_t k) { return inf_slot(T[ids[k]]); });
    last = 0;
    return k;
  }

  // fn(0 .. nt-1), on nt threads (this one included), on the pool's
  // workers, which are kept between sections and builds (creating 15
  // threads costs about 0.5 ms; a build has a dozen sections)
  WorkerPool pool;
  template <class Fn>
  void on_threads(int nt, Fn &&fn) {
    pool.run(nt, std::function<void(int)>(std::forward<Fn>(fn)));
  }

  // Triangulates the nloc points of P in Hilbert order (on nthreads
  // threads, in nslabs slabs along `axis`, when that is more than one).
  int build(const double lo[3], double ext, int nthreads = 1, int axis = 1, int nslabs = 1) {
    T.clear();
    free_slots.clear();
    dropped.clear();
    const int n = nloc;
    last = 0;
    if (mark.size() < 64) mark.assign(64, 0);
    // Hilbert keys (30 bits) on the threads, then a stable three-pass radix
    // sort on the threads (10-bit digits; per-thread histograms, each
    // thread scattering its own range): the order of std::sort on (key, id)
    for (double &x : ph) x = 0.0;
    double t0 = now_ms();
    key.resize(n);
    const int kt = std::max(1, std::min(nthreads, n / 8192 + 1));
    auto part = [&](int w) { return (int64_t)n * w / kt; };
    on_threads(kt, [&](int w) {
      for (int64_t i = part(w); i < part(w + 1); ++i) {
        uint32_t g[3];
        for (int k = 0; k < 3; ++k)
          g[k] = (uint32_t)std::min(1023.0, std::max(0.0, (pt((int)i)[k] - lo[k]) / ext * 1023.0));
        key[(size_t)i] = {hilbert3(g[0], g[1], g[2]), (int)i};
      }
    });
    {
      constexpr int kBits = 10, kDigits = 1 << kBits;
      tmpkey.resize((size_t)n);
      std::vector<std::pair<uint64_t, int>> &tmp = tmpkey;
      std::vector<int64_t> cnt((size_t)kt * kDigits);
      for (int pass = 0; pass < 3; ++pass) {
        const int sh = kBits * pass;
        on_threads(kt, [&](int w) {
          int64_t *c = cnt.data() + (size_t)w * kDigits;
          std::fill(c, c + kDigits, 0);
          for (int64_t i = part(w); i < part(w + 1); ++i) ++c[(key[(size_t)i].first >> sh) & (kDigits - 1)];
        });
        // digit-major, thread-minor offsets keep the sort stable
        int64_t run = 0;
        for (int d = 0; d < kDigits; ++d)
          for (int w = 0; w < kt; ++w) {
            int64_t &c = cnt[(size_t)w * kDigits + d];
            const int64_t x = c;
            c = run;
            run += x;
          }
        on_threads(kt, [&](int w) {
          int64_t *c = cnt.data() + (size_t)w * kDigits;
          for (int64_t i = part(w); i < part(w + 1); ++i)
            tmp[(size_t)c[(key[(size_t)i].first >> sh) & (kDigits - 1)]++] = key[(size_t)i];
        });
        key.swap(tmp);
      }
    }
    // local ids renumbered in curve order, so that the coordinates a walk
    // or a cavity reads lie together in memory (the perturbation orders
    // points by their coordinates, not their ids: the result is the same)
    {
      P2.resize(P.size());
      gid2.resize(gid.size());
      on_threads(kt, [&](int w) {
        for (int64_t k = part(w); k < part(w + 1); ++k) {
          const int i = key[(size_t)k].second;
          for (int j = 0; j < 3; ++j) P2[3 * (size_t)k + j] = P[3 * (size_t)i + j];
          gid2[(size_t)k] = gid[(size_t)i];
          key[(size_t)k].second = (int)k;
        }
      });
      P.swap(P2);
      gid.swap(gid2);
    }
    ph[0] = now_ms() - t0;
    const size_t k0 = start(key);
    if (k0 == 0) {
      T.clear();
      return SM_OK;  // fewer than four points off one plane: no tetrahedra
    }
    retried = 0;
    slab_sizes.clear();
    if (nthreads > 1 && n >= kParallelMin)
      return build_parallel(key, k0, nthreads, axis, nslabs);
    // coarse to fine, as the concurrent build: every 4096th point of the
    // curve, then every 64th, then the rest (shorter walks, fewer cache
    // misses than one sweep along the curve)
    for (const size_t stride : {(size_t)4096, (size_t)64, (size_t)1})
      for (size_t k = k0; k < key.size(); ++k) {
        const size_t j = k - k0;
        if (j % stride != 0 || (stride < 4096 && j % (stride * 64) == 0)) continue;
        int rc = insert(key[k].second);
        if (rc < 0) return rc;
      }
    for (int v : std::vector<int>(deferred)) {
      int rc = insert(v);
      if (rc < 0) return rc;
    }
    return SM_OK;
  }

  // ------------------------------------------------------------------
  // Concurrent insertion into one shared triangulation.  Every thread
  // inserts its own points; a tetrahedron (its vertices, links and cavity
  // mark) is read or written only by the thread holding its lock, taken
  // with a try-lock while walking to the point and growing its cavity.  A
  // thread that finds a lock taken lets go of all its locks and tries that
  // point again later, so no thread ever waits on another.  Since the
  // tetrahedralisation is a function of the point set (the perturbation
  // above), the result is the serial one whatever the interleaving.  The
  // array of tetrahedra has a fixed capacity during the parallel phase (new
  // slots from an atomic counter; freed slots, marked dead, reused by the
  // thread that freed them when it can lock them), and a walk that starts
  // from a dead slot (a stale hint) starts again from a live one.
  static constexpr int kDead = -2;
  static constexpr int kParallelMin = 4096;
  // fresh slots per block: 1024 (32 KiB of tetrahedra), starting on a
  // 64-slot boundary, so that no two threads write the same cache line of
  // a block and each thread's new tetrahedra stay together (at 590 k points
  // on 16 threads: 0.224 s with blocks of 64, 0.20 s with 256, 0.177 s
  // with 1024)
  static constexpr int kBlock = 1024;
  std::vector<int64_t> slab_sizes;  // points per slab of the last parallel build
  std::vector<int> vtet;            // a tetrahedron made when each point went in
  int64_t retried = 0;              // points left to the single-thread pass
  double ph[8] = {};                // phase times of the last build (ms), traced
  struct TCtx {
    uint32_t me = 0;
    std::vector<int> held, cav, stack, ids, freel, retry, retry_pred, recent;
    std::vector<BF> bnd;
    std::vector<Tet> nt;
    std::vector<Slot> table;
    std::vector<int> used;
    int last = 0;
    int64_t blk = 0, blk_end = 0;  // this thread's block of fresh slots
    std::vector<int> skipped;      // fresh slots passed over (found locked)
  };
  std::atomic<uint32_t> *locks = nullptr;  // one per slot of the capacity
  std::unique_ptr<std::atomic<uint32_t>[]> lockbuf;  // kept between builds, all 0
  size_t lockcap = 0;
  std::atomic<int64_t> *ntet = nullptr;
  int64_t cap_par = 0;
  std::mutex *dup_mu = nullptr;

  // a lock word holds its owner (me << 2) and, while held, the cavity mark
  // of the tetrahedron (low two bits), so both share one cache line
  bool trylock(TCtx &c, int t) {
    std::atomic<uint32_t> &l = locks[t];
    uint32_t cur = l.load(std::memory_order_relaxed);
    if ((cur >> 2) == c.me) return true;
    if (cur != 0) return false;
    if (!l.compare_exchange_strong(cur, c.me << 2, std::memory_order_acquire)) return false;
    c.held.push_back(t);
    return true;
  }
  int pmark(int t) const { return (int)(locks[t].load(std::memory_order_relaxed) & 3u); }
  void set_pmark(const TCtx &c, int t, int m) {
    locks[t].store((c.me << 2) | (uint32_t)m, std::memory_order_relaxed);
  }
  // lets go of every lock of c but the one on t
  void keep_only(TCtx &c, int t) {
    for (int h : c.held)
      if (h != t) locks[h].store(0, std::memory_order_release);
    c.held.clear();
    c.held.push_back(t);
  }
  void unlock_all(TCtx &c) {
    for (int t : c.held) locks[t].store(0, std::memory_order_release);
    c.held.clear();
  }
  int par_new_slot(TCtx &c) {
    while (!c.freel.empty()) {
      const int t = c.freel.back();
      c.freel.pop_back();
      // a stale hint of another thread may hold it: then it stays dead
      if (trylock(c, t) && T[t].v[0] == kDead) return t;
    }
    // fresh slots come in blocks, so that threads do not share the
    // cache lines of what they write; nobody knows of a fresh slot
    // before it is handed out, so its lock is free
    for (;;) {
      if (c.blk >= c.blk_end) {
        c.blk = ntet->fetch_add(kBlock);
        if (c.blk >= cap_par) return -1;
        c.blk_end = std::min<int64_t>(c.blk + kBlock, cap_par);
      }
      const int t = (int)c.blk++;
      if (trylock(c, t)) return t;
      c.skipped.push_back(t);
    }
  }
  bool live_locked(TCtx &c, int t) {
    return t >= 0 && t < ntet->load(std::memory_order_relaxed) && t < cap_par &&
           trylock(c, t) && T[t].v[0] != kDead;
  }

  // 1 inserted, 0 duplicate, 2 retry later (a lock was taken), < 0 failure
  int par_insert(TCtx &c, int qi) {
    const double *q = pt(qi);
    int t = c.last;
    if (!live_locked(c, t)) {
      t = -1;
      while (!c.recent.empty() && t < 0) {
        if (live_locked(c, c.recent.back())) t = c.recent.back();
        else c.recent.pop_back();
      }
      if (t < 0) {
        unlock_all(c);
        return 2;
      }
    }
    // walk, hand over hand: the next tetrahedron is locked before the
    // current one is let go
    keep_only(c, t);
    int t0 = -1;
    for (size_t steps = 0; steps < 1000000; ++steps) {
      const Tet &r = T[t];
      const int is = inf_slot(r);
      if (is >= 0) {
        const int nb = r.n[is];
        if (!trylock(c, nb)) {
          unlock_all(c);
          return 2;
        }
        keep_only(c, nb);
        t = nb;
        continue;
      }
      int moved = -1;
      for (int i = 0; i < 4 && moved < 0; ++i)
        if (orient_with(r, i, q) < 0) moved = i;
      if (moved < 0) {
        for (int i = 0; i < 4; ++i) {
          const double *V = pt(r.v[i]);
          if (V[0] == q[0] && V[1] == q[1] && V[2] == q[2]) {
            const int dup = r.v[i];
            {
              std::lock_guard<std::mutex> lk(*dup_mu);
              const int64_t keep = std::min(gid[dup], gid[qi]);
              const int64_t drop = std::max(gid[dup], gid[qi]);
              gid[dup] = keep;
              dropped.push_back(drop);
            }
            c.last = t;
            unlock_all(c);
            return 0;
          }
        }
        t0 = t;
        break;
      }
      const int nb = r.n[moved];
      if (nb < 0) {
        unlock_all(c);
        return SM_ERR_NUMERIC;
      }
      if (!trylock(c, nb)) {
        unlock_all(c);
        return 2;
      }
      if (inf_slot(T[nb]) >= 0) {
        t0 = nb;  // q lies beyond this hull facet
        break;
      }
      keep_only(c, nb);
      t = nb;
    }
    if (t0 < 0) {
      unlock_all(c);
      return SM_ERR_NUMERIC;
    }
    // cavity: every neighbour looked at is locked first (its mark, links
    // and vertices are then this thread's), and so is the finite
    // tetrahedron behind a tested infinite one
    c.cav.clear();
    c.bnd.clear();
    c.stack.assign(1, t0);
    set_pmark(c, t0, 1);
    while (!c.stack.empty()) {
      const int tt = c.stack.back();
      c.stack.pop_back();
      c.cav.push_back(tt);
      for (int i = 0; i < 4; ++i) {
        const int nb = T[tt].n[i];
        if (nb >= 0 && !trylock(c, nb)) {
          unlock_all(c);
          return 2;
        }
        if (nb >= 0 && pmark(nb) == 1) continue;
        bool in = false;
        if (nb >= 0 && pmark(nb) == 0) {
          const int is = inf_slot(T[nb]);
          if (is >= 0 && !trylock(c, T[nb].n[is])) {
            unlock_all(c);
            return 2;
          }
          in = conflict(nb, q);
          set_pmark(c, nb, in ? 1 : 2);
          if (in) c.stack.push_back(nb);
        }
        if (!in) {
          int back = -1;
          if (nb >= 0)
            for (int j = 0; j < 4; ++j)
              if (T[nb].n[j] == tt) back = j;
          c.bnd.push_back({tt, i, nb, back});
        }
      }
    }
    const size_t nbd = c.bnd.size(), nc = c.cav.size();
    c.nt.resize(nbd);
    for (size_t k = 0; k < nbd; ++k) {
      Tet r = T[c.bnd[k].t];
      r.v[c.bnd[k].i] = qi;
      for (int j = 0; j < 4; ++j) r.n[j] = -1;
      r.n[c.bnd[k].i] = c.bnd[k].out;
      c.nt[k] = r;
    }
    c.ids.resize(nbd);
    for (size_t k = 0; k < nbd && k < nc; ++k) c.ids[k] = c.cav[k];
    for (size_t k = nc; k < nbd; ++k) {
      const int s2 = par_new_slot(c);
      if (s2 < 0) {
        unlock_all(c);
        return SM_ERR_CAPACITY;
      }
      c.ids[k] = s2;
    }
    for (size_t k = nbd; k < nc; ++k) {
      T[c.cav[k]].v[0] = kDead;
      c.freel.push_back(c.cav[k]);
    }
    for (size_t k = 0; k < nbd; ++k) T[c.ids[k]] = c.nt[k];
    for (size_t k = 0; k < nbd; ++k)
      if (c.bnd[k].out >= 0) T[c.bnd[k].out].n[c.bnd[k].back] = c.ids[k];
    // link the new tetrahedra through their faces containing q
    size_t tcap = 64;
    while (tcap < 4 * 3 * nbd) tcap <<= 1;
    if (c.table.size() < tcap) c.table.assign(tcap, Slot{~0ull, -1, -1});
    const size_t mask = tcap - 1;
    c.used.clear();
    for (size_t k = 0; k < nbd; ++k) {
      const Tet &r = T[c.ids[k]];
      const int qs = c.bnd[k].i;
      for (int j = 0; j < 4; ++j) {
        if (j == qs) continue;
        const int s0 = kOther[qs][j][0], s1 = kOther[qs][j][1];
        const uint32_t e0 = (uint32_t)r.v[s0], e1 = (uint32_t)r.v[s1];
        const uint64_t key2 = ((uint64_t)std::min(e0, e1) << 32) | std::max(e0, e1);
        size_t h = (size_t)((key2 * 0x9E3779B97F4A7C15ull) >> 32) & mask;
        while (c.table[h].key != ~0ull && c.table[h].key != key2) h = (h + 1) & mask;
        if (c.table[h].key == key2) {
          const int ot = c.table[h].t, oj = c.table[h].j;
          T[c.ids[k]].n[j] = ot;
          T[ot].n[oj] = c.ids[k];
        } else {
          c.table[h] = Slot{key2, c.ids[k], j};
          c.used.push_back((int)h);
        }
      }
    }
    for (int h : c.used) c.table[h].key = ~0ull;
    c.last = c.ids[0];
    __atomic_store_n(&vtet[(size_t)qi], c.ids[0], __ATOMIC_RELAXED);  // read by other threads
    c.recent.push_back(c.ids[0]);
    if (c.recent.size() > 64) c.recent.erase(c.recent.begin(), c.recent.begin() + 32);
    unlock_all(c);
    return 1;
  }

  // Inserts the points order[k0..] on nthreads threads, coarse to fine:
  // every 4096th point on this thread, then every 64th and then the rest
  // concurrently, then, on this thread alone, those that met a taken lock
  // twice.  In each concurrent pass the points are cut into nslabs
  // equal-width slabs along `axis` (each in curve order) and every slab
  // into chunks; the chunks are handed out slab after slab in turn (chunk 0
  // of every slab, then chunk 1, ...), so that the threads insert in
  // different slabs at any time, and each chunk starts its walk at the
  // tetrahedron made by the coarser point before it on the curve.
  int build_parallel(const std::vector<std::pair<uint64_t, int>> &order, size_t k0,
                     int nthreads, int axis, int nslabs) {
    const size_t n = order.size();
    std::vector<int> lvl0, lvl1, pred1, rest, pred2;
    vtet.assign((size_t)nloc, -1);
    int last0 = -1, last1 = -1;
    for (size_t k = k0; k < n; ++k) {
      const int v = order[k].second;
      if ((k - k0) % 4096 == 0) {
        lvl0.push_back(v);
        last0 = last1 = v;
      } else if ((k - k0) % 64 == 0) {
        lvl1.push_back(v);
        pred1.push_back(last0);
        last1 = v;
      } else {
        rest.push_back(v);
        pred2.push_back(last1);
      }
    }
    double xlo = std::numeric_limits<double>::infinity(), xhi = -xlo;
    for (size_t k = k0; k < n; ++k) {
      xlo = std::min(xlo, pt(order[k].second)[axis]);
      xhi = std::max(xhi, pt(order[k].second)[axis]);
    }
    double t0 = now_ms();
    for (int v : lvl0) {
      int rc = insert(v);
      if (rc < 0) return rc;
    }
    for (int v : std::vector<int>(deferred)) {
      int rc = insert(v);
      if (rc < 0) return rc;
    }
    deferred.clear();
    ph[1] = now_ms() - t0;
    t0 = now_ms();
    // fixed capacity, fresh slots left unwritten until handed out; the
    // dead slots of the serial phase go to thread 0
    cap_par = std::max<int64_t>((int64_t)T.size() + 128, 8 * (int64_t)nloc + 4096) +
              (int64_t)std::max(1, nthreads) * kBlock;
    if (cap_par > (int64_t)std::numeric_limits<int>::max())
      return sm_set_error(SM_ERR_CAPACITY, "host 3D Delaunay: too many points");
    std::atomic<int64_t> nt_count{((int64_t)T.size() + 63) / 64 * 64};
    const size_t old_n = T.size();
    T.resize((size_t)cap_par);
    // the slots skipped to reach the boundary are never handed out
    for (int64_t t = (int64_t)old_n; t < nt_count.load(); ++t) T[(size_t)t].v[0] = kDead;
    for (int f : free_slots) T[f].v[0] = kDead;
    // (the concurrent passes keep their cavity marks in the lock words)
    if (lockcap < (size_t)cap_par) {
      lockbuf.reset(new std::atomic<uint32_t>[(size_t)cap_par]);
      for (size_t i = 0; i < (size_t)cap_par; ++i) lockbuf[i].store(0, std::memory_order_relaxed);
      lockcap = (size_t)cap_par;
    }
    std::mutex dmu;
    locks = lockbuf.get();
    ntet = &nt_count;
    dup_mu = &dmu;
    const int nthr = std::max(1, nthreads);
    std::vector<TCtx> ctx((size_t)nthr);
    ctx[0].freel = free_slots;
    free_slots.clear();
    for (int w = 0; w < nthr; ++w) {
      ctx[w].me = (uint32_t)w + 1;
      ctx[w].last = last;
      ctx[w].recent.push_back(last);
    }
    slab_sizes.assign((size_t)std::max(1, nslabs), 0);
    ph[2] = now_ms() - t0;
    t0 = now_ms();
    int rc = concurrent_pass(lvl1, pred1, ctx, axis, nslabs, xlo, xhi);
    ph[3] = now_ms() - t0;
    t0 = now_ms();
    if (rc == 0) rc = concurrent_pass(rest, pred2, ctx, axis, nslabs, xlo, xhi);
    ph[4] = now_ms() - t0;
    t0 = now_ms();
    // points that met taken locks twice: more concurrent passes while there
    // are more than a few of them (contention is a matter of timing), then
    // one thread (slow: the tetrahedra it visits sit in other cores' caches,
    // about 12 us per point against 4 us in a pass)
    retried = 0;
    for (int w = 0; w < nthr; ++w) retried += (int64_t)ctx[w].retry.size();
    // every retried point starts its walk next to the point before it on
    // the curve: vtet is refreshed first to a live tetrahedron of each
    // vertex (the one made at its insertion may be gone)
    if (retried > 0 && rc == 0) refresh_vtet(ctx);
    for (int round = 0; round < 6 && rc == 0; ++round) {
      std::vector<int> pend, ppred;
      for (auto &c : ctx) {
        pend.insert(pend.end(), c.retry.begin(), c.retry.end());
        ppred.insert(ppred.end(), c.retry_pred.begin(), c.retry_pred.end());
        c.retry.clear();
        c.retry_pred.clear();
      }
      if (pend.size() < 64) {
        ctx[0].retry.swap(pend);
        ctx[0].retry_pred.swap(ppred);
        break;
      }
      rc = concurrent_pass(pend, ppred, ctx, axis, nslabs, xlo, xhi, false);
    }
    for (int w = 0; w < nthr && rc == 0; ++w)
      for (size_t i = 0; i < ctx[w].retry.size(); ++i) {
        hint_from(ctx[0], ctx[w].retry_pred[i]);
        const int r2 = par_insert(ctx[0], ctx[w].retry[i]);
        if (r2 == 2 || r2 < 0) {
          rc = r2 < 0 ? r2 : SM_ERR_NUMERIC;
          break;
        }
      }
    ph[5] = now_ms() - t0;
    t0 = now_ms();
    // slots handed out in blocks but never used were never written
    for (auto &c : ctx) {
      for (int64_t t = c.blk; t < c.blk_end; ++t) T[(size_t)t].v[0] = kDead;
      for (int t : c.skipped) T[(size_t)t].v[0] = kDead;
    }
    // back to the serial layout: dead slots become free slots
    const int64_t used_n = std::min<int64_t>(nt_count.load(), cap_par);
    T.resize((size_t)std::max<int64_t>(used_n, (int64_t)old_n));
    {
      // (on the threads: a scan of every slot; the list stays ascending)
      std::vector<std::vector<int>> dead((size_t)nthr);
      const int64_t ns = (int64_t)T.size();
      on_threads(nthr, [&](int w) {
        for (int64_t t = ns * w / nthr; t < ns * (w + 1) / nthr; ++t)
          if (T[(size_t)t].v[0] == kDead) dead[(size_t)w].push_back((int)t);
      });
      free_slots.clear();
      for (auto &d : dead) free_slots.insert(free_slots.end(), d.begin(), d.end());
    }
    // the serial marks are all clear (the concurrent passes kept theirs in
    // the lock words)
    if (mark.size() < T.size() + 64) mark.resize(T.size() + 64, 0);
    locks = nullptr;
    ntet = nullptr;
    dup_mu = nullptr;
    last = 0;
    for (size_t t = 0; t < T.size(); ++t)
      if (T[t].v[0] != kDead) {
        last = (int)t;
        break;
      }
    ph[6] = now_ms() - t0;
    return rc;
  }

  // One concurrent pass over pts (pred[i]: a point inserted before pts[i],
  // next to it on the curve).
  int concurrent_pass(const std::vector<int> &pts, const std::vector<int> &pred,
                      std::vector<TCtx> &ctx, int axis, int nslabs, double xlo, double xhi,
                      bool count = true) {
    const int S = std::max(1, nslabs);
    const double width = (xhi - xlo) / S;
    std::vector<std::vector<int>> slab((size_t)S), spred((size_t)S);
    for (size_t i = 0; i < pts.size(); ++i) {
      const int v = pts[i];
      int k = width > 0 ? (int)std::floor((pt(v)[axis] - xlo) / width) : 0;
      k = std::min(S - 1, std::max(0, k));
      slab[(size_t)k].push_back(v);
      spred[(size_t)k].push_back(pred[i]);
    }
    if (count)
      for (int k = 0; k < S; ++k) slab_sizes[k] += (int64_t)slab[k].size();
    const int nthr = (int)ctx.size();
    const size_t chunk = std::max<size_t>(256, pts.size() / ((size_t)nthr * 8) + 1);
    // each slab's chunks in curve order, taken by one atomic cursor per
    // slab; thread w works through the slabs w, w + nthr, ... first (it
    // stays where its caches are warm) and then takes chunks from the
    // others
    std::vector<size_t> nch((size_t)S);
    for (int k = 0; k < S; ++k) nch[(size_t)k] = (slab[(size_t)k].size() + chunk - 1) / chunk;
    std::unique_ptr<std::atomic<size_t>[]> cursor(new std::atomic<size_t>[(size_t)S]);
    for (int k = 0; k < S; ++k) cursor[(size_t)k].store(0);
    std::atomic<int> fail{0};
    on_threads(nthr, [&](int w) {
      TCtx &c = ctx[w];
      std::vector<int> again;
      int prev_slab = -1;
      for (int round = 0; round < 2; ++round) {
        for (int j = 0; j < S; ++j) {
          // round 0: own slabs; round 1: every slab, starting next to w
          const int k = round == 0 ? w + j * nthr : (w + 1 + j) % S;
          if (round == 0 && k >= S) break;
          for (;;) {
            if (fail.load(std::memory_order_relaxed)) return;
            const size_t r0 = cursor[(size_t)k].fetch_add(1);
            if (r0 >= nch[(size_t)k]) break;
            // odd slabs start half-way along their curve, so that the
            // threads of neighbouring slabs do not meet at the boundary
            const size_t r = (r0 + ((k & 1) ? nch[(size_t)k] / 2 : 0)) % nch[(size_t)k];
            const size_t b = r * chunk, e = std::min(slab[(size_t)k].size(), b + chunk);
            if (k != prev_slab || r0 == 0 || r == 0) {
              // in another slab, start near the chunk when that
              // tetrahedron still lives (the previous position stays as
              // the fallback); the next chunk of the same slab starts
              // where the last one ended
              const int hv = __atomic_load_n(&vtet[(size_t)spred[(size_t)k][b]], __ATOMIC_RELAXED);
              if (hv >= 0) {
                c.recent.push_back(c.last);
                c.last = hv;
              }
            }
            prev_slab = k;
            if (int rc = insert_chunk(c, slab[(size_t)k], spred[(size_t)k], b, e, again,
                                      !count)) {
              fail = rc;
              return;
            }
          }
        }
      }
    });
    return fail.load();
  }

  // Inserts sp[b..e) on thread context c; points that meet a taken lock go
  // once more at the end of the chunk, then (with their entries of pred)
  // to c.retry.  per_point: every point starts its walk from the
  // tetrahedron of its pred (retried points lie far apart on the curve).
  int insert_chunk(TCtx &c, const std::vector<int> &sp, const std::vector<int> &pred,
                   size_t b, size_t e, std::vector<int> &again, bool per_point) {
    again.clear();
    for (size_t k = b; k < e; ++k) {
      if (per_point) hint_from(c, pred[k]);
      const int rc = par_insert(c, sp[k]);
      if (rc == 2) again.push_back((int)k);
      else if (rc < 0) return rc;
    }
    for (int k : again) {
      if (per_point) hint_from(c, pred[(size_t)k]);
      const int rc = par_insert(c, sp[(size_t)k]);
      if (rc == 2) {
        c.retry.push_back(sp[(size_t)k]);
        c.retry_pred.push_back(pred[(size_t)k]);
      } else if (rc < 0) {
        return rc;
      }
    }
    return 0;
  }

  // the next walk of c starts at the tetrahedron recorded for vertex u
  // (when there is one; the current start stays as the fallback)
  void hint_from(TCtx &c, int u) {
    const int hv = u >= 0 ? __atomic_load_n(&vtet[(size_t)u], __ATOMIC_RELAXED) : -1;
    if (hv >= 0 && hv != c.last) {
      c.recent.push_back(c.last);
      if (c.recent.size() > 64) c.recent.erase(c.recent.begin(), c.recent.begin() + 32);
      c.last = hv;
    }
  }

  // vtet[v] := a live tetrahedron with vertex v, for every inserted v
  // (threads over ranges of slots; any of the writers of a vertex will do).
  // The threads' blocks hold slots not yet handed out, never written: they
  // are marked dead first (handing one out later writes it whole).
  void refresh_vtet(std::vector<TCtx> &ctx) {
    const int nthr = (int)ctx.size();
    for (auto &c : ctx) {
      for (int64_t t = c.blk; t < c.blk_end; ++t) T[(size_t)t].v[0] = kDead;
      for (int t : c.skipped) T[(size_t)t].v[0] = kDead;
    }
    const int64_t nslot = std::min<int64_t>(ntet->load(), cap_par);
    on_threads(nthr, [&](int w) {
      for (int64_t t = nslot * w / nthr; t < nslot * (w + 1) / nthr; ++t) {
        const Tet &r = T[(size_t)t];
        if (r.v[0] == kDead) continue;
        for (int j = 0; j < 4; ++j)
          if (r.v[j] >= 0) __atomic_store_n(&vtet[(size_t)r.v[j]], (int)t, __ATOMIC_RELAXED);
      }
    });
  }
};


// ------------------------------------------------------------ driver
DT3::DT3() : B(new Builder3()) {}
DT3::~DT3() = default;

int DT3::run(const double *p, int64_t n, int axis, int nslabs, int nthreads, const Emit &emit,
             int32_t *out, int64_t out_cap) {
  if (n < 0 || (n > 0 && !p) || axis < 0 || axis > 2 || nslabs < 1 || nthreads < 1)
    return sm_set_error(SM_ERR_ARG, "delaunay3d: bad arguments");
  if (n > (int64_t)std::numeric_limits<int32_t>::max() / 8)
    return sm_set_error(SM_ERR_ARG, "delaunay3d: too many points");
  const double r0 = now_ms();
  stats = Stats{};
  inserted.assign((size_t)n, 1);
  Builder3 &b = *B;
  b.nloc = (int)n;
  b.P.assign(p, p + 3 * (size_t)n);
  b.gid.resize((size_t)n);
  for (int64_t i = 0; i < n; ++i) b.gid[i] = i;
  if (n >= 4) {
    double lo[3] = {p[0], p[1], p[2]}, hi[3] = {p[0], p[1], p[2]};
    for (int64_t i = 1; i < n; ++i)
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], p[3 * i + k]);
        hi[k] = std::max(hi[k], p[3 * i + k]);
      }
    double ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    if (!(ext > 0)) ext = 1.0;
    int rc;
    try {
      rc = b.build(lo, ext, nthreads, axis, nslabs);
    } catch (const std::bad_alloc &) {
      rc = SM_ERR_CAPACITY;
    }
    if (rc) return sm_set_error(rc, "host 3D Delaunay failed (%d)", rc);
    for (int64_t g : b.dropped) inserted[g] = 0;
    stats.parallel = b.slab_sizes.empty() ? 0 : 1;
    stats.retried = b.retried;
    stats.slab_points = b.slab_sizes;
  } else {
    b.T.clear();
    b.free_slots.clear();
  }
  // finite tetrahedra, canonical, extracted by the threads in ranges of
  // slots, then put in increasing order of (lowest id, the rest): a
  // counting sort on the lowest id and a sort within each of its short
  // buckets -- the order is then that of the serial build too
  const double t_ex = now_ms();
  const int64_t nslot = (int64_t)b.T.size();
  for (int f : b.free_slots) b.T[f].v[0] = Builder3::kDead;  // free slots: stale contents
  const int nthr = std::max(1, std::min(nthreads, (int)(nslot / 65536 + 1)));
  // Thread w owns slots [nslot w / nthr, ...) for the extraction and the
  // range of lowest ids [n r / nthr, n (r + 1) / nthr) with r = w for the
  // sort.  Pass 1 appends each thread's cells, canonical, to its buffer
  // for their range (buffers kept between calls); pass 2 counting-sorts
  // each range's cells by lowest id into the output (reading the buffers
  // of threads 0, 1, ... in turn) and sorts every bucket.
  auto lo_id = [&](int w) { return n * w / nthr; };
  auto range_of = [&](int32_t v0) {
    int rg = (int)((int64_t)v0 * nthr / std::max<int64_t>(n, 1));
    while (rg + 1 < nthr && lo_id(rg + 1) <= v0) ++rg;
    while (rg > 0 && lo_id(rg) > v0) --rg;
    return rg;
  };
  auto live = [&](const Tet &r) { return r.v[0] >= 0 && Builder3::inf_slot(r) < 0; };
  if ((int)part_.size() != nthr * nthr) part_.assign((size_t)nthr * nthr, {});
  b.on_threads(nthr, [&](int w) {
    const int64_t s0 = nslot * w / nthr, s1 = nslot * (w + 1) / nthr;
    std::vector<int32_t> *row = part_.data() + (size_t)w * nthr;
    for (int r = 0; r < nthr; ++r) row[r].clear();
    for (int64_t t = s0; t < s1; ++t) {
      const Tet &r = b.T[t];
      if (!live(r)) continue;  // dead or infinite
      int32_t v[4];
      for (int j = 0; j < 4; ++j) v[j] = (int32_t)b.gid[r.v[j]];
      canonical_tet(v);
      std::vector<int32_t> &dstv = row[range_of(v[0])];
      dstv.insert(dstv.end(), v, v + 4);
    }
  });
  // below[r]: cells with lowest id below range r
  std::vector<int64_t> below((size_t)nthr + 1, 0);
  for (int r = 0; r < nthr; ++r) {
    int64_t c = 0;
    for (int w = 0; w < nthr; ++w) c += (int64_t)part_[(size_t)w * nthr + r].size() / 4;
    below[(size_t)r + 1] = below[(size_t)r] + c;
  }
  const int64_t m = below[(size_t)nthr];
  int32_t *dst = out;
  if (!out || out_cap < m) {
    sorted_.resize((size_t)m * 4);
    dst = sorted_.data();
  }
  start_.assign((size_t)n + 1, 0);
  auto *cells = reinterpret_cast<std::array<int32_t, 4> *>(dst);
  b.on_threads(nthr, [&](int w) {
    // start_[i + 1] ends bucket i; this thread writes start_[a + 1 .. b] only
    const int64_t a = lo_id(w), bb = lo_id(w + 1), g0 = below[(size_t)w];
    std::vector<int64_t> fill((size_t)(bb - a), 0);
    for (int u = 0; u < nthr; ++u) {
      const std::vector<int32_t> &src = part_[(size_t)u * nthr + w];
      for (size_t k = 0; k < src.size(); k += 4) ++fill[(size_t)(src[k] - a)];
    }
    int64_t run = g0;
    for (int64_t i = a; i < bb; ++i) {
      const int64_t c = fill[(size_t)(i - a)];
      fill[(size_t)(i - a)] = run;
      run += c;
      start_[(size_t)i + 1] = run;
    }
    for (int u = 0; u < nthr; ++u) {
      const std::vector<int32_t> &src = part_[(size_t)u * nthr + w];
      for (size_t k = 0; k < src.size(); k += 4) {
        const int64_t o = fill[(size_t)(src[k] - a)]++;
        std::memcpy(dst + 4 * o, &src[k], 4 * sizeof(int32_t));
      }
    }
    for (int64_t i = a; i < bb; ++i) {
      const int64_t c0 = i == a ? g0 : start_[(size_t)i];
      if (start_[(size_t)i + 1] - c0 > 1) std::sort(cells + c0, cells + start_[(size_t)i + 1]);
    }
  });
  const double r1 = now_ms();
  // handed out in nslabs contiguous pieces
  const int pieces = std::max(1, nslabs);
  for (int k = 0; k < pieces; ++k) {
    const int64_t c0 = m * k / pieces, c1 = m * (k + 1) / pieces;
    emit(k, dst + 4 * c0, c1 - c0);
  }
  emit(-1, nullptr, 0);
  if (trace_dt3()) {
    const double *ph = b.ph;
    std::fprintf(stderr,
                 "dt3: %lld points, %lld cells, %d threads: keys+sort %.1f, serial %.1f, "
                 "setup %.1f, pass 1 %.1f, pass 2 %.1f, retries %.1f (%lld), teardown %.1f, "
                 "extract+sort %.1f, pieces %.1f, total %.1f ms\n",
                 (long long)n, (long long)m, nthreads, ph[0], ph[1], ph[2], ph[3], ph[4], ph[5],
                 (long long)b.retried, ph[6], r1 - t_ex, now_ms() - r1, now_ms() - r0);
  }
  return SM_OK;
}

int DT3::verify(const double *p, int64_t n, const int32_t *cells, int64_t count,
                   int nthreads, int64_t max_bins) {
  if (count == 0) return 1;
  if (n <= 0) return 0;
  bin_points(p, n);
  std::atomic<int> ok{1};
  std::atomic<int64_t> next{0};
  auto worker = [&]() {
    for (;;) {
      const int64_t c = next.fetch_add(1);
      if (c >= count || !ok.load(std::memory_order_relaxed)) return;
      // every point is tested (an empty skip range)
      if (sphere_empty(p, cells + 4 * c, 0, 1.0, 0.0, max_bins) != 1) ok = 0;
    }
  };
  const int nthr = (int)std::max<int64_t>(1, std::min<int64_t>(nthreads, count / 16 + 1));
  std::vector<std::thread> pool;
  for (int t = 1; t < nthr; ++t) pool.emplace_back(worker);
  worker();
  for (auto &t : pool) t.join();
  return ok.load();
}


void DT3::bin_points(const double *p, int64_t n) {
  // bins of about two points each over the points' box
  for (int k = 0; k < 3; ++k) blo_[k] = bhi_[k] = p[k];
  for (int64_t i = 1; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      blo_[k] = std::min(blo_[k], p[3 * i + k]);
      bhi_[k] = std::max(bhi_[k], p[3 * i + k]);
    }
  double vol = 1.0;
  int dims = 0;
  for (int k = 0; k < 3; ++k)
    if (bhi_[k] > blo_[k]) {
      vol *= bhi_[k] - blo_[k];
      ++dims;
    }
  const double cell = dims ? std::pow(vol * 2.0 / (double)n, 1.0 / dims) : 1.0;
  for (int k = 0; k < 3; ++k)
    nb_[k] = std::max<int64_t>(1, std::min<int64_t>(1024, (int64_t)((bhi_[k] - blo_[k]) / cell) + 1));
  const int64_t nbins = nb_[0] * nb_[1] * nb_[2];
  bin_start.assign(nbins + 1, 0);
  bin_pts.resize((size_t)n);
  std::vector<int64_t> idx((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    const double *q = p + 3 * i;
    idx[i] = (bin_of(q[0], 0) * nb_[1] + bin_of(q[1], 1)) * nb_[2] + bin_of(q[2], 2);
    ++bin_start[idx[i] + 1];
  }
  for (int64_t b = 0; b < nbins; ++b) bin_start[b + 1] += bin_start[b];
  std::vector<int32_t> fillp(bin_start.begin(), bin_start.end() - 1);
  for (int64_t i = 0; i < n; ++i) bin_pts[fillp[idx[i]]++] = (int32_t)i;
}

int64_t DT3::bin_of(double x, int k) const {
  const double c = bhi_[k] > blo_[k] ? (x - blo_[k]) / (bhi_[k] - blo_[k]) * (double)nb_[k] : 0.0;
  if (!(c >= 0.0)) return 0;
  return std::min<int64_t>(nb_[k] - 1, (int64_t)c);
}

// 1: no point of p outside [skip_lo, skip_hi] along `axis` (all points when
// the range is empty) lies inside the sphere of the positively oriented
// cell v; 0: one does; -1: the sphere covers more than max_bins bins.
int DT3::sphere_empty(const double *p, const int32_t *v, int axis, double skip_lo,
                         double skip_hi, int64_t max_bins) const {
  const double *a = p + 3 * (size_t)v[0], *b = p + 3 * (size_t)v[1];
  const double *cc = p + 3 * (size_t)v[2], *d = p + 3 * (size_t)v[3];
  double u[3], w[3], z[3];
  for (int k = 0; k < 3; ++k) {
    u[k] = b[k] - a[k];
    w[k] = cc[k] - a[k];
    z[k] = d[k] - a[k];
  }
  const double wz[3] = {w[1] * z[2] - w[2] * z[1], w[2] * z[0] - w[0] * z[2],
                        w[0] * z[1] - w[1] * z[0]};
  const double zu[3] = {z[1] * u[2] - z[2] * u[1], z[2] * u[0] - z[0] * u[2],
                        z[0] * u[1] - z[1] * u[0]};
  const double uw[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2],
                        u[0] * w[1] - u[1] * w[0]};
  const double den = 2.0 * (u[0] * wz[0] + u[1] * wz[1] + u[2] * wz[2]);
  const double lu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const double lw = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  const double lz = z[0] * z[0] + z[1] * z[1] + z[2] * z[2];
  double o[3];
  for (int k = 0; k < 3; ++k) o[k] = (lu * wz[k] + lw * zu[k] + lz * uw[k]) / den;
  const double r = std::sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
  if (!std::isfinite(r)) return -1;
  // a generous margin: the box only limits which points are tested
  const double L = std::sqrt(std::max(lu, std::max(lw, lz)));
  const double rr = r * (1.0 + 1e-6) + 1e-9 * L;
  int64_t b0[3], b1[3], span = 1;
  for (int k = 0; k < 3; ++k) {
    b0[k] = bin_of(a[k] + o[k] - rr, k);
    b1[k] = bin_of(a[k] + o[k] + rr, k);
    span *= b1[k] - b0[k] + 1;
  }
  if (span > max_bins) return -1;
  for (int64_t i = b0[0]; i <= b1[0]; ++i)
    for (int64_t j = b0[1]; j <= b1[1]; ++j)
      for (int64_t k = b0[2]; k <= b1[2]; ++k) {
        const int64_t bb = (i * nb_[1] + j) * nb_[2] + k;
        for (int32_t s = bin_start[bb]; s < bin_start[bb + 1]; ++s) {
          const int32_t q = bin_pts[s];
          if (q == v[0] || q == v[1] || q == v[2] || q == v[3]) continue;
          const double *Q = p + 3 * (size_t)q;
          if (Q[axis] >= skip_lo && Q[axis] <= skip_hi) continue;
          bool dup = false;
          for (int t = 0; t < 4 && !dup; ++t) {
            const double *V = p + 3 * (size_t)v[t];
            dup = V[0] == Q[0] && V[1] == Q[1] && V[2] == Q[2];
          }
          if (!dup && insphere_perturbed(a, b, cc, d, Q) > 0) return 0;
        }
      }
  return 1;
}

}  // namespace smdt

// Returns the number of finite tetrahedra written to `tet` ((m, 4) int32,
// positively oriented for orient3d of the predicates, canonical, in
// increasing order), or an error code.
int64_t sm_host_delaunay3d(const double *p, int64_t n, int32_t *tet, int64_t cap,
                           uint8_t *inserted) {
  return sm_host_delaunay3d_slabs(p, n, 0, 1, 1, tet, cap, inserted);
}

int sm_host_delaunay3d_check(const double *p, int64_t n, const int32_t *tet, int64_t count,
                             int nthreads) {
  if (n < 0 || count < 0 || (n > 0 && !p) || (count > 0 && !tet) || nthreads < 1)
    return sm_set_error(SM_ERR_ARG, "host_delaunay3d_check: bad argument");
  for (int64_t i = 0; i < 4 * count; ++i)
    if (tet[i] < 0 || tet[i] >= n)
      return sm_set_error(SM_ERR_ARG, "host_delaunay3d_check: vertex id out of range");
  smdt::DT3 dt;
  return dt.verify(p, n, tet, count, nthreads, (int64_t)1 << 40);
}

int64_t sm_host_delaunay3d_slabs(const double *p, int64_t n, int axis, int nslabs, int nthreads,
                                 int32_t *tet, int64_t cap, uint8_t *inserted) {
  if (n < 0 || (n > 0 && (!p || !tet || !inserted)))
    return sm_set_error(SM_ERR_ARG, "host_delaunay3d: null argument");
  smdt::DT3 dt;
  int64_t m = 0;
  bool over = false;
  int rc = dt.run(p, n, axis, nslabs, nthreads, [&](int, const int32_t *v, int64_t count) {
    if (m + count > cap) {
      over = true;
      return;
    }
    if (count) std::memcpy(tet + 4 * m, v, (size_t)count * 4 * sizeof(int32_t));
    m += count;
  });
  if (rc) return rc;
  if (over) return sm_set_error(SM_ERR_CAPACITY, "host_delaunay3d: output capacity %lld",
                                (long long)cap);
  for (int64_t i = 0; i < n; ++i) inserted[i] = dt.inserted[i];
  return m;
}
