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
// own points plus every point whose axis coordinate lies in the window
// [lo_k - halo, hi_k + halo].  It reports the
// tetrahedra it owns -- those whose lowest-numbered vertex lies in slab k --
// and certifies each one whose circumsphere spans, along the axis, less than
// the window: every point inside or on that sphere is then in the window,
// so the tetrahedron is Delaunay for the whole point set, and (by the
// uniqueness above) every tetrahedron of the whole set owned by slab k whose
// sphere fits the window is found this way.  Owned tetrahedra that cannot be
// certified are reported after all certified ones; the caller checks that
// none of them survives its centroid filter, and otherwise rebuilds
// serially.
//
// Output tetrahedra are positively oriented and canonical: the lowest vertex
// id first, the second-lowest next, the last two in the order that keeps
// the orientation.  Exact duplicates are skipped: of several equal points
// only the lowest id is used.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
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

inline uint64_t hilbert3(uint32_t x, uint32_t y, uint32_t z, int bits) {
  // Skilling's transpose-form Hilbert index in 3D
  uint32_t X[3] = {x, y, z};
  uint32_t M = 1u << (bits - 1), P, Q, t;
  for (Q = M; Q > 1; Q >>= 1) {
    P = Q - 1;
    for (int i = 0; i < 3; ++i) {
      if (X[i] & Q) X[0] ^= P;
      else {
        t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  for (int i = 1; i < 3; ++i) X[i] ^= X[i - 1];
  t = 0;
  for (Q = M; Q > 1; Q >>= 1)
    if (X[2] & Q) t ^= Q - 1;
  for (int i = 0; i < 3; ++i) X[i] ^= t;
  uint64_t h = 0;
  for (int b = bits - 1; b >= 0; --b)
    for (int i = 0; i < 3; ++i) h = (h << 1) | ((X[i] >> b) & 1u);
  return h;
}

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
  std::vector<int64_t> dropped;   // global ids of skipped duplicates
  std::vector<Tet> T;
  std::vector<int> free_slots;
  std::vector<char> mark;
  std::vector<std::pair<uint64_t, int>> key;
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

  // Triangulates the nloc points of P in Hilbert order.
  int build(const double lo[3], double ext) {
    T.clear();
    free_slots.clear();
    dropped.clear();
    const int n = nloc;
    last = 0;
    if (mark.size() < 64) mark.assign(64, 0);
    key.resize(n);
    for (int i = 0; i < n; ++i) {
      uint32_t g[3];
      for (int k = 0; k < 3; ++k)
        g[k] = (uint32_t)std::min(1023.0, std::max(0.0, (pt(i)[k] - lo[k]) / ext * 1023.0));
      key[i] = {hilbert3(g[0], g[1], g[2], 10), i};
    }
    std::sort(key.begin(), key.end());
    const size_t k0 = start(key);
    if (k0 == 0) {
      T.clear();
      return SM_OK;  // fewer than four points off one plane: no tetrahedra
    }
    for (size_t k = k0; k < key.size(); ++k) {
      int rc = insert(key[k].second);
      if (rc < 0) return rc;
    }
    for (int v : std::vector<int>(deferred)) {
      int rc = insert(v);
      if (rc < 0) return rc;
    }
    return SM_OK;
  }
};

namespace {

// Half-width of the sphere of a..d along `axis` and its centre coordinate,
// with a bound on their rounding error; false when the tetrahedron is too
// flat for the bound to mean anything.
bool sphere_span(const double *a, const double *b, const double *c, const double *d, int axis,
                 double *centre, double *radius, double *err) {
  double u[3], v[3], w[3];
  for (int k = 0; k < 3; ++k) {
    u[k] = b[k] - a[k];
    v[k] = c[k] - a[k];
    w[k] = d[k] - a[k];
  }
  const double vw[3] = {v[1] * w[2] - v[2] * w[1], v[2] * w[0] - v[0] * w[2],
                        v[0] * w[1] - v[1] * w[0]};
  const double wu[3] = {w[1] * u[2] - w[2] * u[1], w[2] * u[0] - w[0] * u[2],
                        w[0] * u[1] - w[1] * u[0]};
  const double uv[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2],
                        u[0] * v[1] - u[1] * v[0]};
  const double den = 2.0 * (u[0] * vw[0] + u[1] * vw[1] + u[2] * vw[2]);
  const double lu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const double lv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  const double lw = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
  const double L2 = std::max(lu, std::max(lv, lw));
  const double L3 = L2 * std::sqrt(L2);
  if (!(std::fabs(den) > 1e-9 * L3)) return false;
  double o[3];
  for (int k = 0; k < 3; ++k) o[k] = (lu * vw[k] + lv * wu[k] + lw * uv[k]) / den;
  const double r = std::sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
  const double eps = std::numeric_limits<double>::epsilon();
  // forward bound: numerator and denominator each within ~10 roundings of
  // their magnitudes (3 L^4 and 2 L^3), plus the final additions
  const double e = 64.0 * eps * (3.0 * L2 * L2 + r * 2.0 * L3) / std::fabs(den) +
                   8.0 * eps * (std::fabs(a[axis]) + r);
  *centre = a[axis] + o[axis];
  *radius = r;
  *err = e;
  return std::isfinite(*centre) && std::isfinite(r) && std::isfinite(e);
}

}  // namespace

// ------------------------------------------------------------ slab driver
Slabs3::Slabs3() = default;
Slabs3::~Slabs3() = default;

int Slabs3::run(const double *p, int64_t n, int axis, int nslabs, double halo, int nthreads,
                const Emit &emit) {
  if (n < 0 || (n > 0 && !p) || axis < 0 || axis > 2 || nslabs < 1 || !(halo >= 0.0))
    return sm_set_error(SM_ERR_ARG, "delaunay3d: bad arguments");
  if (n > (int64_t)std::numeric_limits<int32_t>::max() - 8)
    return sm_set_error(SM_ERR_ARG, "delaunay3d: too many points");
  stats = Stats{};
  inserted.assign((size_t)n, 1);
  const int S = nslabs;
  certified.resize(S);
  uncertified.resize(S);
  for (int k = 0; k < S; ++k) {
    certified[k].clear();
    uncertified[k].clear();
  }
  if (n < 4) {
    for (int k = 0; k < S; ++k) emit(k, nullptr, 0);
    emit(-1, nullptr, 0);
    return SM_OK;
  }
  double lo[3] = {p[0], p[1], p[2]}, hi[3] = {p[0], p[1], p[2]};
  for (int64_t i = 1; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], p[3 * i + k]);
      hi[k] = std::max(hi[k], p[3 * i + k]);
    }
  double ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
  if (!(ext > 0)) ext = 1.0;
  // slab of each point and the window of each slab
  const double x0 = lo[axis];
  const double width = (hi[axis] - lo[axis]) / S;
  core.resize((size_t)n);
  win_lo.assign(S, -std::numeric_limits<double>::infinity());
  win_hi.assign(S, std::numeric_limits<double>::infinity());
  if (S > 1) {
    for (int k = 0; k < S; ++k) {
      win_lo[k] = x0 + k * width - halo;
      win_hi[k] = x0 + (k + 1) * width + halo;
    }
  }
  members.resize(S);
  for (int k = 0; k < S; ++k) members[k].clear();
  for (int64_t i = 0; i < n; ++i) {
    const double x = p[3 * i + axis];
    int s = 0;
    if (S > 1 && width > 0) s = (int)std::floor((x - x0) / width);
    s = std::min(S - 1, std::max(0, s));
    core[i] = s;
    members[s].push_back((int32_t)i);
    if (S > 1) {
      // neighbouring windows that also hold x (and never the core slab twice)
      for (int k = s - 1; k >= 0 && x >= win_lo[k] && x <= win_hi[k]; --k)
        members[k].push_back((int32_t)i);
      for (int k = s + 1; k < S && x >= win_lo[k] && x <= win_hi[k]; ++k)
        members[k].push_back((int32_t)i);
    }
  }
  const int nthr = std::max(1, std::min(nthreads, S));
  if ((int)builders.size() < nthr) builders.resize(nthr);
  for (auto &b : builders)
    if (!b) b.reset(new Builder3());

  // worker: builds slab k on builder w and sorts its owned tetrahedra into
  // the certified and uncertified lists
  auto work = [&](int k, Builder3 &B) -> int {
    // ids in increasing order (the points were visited in order)
    const std::vector<int32_t> &ids = members[k];
    const int m = (int)ids.size();
    B.nloc = m;
    B.P.resize(3 * (size_t)m);
    B.gid.resize((size_t)m);
    double llo[3], lhi[3];
    for (int k2 = 0; k2 < 3; ++k2) {
      llo[k2] = std::numeric_limits<double>::infinity();
      lhi[k2] = -std::numeric_limits<double>::infinity();
    }
    for (int j = 0; j < m; ++j) {
      const double *q = p + 3 * (size_t)ids[j];
      for (int k2 = 0; k2 < 3; ++k2) {
        B.P[3 * (size_t)j + k2] = q[k2];
        llo[k2] = std::min(llo[k2], q[k2]);
        lhi[k2] = std::max(lhi[k2], q[k2]);
      }
      B.gid[j] = ids[j];
    }
    double lext = 0;
    for (int k2 = 0; k2 < 3; ++k2) lext = std::max(lext, lhi[k2] - llo[k2]);
    if (!(lext > 0)) lext = 1.0;
    if (m == 0) return SM_OK;
    int rc = B.build(llo, lext);
    if (rc) return rc;
    for (int64_t g : B.dropped)
      if (core[g] == k) inserted[g] = 0;
    std::vector<char> dead(B.T.size(), 0);
    for (int f : B.free_slots) dead[f] = 1;
    std::vector<int32_t> &cert = certified[k], &unc = uncertified[k];
    for (size_t t = 0; t < B.T.size(); ++t) {
      if (dead[t]) continue;
      const Tet &r = B.T[t];
      if (Builder3::inf_slot(r) >= 0) continue;
      int32_t v[4];
      for (int j = 0; j < 4; ++j) v[j] = (int32_t)B.gid[r.v[j]];
      int32_t vmin = std::min(std::min(v[0], v[1]), std::min(v[2], v[3]));
      if (core[vmin] != k) continue;
      bool ok = true;
      if (S > 1) {
        double cx, rad, err;
        ok = sphere_span(B.pt(r.v[0]), B.pt(r.v[1]), B.pt(r.v[2]), B.pt(r.v[3]), axis, &cx,
                         &rad, &err) &&
             cx - rad - err >= win_lo[k] && cx + rad + err <= win_hi[k];
      }
      canonical_tet(v);
      std::vector<int32_t> &dst = ok ? cert : unc;
      dst.insert(dst.end(), v, v + 4);
    }
    return SM_OK;
  };

  auto guarded = [&](int k, Builder3 &B) -> int {
    try {
      return work(k, B);
    } catch (const std::bad_alloc &) {
      return SM_ERR_CAPACITY;
    }
  };
  std::vector<int> status(S, 1);  // 1 pending, 0 done, < 0 failed
  std::mutex mu;
  std::condition_variable cv;
  std::atomic<int> next{0};
  auto worker = [&](int w) {
    for (;;) {
      const int k = next.fetch_add(1);
      if (k >= S) return;
      int rc = guarded(k, *builders[w]);
      {
        std::lock_guard<std::mutex> lk(mu);
        status[k] = rc < 0 ? rc : 0;
      }
      cv.notify_all();
    }
  };
  std::vector<std::thread> pool;
  for (int w = 1; w < nthr; ++w) pool.emplace_back(worker, w);
  // the calling thread builds too, then hands finished slabs out in order
  int first_err = 0;
  int handed = 0;
  auto hand_out = [&](bool wait) {
    while (handed < S) {
      int st;
      {
        std::unique_lock<std::mutex> lk(mu);
        if (wait)
          cv.wait(lk, [&] { return status[handed] != 1; });
        else if (status[handed] == 1)
          return;
        st = status[handed];
      }
      if (st < 0 && !first_err) first_err = st;
      if (!first_err) {
        const auto &v = certified[handed];
        emit(handed, v.data(), (int64_t)(v.size() / 4));
      }
      ++handed;
    }
  };
  for (;;) {
    const int k = next.fetch_add(1);
    if (k >= S) break;
    int rc = guarded(k, *builders[0]);
    {
      std::lock_guard<std::mutex> lk(mu);
      status[k] = rc < 0 ? rc : 0;
    }
    hand_out(false);
  }
  hand_out(true);
  for (auto &t : pool) t.join();
  if (first_err) return sm_set_error(first_err, "host 3D Delaunay failed (%d)", first_err);
  // the uncertified lists, concatenated, last
  unc_all.clear();
  for (int k = 0; k < S; ++k) {
    unc_all.insert(unc_all.end(), uncertified[k].begin(), uncertified[k].end());
    stats.certified += (int64_t)certified[k].size() / 4;
    stats.local_points += (int64_t)members[k].size();
  }
  stats.uncertified = (int64_t)unc_all.size() / 4;
  emit(-1, unc_all.data(), stats.uncertified);
  return SM_OK;
}

}  // namespace smdt

// Returns the number of finite tetrahedra written to `tet` ((m, 4) int32,
// positively oriented for orient3d of the predicates, canonical order), or
// an error code.
int64_t sm_host_delaunay3d(const double *p, int64_t n, int32_t *tet, int64_t cap,
                           uint8_t *inserted) {
  int64_t ncert = 0;
  return sm_host_delaunay3d_slabs(p, n, 0, 1, 0.0, 1, tet, cap, inserted, &ncert);
}

int64_t sm_host_delaunay3d_slabs(const double *p, int64_t n, int axis, int nslabs, double halo,
                                 int nthreads, int32_t *tet, int64_t cap, uint8_t *inserted,
                                 int64_t *ncertified) {
  if (n < 0 || (n > 0 && (!p || !tet || !inserted)) || !ncertified)
    return sm_set_error(SM_ERR_ARG, "host_delaunay3d: null argument");
  smdt::Slabs3 sl;
  int64_t m = 0;
  bool over = false;
  int rc = sl.run(p, n, axis, nslabs, halo, nthreads,
                  [&](int slab, const int32_t *v, int64_t count) {
                    if (slab < 0) *ncertified = m;
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
  for (int64_t i = 0; i < n; ++i) inserted[i] = sl.inserted[i];
  return m;
}
