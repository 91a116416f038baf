// Host 3D Delaunay tetrahedralisation (incremental Bowyer-Watson, exact
// predicates).
//
// Replaces the role of CGAL's Delaunay_triangulation_3 behind
// SeismicMesh/generation/cpp/delaunay_class3.cpp:116-129 (insert) and
// :276-344 (get_finite_cells / get_finite_vertices), which the 3D loop and
// sliver_removal call every iteration (mesh_generator.py:202-215, :465-469).
//
// The points are enclosed in a tetrahedron of four far vertices (1e5 times
// the bounding box), inserted in Hilbert-curve order with a walk from the
// last created tetrahedron, and the tetrahedra touching the far vertices are
// dropped at the end.  Every tetrahedron whose circumsphere is smaller than
// the far vertices' distance is exactly a Delaunay tetrahedron of the
// points; only hull slivers with enormous circumspheres can differ from the
// triangulation of the convex hull, and those never pass the centroid
// filter of the force loop.  Exact duplicates are skipped.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

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

struct Builder3 {
  std::vector<double> P;  // points + 4 far vertices, (n + 4, 3)
  std::vector<Tet> T;
  std::vector<int> free_slots;
  std::vector<char> mark;
  int last = 0;

  const double *pt(int v) const { return P.data() + 3 * (size_t)v; }

  int orient(int a, int b, int c, int d) const {
    return smpred::orient3d_h(pt(a), pt(b), pt(c), pt(d));
  }

  bool conflict(int t, const double *q) const {
    const Tet &r = T[t];
    return smpred::insphere_h(pt(r.v[0]), pt(r.v[1]), pt(r.v[2]), pt(r.v[3]), q) > 0;
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

  // visibility walk; -1 when q duplicates a vertex
  int locate(int qi) {
    const double *q = pt(qi);
    int t = last;
    for (size_t steps = 0; steps < 8 * T.size() + 64; ++steps) {
      const Tet &r = T[t];
      int moved = -1;
      for (int i = 0; i < 4 && moved < 0; ++i) {
        int w[4] = {r.v[0], r.v[1], r.v[2], r.v[3]};
        w[i] = qi;
        if (smpred::orient3d_h(pt(w[0]), pt(w[1]), pt(w[2]), pt(w[3])) < 0) moved = i;
      }
      if (moved < 0) {
        for (int i = 0; i < 4; ++i) {
          const double *V = pt(r.v[i]);
          if (V[0] == q[0] && V[1] == q[1] && V[2] == q[2]) return -1;
        }
        return t;
      }
      t = r.n[moved];
      if (t < 0) return -2;
    }
    return -2;
  }

  // per-insertion scratch, reused across insertions
  struct BF {
    int t, i, out, back;  // cavity tet, its face, the outside tet, its face
  };
  std::vector<int> cav, stack, ids;
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

  bool insert(int qi) {
    int t0 = locate(qi);
    if (t0 < 0) return false;
    const double *q = pt(qi);
    cav.clear();
    bnd.clear();
    stack.assign(1, t0);
    if ((size_t)T.size() > mark.size()) mark.resize(T.size() * 2, 0);
    mark[t0] = 1;
    while (!stack.empty()) {
      int t = stack.back();
      stack.pop_back();
      cav.push_back(t);
      for (int i = 0; i < 4; ++i) {
        int nb = T[t].n[i];
        if (nb >= 0 && mark[nb]) continue;
        if (nb >= 0 && conflict(nb, q)) {
          mark[nb] = 1;
          stack.push_back(nb);
        } else {
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
    ids.resize(nb);
    for (size_t k = 0; k < nb && k < nc; ++k) ids[k] = cav[k];
    for (size_t k = nc; k < nb; ++k) ids[k] = new_tet();
    for (size_t k = nb; k < nc; ++k) free_slots.push_back(cav[k]);
    if ((size_t)T.size() > mark.size()) mark.resize(T.size() * 2, 0);
    for (size_t k = 0; k < nb; ++k) T[ids[k]] = nt[k];
    // outside neighbours point back to the new tetrahedra
    for (size_t k = 0; k < nb; ++k)
      if (bnd[k].out >= 0) T[bnd[k].out].n[bnd[k].back] = ids[k];
    // link new tetrahedra through their faces containing q: the face of
    // new tet k opposite vertex j (j != slot of q) is {q} + the boundary
    // face's edge opposite that vertex; each such edge is shared by exactly
    // two new tetrahedra
    size_t cap = 64;
    while (cap < 4 * 3 * nb) cap <<= 1;
    if (table.size() < cap) table.assign(cap, Slot{~0ull, -1, -1});
    const size_t mask = cap - 1;
    used.clear();
    for (size_t k = 0; k < nb; ++k) {
      const Tet &r = T[ids[k]];
      const int qs = bnd[k].i;
      for (int j = 0; j < 4; ++j) {
        if (j == qs) continue;
        // the two slots other than j and qs
        const int s0 = kOther[qs][j][0], s1 = kOther[qs][j][1];
        const int e[2] = {r.v[s0], r.v[s1]};
        const uint64_t key = ((uint64_t)(uint32_t)std::min(e[0], e[1]) << 32) |
                             (uint32_t)std::max(e[0], e[1]);
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
    last = ids[0];
    return true;
  }
};

}  // namespace

// Returns the number of finite tetrahedra written to `tet` ((m, 4) int32,
// positively oriented for orient3d of the predicates), or an error code.
int64_t sm_host_delaunay3d(const double *p, int64_t n, int32_t *tet, int64_t cap,
                           uint8_t *inserted) {
  if (n < 0 || (n > 0 && !p)) return SM_ERR_ARG;
  for (int64_t i = 0; i < n; ++i) inserted[i] = 0;
  if (n < 4) return 0;
  double lo[3] = {p[0], p[1], p[2]}, hi[3] = {p[0], p[1], p[2]};
  for (int64_t i = 1; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], p[3 * i + k]);
      hi[k] = std::max(hi[k], p[3 * i + k]);
    }
  double ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
  if (!(ext > 0)) ext = 1.0;
  Builder3 B;
  B.P.assign(p, p + 3 * n);
  const double c[3] = {(lo[0] + hi[0]) / 2, (lo[1] + hi[1]) / 2, (lo[2] + hi[2]) / 2};
  const double K = 1e5 * ext;
  const double far[4][3] = {{1, 1, 1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, 1}};
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 3; ++k) B.P.push_back(c[k] + K * far[i][k]);
  int f0 = (int)n, f1 = (int)n + 1, f2 = (int)n + 2, f3 = (int)n + 3;
  if (B.orient(f0, f1, f2, f3) < 0) std::swap(f2, f3);
  B.T.push_back(Tet{{f0, f1, f2, f3}, {-1, -1, -1, -1}});
  B.mark.assign(64, 0);
  // Hilbert order on a 2^10 grid
  std::vector<std::pair<uint64_t, int>> key(n);
  for (int64_t i = 0; i < n; ++i) {
    uint32_t g[3];
    for (int k = 0; k < 3; ++k)
      g[k] = (uint32_t)std::min(1023.0, (p[3 * i + k] - lo[k]) / ext * 1023.0);
    key[i] = {hilbert3(g[0], g[1], g[2], 10), (int)i};
  }
  std::sort(key.begin(), key.end());
  for (int64_t k = 0; k < n; ++k) {
    int q = key[k].second;
    if (B.insert(q)) inserted[q] = 1;
  }
  int64_t m = 0;
  std::vector<char> dead(B.T.size(), 0);
  for (int f : B.free_slots) dead[f] = 1;
  for (size_t t = 0; t < B.T.size(); ++t) {
    if (dead[t]) continue;
    const Tet &r = B.T[t];
    if (r.v[0] >= n || r.v[1] >= n || r.v[2] >= n || r.v[3] >= n) continue;
    if (m >= cap) return SM_ERR_CAPACITY;
    for (int j = 0; j < 4; ++j) tet[4 * m + j] = r.v[j];
    ++m;
  }
  return m;
}
