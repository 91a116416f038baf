// Host 2D Delaunay triangulation (incremental Bowyer-Watson, exact predicates).
//
// Replaces the role of CGAL's Delaunay_triangulation_2 behind
// SeismicMesh/generation/cpp/delaunay_class.cpp:108-121 (insert) and
// :232-298 (get_finite_cells / get_finite_vertices): it triangulates the
// convex hull of all points.  The result is stored the way the device flip
// repair consumes it: every triangle of the triangulated sphere, i.e. real
// triangles plus one "ghost" triangle (u, v, INF) per convex-hull edge, in
// counter-clockwise order, with neighbour j opposite vertex j.
//
// Points are inserted along a Hilbert curve, coarse to fine (every 4096th
// point of the curve, then every 64th, then the rest), from a copy stored
// in curve order; point location walks from the last created triangle; of
// exact duplicates the lowest id is kept.
// Cavities use the strict in-circle test, so the
// triangulation of cocircular configurations depends on insertion order
// (exactly as for any incremental method); for points in general position the
// Delaunay triangulation is unique and this builder returns it.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "predicates.h"
#include "smgpu_internal.h"

namespace {

constexpr int INF = -1;

struct Tri {
  int v[3];
  int n[3];
};

inline int hilbert_d(uint32_t n, uint32_t x, uint32_t y) {
  // index of (x, y) along a Hilbert curve filling an n x n grid (n = 2^k)
  uint32_t d = 0;
  for (uint32_t s = n / 2; s > 0; s /= 2) {
    uint32_t rx = (x & s) > 0;
    uint32_t ry = (y & s) > 0;
    d += s * s * ((3 * rx) ^ ry);
    if (ry == 0) {
      if (rx == 1) {
        x = s - 1 - x;
        y = s - 1 - y;
      }
      uint32_t t = x;
      x = y;
      y = t;
    }
  }
  return (int)d;
}

struct Builder {
  const double *P;
  std::vector<Tri> T;
  std::vector<int> free_slots;
  int last = 0;

  const double *pt(int v) const { return P + 2 * (size_t)v; }

  bool is_ghost(int t) const {
    const Tri &r = T[t];
    return r.v[0] == INF || r.v[1] == INF || r.v[2] == INF;
  }

  // conflict of point q (index) with triangle t
  bool conflict(int t, int q) const {
    const Tri &r = T[t];
    const double *Q = pt(q);
    int g = r.v[0] == INF ? 0 : r.v[1] == INF ? 1 : r.v[2] == INF ? 2 : -1;
    if (g < 0) return smpred::incircle_h(pt(r.v[0]), pt(r.v[1]), pt(r.v[2]), Q) > 0;
    // ghost (u, v, INF): region strictly left of u->v, plus the open segment uv
    int u = r.v[(g + 1) % 3], v = r.v[(g + 2) % 3];
    int o = smpred::orient2d_h(pt(u), pt(v), Q);
    if (o > 0) return true;
    if (o < 0) return false;
    // Q is exactly on the line through U and V: conflict iff strictly
    // between them (the ghost's circle degenerates to the open segment)
    const double *U = pt(u), *V = pt(v);
    if (U[0] != V[0])
      return Q[0] > std::min(U[0], V[0]) && Q[0] < std::max(U[0], V[0]);
    return Q[1] > std::min(U[1], V[1]) && Q[1] < std::max(U[1], V[1]);
  }

  int new_tri() {
    if (!free_slots.empty()) {
      int t = free_slots.back();
      free_slots.pop_back();
      return t;
    }
    T.push_back(Tri{});
    return (int)T.size() - 1;
  }

  // returns a triangle in conflict with q, or -1 when q duplicates a vertex
  int locate(int q) {
    const double *Q = pt(q);
    int t = last;
    for (size_t steps = 0; steps < 4 * T.size() + 64; ++steps) {
      const Tri &r = T[t];
      int g = r.v[0] == INF ? 0 : r.v[1] == INF ? 1 : r.v[2] == INF ? 2 : -1;
      if (g >= 0) {
        if (conflict(t, q)) return t;
        // step back into the hull through the real edge (opposite INF)
        t = r.n[g];
        continue;
      }
      int moved = 0;
      for (int k = 0; k < 3; ++k) {
        int a = r.v[(k + 1) % 3], b = r.v[(k + 2) % 3];
        if (smpred::orient2d_h(pt(a), pt(b), Q) < 0) {
          t = r.n[k];
          moved = 1;
          break;
        }
      }
      if (moved) continue;
      for (int k = 0; k < 3; ++k) {
        const double *V = pt(r.v[k]);
        if (V[0] == Q[0] && V[1] == Q[1]) return -1;
      }
      return t;
    }
    // walk did not terminate (should not happen with exact predicates):
    // fall back to a linear scan for a conflicting triangle
    for (size_t i = 0; i < T.size(); ++i) {
      if (std::find(free_slots.begin(), free_slots.end(), (int)i) != free_slots.end()) continue;
      for (int k = 0; k < 3; ++k) {
        int v = T[i].v[k];
        if (v != INF && pt(v)[0] == Q[0] && pt(v)[1] == Q[1]) return -1;
      }
      if (conflict((int)i, q)) return (int)i;
    }
    return -2;
  }

  bool insert(int q) {
    int t0 = locate(q);
    if (t0 == -1) return false;  // duplicate point
    if (t0 < 0) return false;
    // cavity by depth-first search over conflicting triangles
    std::vector<int> cav, stack{t0};
    std::vector<char> &mark = marks;
    struct BE {
      int u, v, out;
    };
    std::vector<BE> bnd;
    mark_set(t0);
    while (!stack.empty()) {
      int t = stack.back();
      stack.pop_back();
      cav.push_back(t);
      for (int k = 0; k < 3; ++k) {
        int nb = T[t].n[k];
        int u = T[t].v[(k + 1) % 3], v = T[t].v[(k + 2) % 3];
        if (is_marked(nb)) continue;
        if (conflict(nb, q)) {
          mark_set(nb);
          stack.push_back(nb);
        } else {
          bnd.push_back({u, v, nb});
        }
      }
    }
    (void)mark;
    // new triangles (u, v, q), one per cavity boundary edge
    std::vector<int> ids(bnd.size());
    for (size_t i = 0; i < bnd.size(); ++i) ids[i] = i < cav.size() ? cav[i] : -1;
    for (size_t i = cav.size(); i < bnd.size(); ++i) ids[i] = new_tri();
    for (size_t i = bnd.size(); i < cav.size(); ++i) free_slots.push_back(cav[i]);
    for (int t : cav) mark_clear(t);
    for (size_t i = 0; i < bnd.size(); ++i) {
      Tri &r = T[ids[i]];
      r.v[0] = bnd[i].u;
      r.v[1] = bnd[i].v;
      r.v[2] = q;
      r.n[2] = bnd[i].out;
      // outside neighbour: repoint its edge {u, v} to the new triangle
      Tri &o = T[bnd[i].out];
      for (int k = 0; k < 3; ++k) {
        int a = o.v[(k + 1) % 3], b = o.v[(k + 2) % 3];
        if ((a == bnd[i].v && b == bnd[i].u) || (a == bnd[i].u && b == bnd[i].v)) {
          o.n[k] = ids[i];
          break;
        }
      }
    }
    // link new triangles around q: (u,v,q) shares {v,q} with the one starting at v
    for (size_t i = 0; i < bnd.size(); ++i) {
      for (size_t j = 0; j < bnd.size(); ++j) {
        if (bnd[j].u == bnd[i].v) T[ids[i]].n[0] = ids[j];
        if (bnd[j].v == bnd[i].u) T[ids[i]].n[1] = ids[j];
      }
    }
    last = ids[0];
    for (size_t i = 0; i < bnd.size(); ++i)
      if (!is_ghost(ids[i])) {
        last = ids[i];
        break;
      }
    return true;
  }

  std::vector<char> marks;
  void mark_set(int t) {
    if ((size_t)t >= marks.size()) marks.resize(std::max<size_t>(2 * marks.size(), t + 1), 0);
    marks[t] = 1;
  }
  void mark_clear(int t) {
    if ((size_t)t < marks.size()) marks[t] = 0;
  }
  bool is_marked(int t) const { return (size_t)t < marks.size() && marks[t]; }
};

}  // namespace

// Builds the full (ghost-augmented) Delaunay triangulation.  Returns the
// number of triangles written (real + ghost), or a negative error code.
// `inserted[i]` is set to 0 for points skipped as exact duplicates.
int64_t sm_host_delaunay2d(const double *p, int64_t n, int32_t *tri, int32_t *adj,
                           int64_t cap, uint8_t *inserted) {
  if (n < 0 || (n > 0 && p == nullptr)) return SM_ERR_ARG;
  for (int64_t i = 0; i < n; ++i) inserted[i] = 0;
  if (n < 3) return 0;
  // Hilbert order on a 2^16 grid over the bounding box
  double lo[2] = {p[0], p[1]}, hi[2] = {p[0], p[1]};
  for (int64_t i = 1; i < n; ++i)
    for (int k = 0; k < 2; ++k) {
      lo[k] = std::min(lo[k], p[2 * i + k]);
      hi[k] = std::max(hi[k], p[2 * i + k]);
    }
  double ext = std::max(hi[0] - lo[0], hi[1] - lo[1]);
  if (!(ext > 0)) ext = 1.0;
  std::vector<std::pair<int, int>> key(n);
  for (int64_t i = 0; i < n; ++i) {
    uint32_t x = (uint32_t)std::min(65535.0, (p[2 * i] - lo[0]) / ext * 65535.0);
    uint32_t y = (uint32_t)std::min(65535.0, (p[2 * i + 1] - lo[1]) / ext * 65535.0);
    key[i] = {hilbert_d(65536, x, y), (int)i};
  }
  // stable two-pass radix sort on the 32-bit keys (ids ascending within a
  // key): the order of std::sort on (key, id); the keys are signed ints
  // (the curve index wraps past 2^31), so the sign bit is flipped
  {
    std::vector<std::pair<int, int>> tmp((size_t)n);
    std::vector<int64_t> cnt(1 << 16);
    for (int pass = 0; pass < 2; ++pass) {
      const int sh = 16 * pass;
      std::fill(cnt.begin(), cnt.end(), 0);
      for (int64_t i = 0; i < n; ++i) ++cnt[(((uint32_t)key[i].first ^ 0x80000000u) >> sh) & 0xFFFF];
      int64_t run = 0;
      for (int64_t &c : cnt) {
        const int64_t x = c;
        c = run;
        run += x;
      }
      for (int64_t i = 0; i < n; ++i)
        tmp[(size_t)cnt[(((uint32_t)key[i].first ^ 0x80000000u) >> sh) & 0xFFFF]++] = key[i];
      key.swap(tmp);
    }
  }
  std::vector<int> order(n);
  for (int64_t i = 0; i < n; ++i) order[i] = key[i].second;
  // exact duplicates share a curve cell, where the lowest id comes first:
  // every later copy is skipped, whatever order the points go in below
  std::vector<char> dupl((size_t)n, 0);
  for (int64_t a = 0; a < n;) {
    int64_t e = a + 1;
    while (e < n && key[e].first == key[a].first) ++e;
    for (int64_t u = a + 1; u < e; ++u)
      for (int64_t v = a; v < u; ++v) {
        const int iu = order[u], iv = order[v];
        if (!dupl[v] && p[2 * iu] == p[2 * iv] && p[2 * iu + 1] == p[2 * iv + 1]) {
          dupl[u] = 1;
          break;
        }
      }
    a = e;
  }

  // the points in curve order: the builder works on positions along the
  // curve (local ids), so that points near each other sit together in memory
  std::vector<double> L(2 * (size_t)n);
  for (int64_t k = 0; k < n; ++k) {
    L[2 * k] = p[2 * (int64_t)order[k]];
    L[2 * k + 1] = p[2 * (int64_t)order[k] + 1];
  }
  const double *lp = L.data();
  Builder B;
  B.P = lp;
  // first triangle: two distinct points and a third not collinear with them
  int i0 = 0, i1 = -1, i2 = -1;
  for (int64_t k = 1; k < n; ++k)
    if (lp[2 * k] != lp[0] || lp[2 * k + 1] != lp[1]) {
      i1 = (int)k;
      break;
    }
  if (i1 < 0) {
    inserted[order[0]] = 1;
    return 0;
  }
  int o = 0;
  for (int64_t k = i1 + 1; k < n; ++k) {
    o = smpred::orient2d_h(lp + 2 * i0, lp + 2 * i1, lp + 2 * k);
    if (o != 0) {
      i2 = (int)k;
      break;
    }
  }
  if (i2 < 0) {
    // all points collinear: no triangles (CGAL reports no finite faces)
    return 0;
  }
  if (o < 0) std::swap(i1, i2);
  // real triangle 0 = (i0, i1, i2); ghosts 1..3 on its three edges
  B.T.resize(4);
  B.T[0] = Tri{{i0, i1, i2}, {2, 3, 1}};
  // ghost across edge (i0,i1) [opposite i2]: (i1, i0, INF)
  B.T[1] = Tri{{i1, i0, INF}, {3, 2, 0}};
  // ghost across edge (i1,i2) [opposite i0]: (i2, i1, INF)
  B.T[2] = Tri{{i2, i1, INF}, {1, 3, 0}};
  // ghost across edge (i2,i0) [opposite i1]: (i0, i2, INF)
  B.T[3] = Tri{{i0, i2, INF}, {2, 1, 0}};
  inserted[order[i0]] = inserted[order[i1]] = inserted[order[i2]] = 1;
  // coarse to fine along the curve: every 4096th point, then every 64th,
  // then the rest (shorter walks and fewer cache misses than one sweep)
  for (const size_t stride : {(size_t)4096, (size_t)64, (size_t)1})
    for (size_t k = 0; k < (size_t)n; ++k) {
      if (k % stride != 0 || (stride < 4096 && k % (stride * 64) == 0)) continue;
      const int q = (int)k;
      if (dupl[k] || q == i0 || q == i1 || q == i2) continue;
      if (B.insert(q)) inserted[order[k]] = 1;
    }
  // compact
  std::vector<int> remap(B.T.size(), -1);
  std::vector<char> dead(B.T.size(), 0);
  for (int f : B.free_slots) dead[f] = 1;
  int64_t m = 0;
  for (size_t t = 0; t < B.T.size(); ++t)
    if (!dead[t]) remap[t] = (int)m++;
  if (m > cap) return SM_ERR_CAPACITY;
  for (size_t t = 0; t < B.T.size(); ++t) {
    if (dead[t]) continue;
    int64_t r = remap[t];
    for (int k = 0; k < 3; ++k) {
      const int v = B.T[t].v[k];
      tri[3 * r + k] = v >= 0 ? order[v] : v;  // back to the caller's ids
      adj[3 * r + k] = remap[B.T[t].n[k]];
    }
  }
  return m;
}
