// Host repair of 2D triangulations that edge flips cannot untangle.
//
// When the parallel flips end with inverted triangles left -- typically a
// vertex that slid across a nearly flat triangle along a straight piece of
// boundary, or one driven almost onto a domain corner -- a vertex of each
// inverted triangle is taken out of the triangulation (its link polygon is
// re-triangulated by ear clipping) and put back in (visibility walk to the
// triangle that strictly contains it, split into three).  The slots stay the
// same (k star triangles become k - 2 polygon triangles; the split uses the
// two freed slots), so the triangulation keeps its layout on the device and
// the device flip rounds then restore the Delaunay property.  All
// orientation tests are exact (predicates.h).  Ghost triangles (vertex -1)
// and dead slots (vertex -2) are never modified; a vertex whose star holds
// one is left alone.
#include <algorithm>
#include <array>
#include <cstdint>
#include <unordered_map>
#include <vector>

#include "predicates.h"
#include "relocate2d.h"

namespace {

struct Mesh {
  int32_t *tri, *adj;  // slot range [t0, t0 + nt), indexed by global slot
  int64_t t0, nt;
  const double *p;     // points by global vertex id

  int32_t &v(int64_t s, int k) { return tri[3 * (s - t0) + k]; }
  int32_t &n(int64_t s, int k) { return adj[3 * (s - t0) + k]; }
  const double *pt(int32_t x) const { return p + 2 * (int64_t)x; }
  int orient(int32_t a, int32_t b, int32_t c) const {
    return smpred::orient2d_h(pt(a), pt(b), pt(c));
  }
  bool real(int64_t s) { return v(s, 0) >= 0 && v(s, 1) >= 0 && v(s, 2) >= 0; }
};

int index_of(Mesh &m, int64_t s, int32_t x) {
  for (int k = 0; k < 3; ++k)
    if (m.v(s, k) == x) return k;
  return -1;
}

// Removes vertex x and re-inserts it; returns false (mesh unchanged) when
// any step does not apply.
bool relocate_vertex(Mesh &m, int32_t x) {
  // star of x
  std::vector<int64_t> star;
  for (int64_t s = m.t0; s < m.t0 + m.nt; ++s)
    if (index_of(m, s, x) >= 0) {
      if (!m.real(s)) return false;
      star.push_back(s);
    }
  const int k = (int)star.size();
  if (k < 3) return false;
  // link edges a -> b of the star triangles rotated to (x, a, b)
  struct Link {
    int32_t b, outer;
    int64_t slot;
  };
  std::unordered_map<int32_t, Link> link;
  for (int64_t s : star) {
    const int j = index_of(m, s, x);
    link[m.v(s, (j + 1) % 3)] = {m.v(s, (j + 2) % 3), m.n(s, j), s};
  }
  if ((int)link.size() != k) return false;
  std::vector<int32_t> poly{link.begin()->first};
  for (;;) {
    auto it = link.find(poly.back());
    if (it == link.end()) return false;
    if (it->second.b == poly[0]) break;
    poly.push_back(it->second.b);
    if ((int)poly.size() > k) return false;
  }
  if ((int)poly.size() != k) return false;
  // ear clipping
  std::vector<int32_t> ring = poly;
  std::vector<std::array<int32_t, 3>> tris;
  while (ring.size() > 3) {
    const int r = (int)ring.size();
    bool clipped = false;
    for (int i = 0; i < r && !clipped; ++i) {
      const int32_t a = ring[(i + r - 1) % r], b = ring[i], c = ring[(i + 1) % r];
      if (m.orient(a, b, c) <= 0) continue;
      bool empty = true;
      for (int32_t w : ring) {
        if (w == a || w == b || w == c) continue;
        if (m.orient(a, b, w) >= 0 && m.orient(b, c, w) >= 0 && m.orient(c, a, w) >= 0) {
          empty = false;
          break;
        }
      }
      if (!empty) continue;
      tris.push_back({a, b, c});
      ring.erase(ring.begin() + i);
      clipped = true;
    }
    if (!clipped) return false;
  }
  if (m.orient(ring[0], ring[1], ring[2]) <= 0) return false;
  tris.push_back({ring[0], ring[1], ring[2]});
  // write the polygon triangles into the first k - 2 star slots
  std::vector<int32_t> save_tri, save_adj;
  for (int64_t s : star)
    for (int q = 0; q < 3; ++q) {
      save_tri.push_back(m.v(s, q));
      save_adj.push_back(m.n(s, q));
    }
  struct Back {
    int32_t outer;
    int q;
    int32_t old;
  };
  std::vector<Back> backs;
  auto undo = [&]() {
    for (int i = 0; i < k; ++i)
      for (int q = 0; q < 3; ++q) {
        m.v(star[i], q) = save_tri[3 * i + q];
        m.n(star[i], q) = save_adj[3 * i + q];
      }
    for (const Back &bk : backs) m.n(bk.outer, bk.q) = bk.old;
  };
  std::unordered_map<int64_t, std::pair<int64_t, int>> owner;  // edge u->w
  auto key = [](int32_t u, int32_t w) { return ((int64_t)u << 32) | (uint32_t)w; };
  for (int i = 0; i < k - 2; ++i) {
    for (int q = 0; q < 3; ++q) m.v(star[i], q) = tris[i][q];
    for (int q = 0; q < 3; ++q)
      owner[key(tris[i][(q + 1) % 3], tris[i][(q + 2) % 3])] = {star[i], q};
  }
  for (int i = 0; i < k - 2; ++i)
    for (int q = 0; q < 3; ++q) {
      const int32_t u = tris[i][(q + 1) % 3], w = tris[i][(q + 2) % 3];
      auto it = owner.find(key(w, u));
      if (it != owner.end()) {
        m.n(star[i], q) = (int32_t)it->second.first;
        continue;
      }
      // a link edge: the outer neighbour now borders this slot
      const Link &lk = link[u];
      m.n(star[i], q) = lk.outer;
      const int bq = [&]() {
        for (int z = 0; z < 3; ++z)
          if (m.n(lk.outer, z) == (int32_t)lk.slot) return z;
        return -1;
      }();
      if (bq < 0) {
        undo();
        return false;
      }
      backs.push_back({lk.outer, bq, m.n(lk.outer, bq)});
      m.n(lk.outer, bq) = (int32_t)star[i];
    }
  // walk from the polygon to the triangle strictly containing x
  int64_t cur = star[0];
  for (int64_t step = 0; step < 4 * m.nt + 16; ++step) {
    if (!m.real(cur)) break;
    bool moved = false;
    for (int q = 0; q < 3 && !moved; ++q) {
      const int32_t a = m.v(cur, (q + 1) % 3), b = m.v(cur, (q + 2) % 3);
      if (m.orient(a, b, x) < 0) {
        cur = m.n(cur, q);
        moved = true;
      }
    }
    if (!moved) break;
  }
  if (cur < m.t0 || cur >= m.t0 + m.nt || !m.real(cur)) {
    undo();
    return false;
  }
  int32_t z0 = m.v(cur, 0), z1 = m.v(cur, 1), z2 = m.v(cur, 2);
  const int o0 = m.orient(z1, z2, x), o1 = m.orient(z2, z0, x), o2 = m.orient(z0, z1, x);
  if (o0 < 0 || o1 < 0 || o2 < 0 || (o0 == 0) + (o1 == 0) + (o2 == 0) > 1) {
    undo();
    return false;
  }
  if (o0 == 0 || o1 == 0 || o2 == 0) {
    // x on an edge of Z (a straight run of boundary points): split Z and its
    // neighbour across that edge into four
    const int e = o0 == 0 ? 0 : (o1 == 0 ? 1 : 2);  // edge opposite z_e
    const int64_t Z = cur;
    const int32_t a = m.v(Z, e), b = m.v(Z, (e + 1) % 3), c = m.v(Z, (e + 2) % 3);
    // Z = (a, b, c), x on (b, c); W = neighbour across (b, c) = (c, b, w)
    const int32_t W = m.n(Z, e);
    if (W < m.t0 || W >= m.t0 + m.nt || !m.real(W)) {
      undo();
      return false;
    }
    int jw = -1;
    for (int q = 0; q < 3; ++q)
      if (m.v(W, q) != b && m.v(W, q) != c) jw = q;
    if (jw < 0) {
      undo();
      return false;
    }
    const int32_t w = m.v(W, jw);
    if (!(m.orient(c, w, x) > 0 && m.orient(w, b, x) > 0)) {
      undo();
      return false;
    }
    const int32_t Zab = m.n(Z, (e + 2) % 3), Zca = m.n(Z, (e + 1) % 3);  // across (a,b), (c,a)
    int32_t Wcw = -1, Wwb = -1;  // W's neighbours across (c, w) and (w, b)
    for (int q = 0; q < 3; ++q) {
      const int32_t u = m.v(W, (q + 1) % 3), t = m.v(W, (q + 2) % 3);
      if ((u == c && t == w) || (u == w && t == c)) Wcw = m.n(W, q);
      if ((u == w && t == b) || (u == b && t == w)) Wwb = m.n(W, q);
    }
    const int64_t B = star[k - 2], C = star[k - 1];
    // T1 = (a, b, x) in Z, T2 = (c, a, x) in B, T3 = (w, c, x) in W, T4 = (b, w, x) in C
    m.v(Z, 0) = a; m.v(Z, 1) = b; m.v(Z, 2) = x;
    m.n(Z, 0) = (int32_t)C; m.n(Z, 1) = (int32_t)B; m.n(Z, 2) = Zab;
    m.v(B, 0) = c; m.v(B, 1) = a; m.v(B, 2) = x;
    m.n(B, 0) = (int32_t)Z; m.n(B, 1) = W; m.n(B, 2) = Zca;
    m.v(W, 0) = w; m.v(W, 1) = c; m.v(W, 2) = x;
    m.n(W, 0) = (int32_t)B; m.n(W, 1) = (int32_t)C; m.n(W, 2) = Wcw;
    m.v(C, 0) = b; m.v(C, 1) = w; m.v(C, 2) = x;
    m.n(C, 0) = W; m.n(C, 1) = (int32_t)Z; m.n(C, 2) = Wwb;
    auto relink2 = [&](int32_t nb, int32_t oldt, int32_t u, int32_t t, int64_t to) {
      if (nb < 0) return;
      for (int z = 0; z < 3; ++z) {
        if (m.n(nb, z) != oldt) continue;
        const int32_t p1 = m.v(nb, (z + 1) % 3), p2 = m.v(nb, (z + 2) % 3);
        if ((p1 == u && p2 == t) || (p1 == t && p2 == u)) m.n(nb, z) = (int32_t)to;
      }
    };
    relink2(Zca, (int32_t)Z, c, a, B);
    relink2(Wwb, W, w, b, C);
    return true;
  }
  const int64_t Z = cur, B = star[k - 2], C = star[k - 1];
  const int32_t N0 = m.n(Z, 0), N1 = m.n(Z, 1), N2 = m.n(Z, 2);
  // A = (z1, z2, x) in Z, B = (z2, z0, x), C = (z0, z1, x)
  m.v(Z, 0) = z1; m.v(Z, 1) = z2; m.v(Z, 2) = x;
  m.n(Z, 0) = (int32_t)B; m.n(Z, 1) = (int32_t)C; m.n(Z, 2) = N0;
  m.v(B, 0) = z2; m.v(B, 1) = z0; m.v(B, 2) = x;
  m.n(B, 0) = (int32_t)C; m.n(B, 1) = (int32_t)Z; m.n(B, 2) = N1;
  m.v(C, 0) = z0; m.v(C, 1) = z1; m.v(C, 2) = x;
  m.n(C, 0) = (int32_t)Z; m.n(C, 1) = (int32_t)B; m.n(C, 2) = N2;
  // neighbours across (z2, z0) and (z0, z1) now border B and C: fix the
  // link whose edge is that pair (a triangle may border Z twice)
  auto relink = [&](int32_t nb, int32_t u, int32_t w, int64_t to) {
    if (nb < 0) return;
    for (int z = 0; z < 3; ++z) {
      if (m.n(nb, z) != (int32_t)Z) continue;
      const int32_t a = m.v(nb, (z + 1) % 3), b = m.v(nb, (z + 2) % 3);
      if ((a == u && b == w) || (a == w && b == u)) m.n(nb, z) = (int32_t)to;
    }
  };
  relink(N1, z2, z0, B);
  relink(N2, z0, z1, C);
  return true;
}

}  // namespace

int64_t sm_relocate2d(const double *p, const double *pold, int32_t *tri, int32_t *adj,
                      int64_t t0, int64_t nt, int64_t *failed) {
  Mesh m{tri, adj, t0, nt, p};
  int64_t done = 0;
  *failed = 0;
  for (int64_t s = t0; s < t0 + nt; ++s) {
    if (!m.real(s) || m.orient(m.v(s, 0), m.v(s, 1), m.v(s, 2)) > 0) continue;
    // the vertex that moved most first
    int32_t vs[3] = {m.v(s, 0), m.v(s, 1), m.v(s, 2)};
    double mv[3];
    for (int q = 0; q < 3; ++q) {
      const double dx = p[2 * (int64_t)vs[q]] - pold[2 * (int64_t)vs[q]];
      const double dy = p[2 * (int64_t)vs[q] + 1] - pold[2 * (int64_t)vs[q] + 1];
      mv[q] = dx * dx + dy * dy;
    }
    int order[3] = {0, 1, 2};
    std::sort(order, order + 3, [&](int a, int b) { return mv[a] > mv[b]; });
    bool ok = false;
    for (int q = 0; q < 3 && !ok; ++q) ok = relocate_vertex(m, vs[order[q]]);
    if (ok) ++done;
    else ++*failed;
  }
  return done;
}
