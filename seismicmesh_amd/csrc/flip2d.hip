// Device repair of the 2D Delaunay triangulation after the points move.
//
// Replaces the per-iteration CGAL re-triangulation of
// SeismicMesh/generation/mesh_generator.py:464-469 (delaunay_class.cpp
// insert + get_finite_cells).  The previous iteration's triangulation of the
// whole convex hull is kept on the device as a triangulated sphere: real
// triangles plus one ghost triangle (u, v, INF) per hull edge.  After the
// points move, rounds of parallel edge flips restore it:
//
//  * untangling flips first: an edge next to an inverted (orientation <= 0)
//    triangle is flipped when both new triangles are valid -- for a thin hull
//    sliver whose middle vertex crossed the hull edge this turns the vertex
//    into a hull vertex;
//  * then Lawson flips: an edge is flipped when the opposite vertex lies
//    strictly inside the circumcircle of the other triangle, where a ghost's
//    circle is the open half-plane outside its hull edge (so reflex hull
//    vertices are filled in).
//
// Every flip claims its two triangles with a 64-bit priority (class, then a
// per-round scramble of the edge id); only edges owning both triangles flip,
// so a round's flips are disjoint.  Neighbour links of the triangles around
// a flipped pair are repaired in a separate pass.  Predicates are filtered in
// FP64; signs the filter cannot certify are queued and settled by an exact
// expansion pass (Shewchuk's method, see predicates.h).  When the rounds end
// with no flip and no inverted triangle, every edge is locally Delaunay and
// the triangulation is the Delaunay triangulation of the points; otherwise
// the host rebuilds it from scratch with exact predicates.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <thread>

#include "common.h"
#include "mesher.h"
#include "predicates.h"
#include "relocate2d.h"

namespace smk {

namespace {

constexpr int kGhostOri = 2;
// vertex id of an unused triangle slot (batch handles keep each domain's
// triangles in a fixed slot range; a slot the domain does not use is dead:
// vertices kDeadVertex, neighbours itself, never flipped or kept)
constexpr int32_t kDeadVertex = -2;

__device__ inline const double *PT(const double *p, int v) { return p + 2 * (int64_t)v; }

// A round's counter block (8 ints): [0] flips, [1] inverted triangles,
// [2] orientation queue, [3] in-circle queue, [5] unresolved predicates,
// [6] round id (unique over the handle's life), [7] skip.  A round is
// skipped when the previous one flipped nothing, was skipped or hit an
// unresolved predicate, so a batch of rounds can be enqueued (or replayed
// from a graph) without the host looking at the counters.  Round 0 of a
// batch is opened by k_batch_begin; round r > 0 takes its skip flag and id
// from round r - 1's final counters (its first kernel computes and stores
// them), and its counters are zeroed by round r - 1's last kernel.
__global__ void k_batch_begin(int32_t *c0, const int32_t *round_counter) {
  for (int i = 0; i < 6; ++i) c0[i] = 0;
  c0[6] = *round_counter + 1;
  c0[7] = 0;
}

// Per-domain activity of a batch handle's flip rounds.  act holds four
// arrays of ndom flags: flipped in a round of parity 0 / 1, then has an
// inverted or uncertain triangle in a round of parity 0 / 1.  Round r
// (parity p) marks and applies flips only in domains that flipped in round
// r - 1 or have an inverted triangle now: a domain without a flip in round
// r - 1 had no flip candidate then (its best candidate would have won both
// its triangles, which no other domain shares), and without a change to
// its triangles or points it has none in round r either.  Null for a
// single-domain handle (every triangle active).
struct ActView {
  const int32_t *tdom = nullptr;
  int32_t *act = nullptr;
  int ndom = 0;
  __device__ bool active(int64_t t, int par) const {
    if (!act) return true;
    const int d = tdom[t];
    return act[(1 - par) * ndom + d] || act[(2 + par) * ndom + d];
  }
};

// Opens a batch: every domain counts as having flipped in the round before
// round 0 (parity 1), and round 0's inverted flags start clear.
__global__ void k_act_begin(int32_t *act, int ndom) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ndom) return;
  act[1 * ndom + d] = 1;
  act[2 * ndom + d] = 0;
}

// Neighbour links of triangle s after the flips of round `round`: a link to
// a flipped triangle that no longer holds the shared edge now belongs to its
// partner.
__device__ inline void fix_links(int64_t s, const int32_t *tri, int32_t *adj,
                                 const int32_t *partner, const int32_t *mark, int round) {
  for (int k = 0; k < 3; ++k) {
    int nb = adj[3 * s + k];
    if (mark[nb] != round) continue;
    int u = tri[3 * s + (k + 1) % 3], v = tri[3 * s + (k + 2) % 3];
    bool hu = false, hv = false;
    for (int j = 0; j < 3; ++j) {
      int w = tri[3 * (int64_t)nb + j];
      hu |= (w == u);
      hv |= (w == v);
    }
    if (!(hu && hv)) adj[3 * s + k] = partner[nb];
  }
}

// Round r's first kernel: the link fix of round r - 1, then (unless round r
// is skipped) the claim reset and the orientation of every triangle.
// Uncertain orientations are queued for the exact pass, or, with
// queue_exact == 0, counted as unresolved (the batch then reports failure).
__global__ void k_fix_orient(int64_t T, const int32_t *tri, int32_t *adj,
                             const int32_t *partner, const int32_t *mark, const double *p,
                             int32_t *ori, uint64_t *claim, int32_t *queue,
                             const int32_t *prev, int32_t *c, int queue_exact, ActView av,
                             int par) {
  int skip = c[7], id = c[6];
  if (prev) {
    skip = prev[7] || prev[0] == 0 || prev[5] > 0;
    id = prev[6] + 1;
  }
  int64_t t = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (prev && t == 0) {
    c[6] = id;
    c[7] = skip;
  }
  // this round's flip flags start clear (round r - 1 read the other parity)
  if (av.act && t < av.ndom) av.act[par * av.ndom + t] = 0;
  if (t >= T) return;
  // a domain without flips and without inverted triangles in round r - 1
  // is unchanged: no links to fix, orientations and claims stand
  if (av.act) {
    const int d = av.tdom[t];
    if (!av.act[(1 - par) * av.ndom + d] && !av.act[(3 - par) * av.ndom + d]) return;
  }
  if (prev && !prev[7]) fix_links(t, tri, adj, partner, mark, prev[6]);
  if (skip) return;
  claim[t] = 0ull;
  int a = tri[3 * t], b = tri[3 * t + 1], cc = tri[3 * t + 2];
  // Two triangles sharing two edges (a vertex left with only two ghosts
  // glued around it, after flips next to collinear hull points) make the
  // triangulation invalid although no triangle is inverted: counted as an
  // inverted triangle, so that the repair fails and the exact rebuild takes
  // over.
  if (a != kDeadVertex) {
    const int32_t n0 = adj[3 * t], n1 = adj[3 * t + 1], n2 = adj[3 * t + 2];
    if (n0 == n1 || n1 == n2 || n0 == n2) {
      atomicAdd(&c[1], 1);
      if (av.act) av.act[(2 + par) * av.ndom + av.tdom[t]] = 1;
    }
  }
  if (a < 0 || b < 0 || cc < 0) {
    ori[t] = kGhostOri;
    return;
  }
  int o = smpred::orient2d_filter(PT(p, a), PT(p, b), PT(p, cc));
  if (o == SM_PRED_UNRESOLVED) o = smpred::orient2d_if_exact(PT(p, a), PT(p, b), PT(p, cc));
  if (o == SM_PRED_UNRESOLVED) {
    if (queue_exact) {
      int slot = atomicAdd(&c[2], 1);
      queue[slot] = (int32_t)t;
    } else {
      atomicAdd(&c[5], 1);
    }
    ori[t] = 0;  // provisional, settled by k_orient_exact
    if (av.act) av.act[(2 + par) * av.ndom + av.tdom[t]] = 1;
    return;
  }
  ori[t] = o;
  if (o <= 0) {
    atomicAdd(&c[1], 1);
    if (av.act) av.act[(2 + par) * av.ndom + av.tdom[t]] = 1;
  }
}

// Summary of a batch of rounds into status[0..7]: [0] rounds executed,
// [1] flips, [2] converged (the last executed round flipped nothing),
// [3] inverted triangles seen by that round, [4] unresolved predicates,
// [5] exact-predicate evaluations.
__device__ inline void flip_status(const int32_t *ctr, int nrounds, int32_t *status,
                                   int32_t *round_counter) {
  int last = -1, flips = 0, unres = 0, exact = 0;
  for (int r = 0; r < nrounds; ++r) {
    const int32_t *c = ctr + 8 * r;
    if (c[7]) break;
    last = r;
    flips += c[0];
    unres += c[5];
    exact += c[2] + c[3];
  }
  status[0] = last + 1;
  status[1] = flips;
  status[2] = last >= 0 && ctr[8 * last] == 0;
  status[3] = last >= 0 ? ctr[8 * last + 1] : 0;
  status[4] = unres;
  status[5] = exact;
  status[6] = last >= 0 ? ctr[8 * last + 6] : 0;
  if (last >= 0) *round_counter = ctr[8 * last + 6];
}

// The link fix of a batch's last round, and (one thread) the batch's
// summary (every round's counters are final by now).
__global__ void k_fix_last(int64_t T, const int32_t *tri, int32_t *adj, const int32_t *partner,
                           const int32_t *mark, const int32_t *c, const int32_t *base,
                           int nrounds, int32_t *status, int32_t *round_counter) {
  if (blockIdx.x == 0 && threadIdx.x == 0) flip_status(base, nrounds, status, round_counter);
  if (c[7]) return;
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < T) fix_links(t, tri, adj, partner, mark, c[6]);
}

__global__ void k_orient_exact(const int32_t *tri, const double *p, int32_t *ori,
                               const int32_t *queue, int32_t *ctr) {
  if (ctr[7]) return;
  const int n = ctr[2];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    int t = queue[i];
    const double *pa = PT(p, tri[3 * t]), *pb = PT(p, tri[3 * t + 1]), *pc = PT(p, tri[3 * t + 2]);
    int o = smpred::orient2d_fix(pa, pb, pc);
    if (o == SM_PRED_UNRESOLVED) o = smpred::orient2d_exact(pa, pb, pc);
    if (o == SM_PRED_UNRESOLVED) {
      o = 0;
      atomicAdd(&ctr[5], 1);
    }
    ori[t] = o;
    if (o <= 0) atomicAdd(&ctr[1], 1);
  }
}

__device__ inline uint32_t scramble(uint32_t id, int round) {
  return (id * 2654435761u) ^ ((uint32_t)round * 0x9E3779B9u);
}

// quad of edge k of triangle t: t = (x, a, b), o = (b, a, y)
struct Quad {
  int x, a, b, y;
  int o;
};

__device__ inline Quad quad_of(const int32_t *tri, const int32_t *adj, int64_t t, int k) {
  Quad q;
  q.x = tri[3 * t + k];
  q.a = tri[3 * t + (k + 1) % 3];
  q.b = tri[3 * t + (k + 2) % 3];
  q.o = adj[3 * t + k];
  int y = -2;
  for (int j = 0; j < 3; ++j) {
    int v = tri[3 * (int64_t)q.o + j];
    if (v != q.a && v != q.b) y = v;
  }
  q.y = y;
  return q;
}

// Lawson test "y inside the circle of t = (x, a, b)"; t may be a ghost.
// Returns 1/0, or SM_PRED_UNRESOLVED.
__device__ inline int in_circle_of(const double *p, int x, int a, int b, int y) {
  if (y < 0) return 0;  // INF is never inside a finite circle
  if (x >= 0 && a >= 0 && b >= 0) {
    int s = smpred::incircle_filter(PT(p, x), PT(p, a), PT(p, b), PT(p, y));
    if (s == SM_PRED_UNRESOLVED) s = smpred::incircle_if_exact(PT(p, x), PT(p, a), PT(p, b), PT(p, y));
    if (s == SM_PRED_UNRESOLVED) return s;
    return s > 0;
  }
  // ghost: rotate so INF is last -> (u, v, INF); inside = strictly left of u->v
  int u, v;
  if (x < 0) { u = a; v = b; }
  else if (a < 0) { u = b; v = x; }
  else { u = x; v = a; }
  int s = smpred::orient2d_filter(PT(p, u), PT(p, v), PT(p, y));
  if (s == SM_PRED_UNRESOLVED) s = smpred::orient2d_if_exact(PT(p, u), PT(p, v), PT(p, y));
  if (s == SM_PRED_UNRESOLVED) return s;
  return s > 0;
}

// orientation of a would-be triangle; INF triangles count as valid
__device__ inline int valid_new(const double *p, int a, int b, int c) {
  if (a < 0 || b < 0 || c < 0) return 1;
  int s = smpred::orient2d_filter(PT(p, a), PT(p, b), PT(p, c));
  if (s == SM_PRED_UNRESOLVED) s = smpred::orient2d_if_exact(PT(p, a), PT(p, b), PT(p, c));
  if (s == SM_PRED_UNRESOLVED) return 0;
  return s > 0;
}

__device__ inline bool near_inverted(const int32_t *adj, const int32_t *ori, int64_t t) {
  return ori[adj[3 * t]] <= 0 || ori[adj[3 * t + 1]] <= 0 || ori[adj[3 * t + 2]] <= 0;
}

__global__ void k_mark(int64_t T, const int32_t *tri, const int32_t *adj, const int32_t *ori,
                       const double *p, uint64_t *claim, uint64_t *cand, int32_t *queue,
                       int32_t *ctr, int queue_exact, ActView av, int par) {
  if (ctr[7]) return;
  const int round = ctr[6];
  // no triangle with orientation <= 0 and none left uncertain (the
  // orientation pass counted both; a domain skipped this round has none):
  // the inverted-triangle tests below are all false and their loads are
  // skipped.  An uncertain incircle counted by this kernel meanwhile only
  // sends a thread down the full path.
  const bool tangled = ctr[1] != 0 || ctr[5] != 0;
  int64_t t = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  if (t >= T) return;
  // an idle domain's claims stay 0 (reset by k_fix_orient), so its stale
  // candidates never match in k_apply
  if (!av.active(t, par)) return;
  for (int k = 0; k < 3; ++k) {
    uint64_t prio = 0;
    int64_t o = adj[3 * t + k];
    if (o > t) {
      Quad q = quad_of(tri, adj, t, k);
      if (q.y == -2) {  // inconsistent link: leave it to the validity check
        cand[3 * t + k] = 0;
        continue;
      }
      const bool inv_t = tangled && ori[t] <= 0;
      const bool inv_o = tangled && ori[o] <= 0;
      if (inv_t || inv_o) {
        // untangling flip: both new triangles (x, a, y) and (y, b, x) valid
        bool ghost_new = q.x < 0 || q.y < 0;
        bool ok;
        if (ghost_new) ok = true;  // the real old triangle is the inverted one
        else ok = valid_new(p, q.x, q.a, q.y) && valid_new(p, q.y, q.b, q.x);
        if (q.a < 0 || q.b < 0) {
          // edge through INF between two ghosts: the new real triangle
          // (x, a, y) or (y, b, x) must be positive
          ok = valid_new(p, q.x, q.a, q.y) && valid_new(p, q.y, q.b, q.x);
        }
        if (ok) prio = (2ull << 32) | scramble((uint32_t)(3 * t + k), round);
      } else if (tangled && (near_inverted(adj, ori, t) || near_inverted(adj, ori, o))) {
        // no Lawson flips next to an inverted triangle: in a tangled
        // neighbourhood they can fold two triangles onto each other; the
        // untangling flips there go first
      } else {
        int s = in_circle_of(p, q.x, q.a, q.b, q.y);
        if (s == SM_PRED_UNRESOLVED) {
          if (queue_exact) {
            int slot = atomicAdd(&ctr[3], 1);
            queue[slot] = (int32_t)(3 * t + k);
          } else {
            atomicAdd(&ctr[5], 1);
          }
        } else if (s) {
          prio = (1ull << 32) | scramble((uint32_t)(3 * t + k), round);
        }
      }
      if (prio) {
        atomicMax((unsigned long long *)&claim[t], (unsigned long long)prio);
        atomicMax((unsigned long long *)&claim[o], (unsigned long long)prio);
      }
    }
    cand[3 * t + k] = prio;
  }
}

__global__ void k_mark_exact(const int32_t *tri, const int32_t *adj, const double *p,
                             uint64_t *claim, uint64_t *cand, const int32_t *queue,
                             int32_t *ctr) {
  if (ctr[7]) return;
  const int round = ctr[6];
  const int n = ctr[3];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    int id = queue[i];
    int64_t t = id / 3;
    int k = id % 3;
    Quad q = quad_of(tri, adj, t, k);
    // only real-real quads reach here (ghost tests use orient2d_filter,
    // whose uncertain cases are also settled exactly below)
    int s;
    if (q.x >= 0 && q.a >= 0 && q.b >= 0) {
      s = smpred::incircle_fix(PT(p, q.x), PT(p, q.a), PT(p, q.b), PT(p, q.y));
      if (s == SM_PRED_UNRESOLVED)
        s = smpred::incircle_exact(PT(p, q.x), PT(p, q.a), PT(p, q.b), PT(p, q.y));
    } else {
      int u, v;
      if (q.x < 0) { u = q.a; v = q.b; }
      else if (q.a < 0) { u = q.b; v = q.x; }
      else { u = q.x; v = q.a; }
      s = smpred::orient2d_fix(PT(p, u), PT(p, v), PT(p, q.y));
      if (s == SM_PRED_UNRESOLVED) s = smpred::orient2d_exact(PT(p, u), PT(p, v), PT(p, q.y));
    }
    if (s == SM_PRED_UNRESOLVED) {
      atomicAdd(&ctr[5], 1);
      continue;
    }
    if (s > 0) {
      uint64_t prio = (1ull << 32) | scramble((uint32_t)id, round);
      cand[id] = prio;
      atomicMax((unsigned long long *)&claim[t], (unsigned long long)prio);
      atomicMax((unsigned long long *)&claim[q.o], (unsigned long long)prio);
    }
  }
}

__global__ void k_apply(int64_t T, int32_t *tri, int32_t *adj, const uint64_t *claim,
                        const uint64_t *cand, int32_t *partner, int32_t *mark, int32_t *ctr,
                        int32_t *next, ActView av, int par) {
  if (next && blockIdx.x == 0 && threadIdx.x < 6) next[threadIdx.x] = 0;
  if (ctr[7]) return;
  const int round = ctr[6];
  int64_t t = (int64_t)xcd_block() * blockDim.x + threadIdx.x;
  // the next round's inverted flags start clear (nothing reads them before)
  if (av.act && t < av.ndom) av.act[(3 - par) * av.ndom + t] = 0;
  if (t >= T) return;
  if (!av.active(t, par)) return;
  for (int k = 0; k < 3; ++k) {
    uint64_t pr = cand[3 * t + k];
    if (!pr) continue;
    int64_t o = adj[3 * t + k];
    if (claim[t] != pr || claim[o] != pr) continue;
    // t = (x, a, b) with neighbours: opp x -> o, opp a -> na, opp b -> nb
    int x = tri[3 * t + k], a = tri[3 * t + (k + 1) % 3], b = tri[3 * t + (k + 2) % 3];
    int na = adj[3 * t + (k + 1) % 3], nb = adj[3 * t + (k + 2) % 3];
    // o = (y, b, a) rotated: find y's slot j
    int j = 0;
    for (int s = 0; s < 3; ++s)
      if (adj[3 * o + s] == t) j = s;
    int y = tri[3 * o + j];
    // in o = (y, b, a): the vertex after y is b, then a
    int mb = adj[3 * o + (j + 1) % 3];  // opposite b: edge (a, y)
    int ma = adj[3 * o + (j + 2) % 3];  // opposite a: edge (y, b)
    // t' = (x, a, y): opp x -> mb (a,y); opp a -> o' (y,x); opp y -> nb (x,a)
    tri[3 * t] = x; tri[3 * t + 1] = a; tri[3 * t + 2] = y;
    adj[3 * t] = mb; adj[3 * t + 1] = (int32_t)o; adj[3 * t + 2] = nb;
    // o' = (y, b, x): opp y -> na (b,x); opp b -> t' (x,y); opp x -> ma (y,b)
    tri[3 * o] = y; tri[3 * o + 1] = b; tri[3 * o + 2] = x;
    adj[3 * o] = na; adj[3 * o + 1] = (int32_t)t; adj[3 * o + 2] = ma;
    partner[t] = (int32_t)o;
    partner[o] = (int32_t)t;
    mark[t] = round;
    mark[o] = round;
    atomicAdd(&ctr[0], 1);
    if (av.act) av.act[par * av.ndom + av.tdom[t]] = 1;
    break;
  }
}

// After a flip batch that did not end valid: flags the domains (one flag
// for a single-domain handle) holding an inverted or degenerate real
// triangle, a triangle flipped by the last round, or two vertices at the
// same position; dup is set for the latter (flips cannot separate them).
__global__ void k_domain_fail(int64_t T, const int32_t *tri, const double *p,
                              const int32_t *mark, const int32_t *status,
                              const int32_t *vdom, int32_t *dflag, int32_t *dup) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  const int32_t a = tri[3 * t], b = tri[3 * t + 1], c = tri[3 * t + 2];
  if (a == kDeadVertex) return;
  const int last = status[6];
  // reason bits: 1 flipped by the last round, 2 inverted or degenerate,
  // 4 two coincident vertices
  int why = (last > 0 && mark[t] == last) ? 1 : 0;
  if (a >= 0 && b >= 0 && c >= 0) {
    const double *pa = PT(p, a), *pb = PT(p, b), *pc = PT(p, c);
    const bool same = (pa[0] == pb[0] && pa[1] == pb[1]) || (pb[0] == pc[0] && pb[1] == pc[1]) ||
                      (pa[0] == pc[0] && pa[1] == pc[1]);
    if (same) {
      atomicOr(dup, 1);
      why |= 4;
    } else {
      int o = smpred::orient2d_filter(pa, pb, pc);
      if (o == SM_PRED_UNRESOLVED) o = smpred::orient2d_if_exact(pa, pb, pc);
      if (o == SM_PRED_UNRESOLVED) o = smpred::orient2d_exact(pa, pb, pc);
      if (o <= 0) why |= 2;
    }
  }
  if (!why) return;
  const int32_t v = a >= 0 ? a : (b >= 0 ? b : c);
  atomicOr(&dflag[vdom ? vdom[v] : 0], why);
}

}  // namespace

__global__ void k_lerp(int64_t n, const double *a, const double *b, double s, double *out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i] + s * (b[i] - a[i]);
}

// Batch handles: sub-step only the vertices of masked domains (the others
// keep their current positions) ...
__global__ void k_lerp_masked(int64_t n2, const double *a, const double *b, double s,
                              double *out, const int32_t *vdom, const int32_t *mask) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n2) return;
  out[i] = mask[vdom[i / 2]] ? a[i] + s * (b[i] - a[i]) : b[i];
}

// ... and copy only their triangle slots (tdom: domain of each slot).
__global__ void k_copy_tri_masked(int64_t T, const int32_t *tri_src, const int32_t *adj_src,
                                  int32_t *tri_dst, int32_t *adj_dst, const int32_t *tdom,
                                  const int32_t *mask) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T || !mask[tdom[t]]) return;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    tri_dst[3 * t + k] = tri_src[3 * t + k];
    adj_dst[3 * t + k] = adj_src[3 * t + k];
  }
}

// Copies the triangulation and the points it is valid for into the
// snapshot buffers, but only when the flip status says the repair
// converged to a valid triangulation (so a speculative graph iteration
// never overwrites the last good state with a tangled one).
// p_any (optional) receives the points whatever the status: the rollback
// copy of a graph iteration, whose points do not move before its update.
__global__ void k_snapshot_if_valid(const int32_t *status, int64_t T, const int32_t *tri,
                                    const int32_t *adj, int32_t *tri_s, int32_t *adj_s,
                                    int64_t n2, const double *p, double *p_s, double *p_any) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  if (p_any)
    for (int64_t k = i; k < n2; k += stride) p_any[k] = p[k];
  if (!(status[2] && status[3] == 0 && status[4] == 0)) return;
  for (int64_t k = i; k < 3 * T; k += stride) {
    tri_s[k] = tri[k];
    adj_s[k] = adj[k];
  }
  for (int64_t k = i; k < n2; k += stride) p_s[k] = p[k];
}

int enqueue_snapshot(sm_mesher *m, const int32_t *status, double *p_any) {
  const int64_t work = std::max<int64_t>(3 * m->T, m->NA * 2);
  unsigned g = blocks_for(work, 256);
  if (g > 2048) g = 2048;
  hipLaunchKernelGGL(k_snapshot_if_valid, dim3(g), dim3(256), 0, m->stream, status, m->T,
                     m->d_tri, m->d_adj, m->d_tri_s, m->d_adj_s, m->NA * 2, m->d_p, m->d_pvalid,
                     p_any);
  SM_LAUNCH_CHECK();
  return SM_OK;
}

namespace {

// Runs flip batches on `pts` until the rounds converge; returns 1 when the
// triangulation is then valid (Delaunay) for `pts`, 0 when it is not and
// -1 when an exact predicate could not be settled.
int repair_on(sm_mesher *m, const double *pts, int max_rounds) {
  hipStream_t s = m->stream;
  int32_t *dstat = m->d_ctr + kCtrStatus;
  for (int done = 0; done < max_rounds;) {
    int rc = enqueue_flip_rounds(m, 4, dstat, pts, true);
    if (rc) return rc > 0 ? -rc : rc - 100;
    if (hipMemcpyAsync(m->h_ctr, dstat, 8 * sizeof(int32_t), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return -2;
    account_flips(m, m->h_ctr);
    done += m->h_ctr[0];
    if (m->h_ctr[4] > 0) return -1;
    if (m->h_ctr[2]) return m->h_ctr[3] == 0 ? 1 : 0;
  }
  return 0;
}

}  // namespace

// Kinetic repair: restore the last valid triangulation (valid for the
// points in d_pvalid) and move the points towards their current positions
// in sub-steps, repairing by flips after each; a failed sub-step is retried
// from the last good state with half the step.  Returns true on success.
static bool trace_flips();

bool kinetic_repair(sm_mesher *m) {
  hipStream_t s = m->stream;
  const int64_t n2 = m->NA * 2;
  double s0 = 0.0, step = 0.5;
  // batches sub-step only the domains flagged by flag_failed_domains
  const bool masked = m->ndom > 0 && m->d_dfail;
  const unsigned gT = blocks_for(m->T, 256);
  auto restore = [&]() {
    if (masked) {
      hipLaunchKernelGGL(k_copy_tri_masked, dim3(gT), dim3(256), 0, s, m->T, m->d_tri_s,
                         m->d_adj_s, m->d_tri, m->d_adj, m->d_tdom, m->d_dfail);
      return;
    }
    (void)hipMemcpyAsync(m->d_tri, m->d_tri_s, m->T * 3 * sizeof(int32_t),
                         hipMemcpyDeviceToDevice, s);
    (void)hipMemcpyAsync(m->d_adj, m->d_adj_s, m->T * 3 * sizeof(int32_t),
                         hipMemcpyDeviceToDevice, s);
  };
  auto save = [&]() {
    if (masked) {
      hipLaunchKernelGGL(k_copy_tri_masked, dim3(gT), dim3(256), 0, s, m->T, m->d_tri,
                         m->d_adj, m->d_tri_s, m->d_adj_s, m->d_tdom, m->d_dfail);
      return;
    }
    (void)hipMemcpyAsync(m->d_tri_s, m->d_tri, m->T * 3 * sizeof(int32_t),
                         hipMemcpyDeviceToDevice, s);
    (void)hipMemcpyAsync(m->d_adj_s, m->d_adj, m->T * 3 * sizeof(int32_t),
                         hipMemcpyDeviceToDevice, s);
  };
  restore();
  int stages = 0;
  while (s0 < 1.0) {
    double s1 = std::min(1.0, s0 + step);
    const double *pts = m->d_p;
    if (s1 < 1.0) {
      if (masked)
        hipLaunchKernelGGL(k_lerp_masked, dim3(blocks_for(n2, 256)), dim3(256), 0, s, n2,
                           m->d_pvalid, m->d_p, s1, m->d_pstage, m->d_vdom, m->d_dfail);
      else
        hipLaunchKernelGGL(k_lerp, dim3(blocks_for(n2, 256)), dim3(256), 0, s, n2,
                           m->d_pvalid, m->d_p, s1, m->d_pstage);
      pts = m->d_pstage;
    }
    const int64_t r0 = m->stats[1];
    int ok = repair_on(m, pts, 64);
    ++stages;
    if (trace_flips())
      std::fprintf(stderr, "kinetic_repair: stage %d, %.4f -> %.4f: %s after %lld rounds\n",
                   stages, s0, s1, ok == 1 ? "ok" : "failed", (long long)(m->stats[1] - r0));
    if (ok == 1) {
      s0 = s1;
      if (s0 < 1.0) save();
      step = std::min(2.0 * step, 1.0);
    } else {
      if (ok < 0 || step < 1.0 / 512 || stages > 200) return false;
      restore();
      // half the sub-step just tried (the step may exceed what was left
      // before the end, and halving it alone would try the same one again)
      step = 0.5 * (s1 - s0);
    }
  }
  m->stats[6] += 1;
  return true;
}

int time_flip_mark(sm_mesher *m, int reps, double *ms) {
  SM_CHECK(m->dim == 2 && m->have_tri && m->T > 0, SM_ERR_STATE,
           "time_kernel: needs a 2D handle with a triangulation");
  SM_CHECK(reps >= 1 && reps <= 100000, SM_ERR_ARG, "time_kernel: bad repetition count");
  hipStream_t s = m->stream;
  if (!m->d_tctr) SM_HIP(hipMalloc((void **)&m->d_tctr, 8 * sizeof(int32_t)));
  // The launches start from the counters of an open round (not finished,
  // round id 0, no inverted or uncertain triangle: the loop's usual state,
  // in which k_mark does not read orientations) and run back to back.  An
  // uncertain incircle test counted by one launch would switch every later
  // one to the tangled path, which also loads the orientations: then every
  // launch is timed alone from a cleared counter block (its event pair adds
  // a few microseconds).  The launches write claims and candidates, which
  // every flip round recomputes before use.
  const int64_t T = m->T;
  const int B = m->blk[kBlkMark];
  const unsigned g = blocks_for(T, B);
  hipEvent_t e0, e1;
  SM_HIP(hipEventCreate(&e0));
  SM_HIP(hipEventCreate(&e1));
  int rc = fill<int32_t>(m->d_tctr, 8, 0, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_mark, dim3(g), dim3(B), 0, s, T, m->d_tri, m->d_adj, m->d_ori, m->d_p,
                     m->d_claim, m->d_cand, m->d_queue + T, m->d_tctr, 0, ActView(), 0);
  SM_HIP(hipEventRecord(e0, s));
  for (int i = 0; i < reps; ++i)
    hipLaunchKernelGGL(k_mark, dim3(g), dim3(B), 0, s, T, m->d_tri, m->d_adj, m->d_ori, m->d_p,
                       m->d_claim, m->d_cand, m->d_queue + T, m->d_tctr, 0, ActView(), 0);
  SM_HIP(hipEventRecord(e1, s));
  SM_LAUNCH_CHECK();
  SM_HIP(hipEventSynchronize(e1));
  float back = 0.0f;
  SM_HIP(hipEventElapsedTime(&back, e0, e1));
  int32_t tc[8];
  SM_HIP(hipMemcpy(tc, m->d_tctr, sizeof(tc), hipMemcpyDeviceToHost));
  if (tc[5] == 0) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = (double)back / reps;
    return SM_OK;
  }
  double total = 0.0;
  for (int i = 0; i <= reps && rc == SM_OK; ++i) {
    if ((rc = fill<int32_t>(m->d_tctr, 8, 0, s))) break;
    if (hipEventRecord(e0, s) != hipSuccess) rc = sm_set_error(SM_ERR_HIP, "time_kernel: event");
    hipLaunchKernelGGL(k_mark, dim3(g), dim3(B), 0, s, T, m->d_tri, m->d_adj, m->d_ori, m->d_p,
                       m->d_claim, m->d_cand, m->d_queue + T, m->d_tctr, 0, ActView(), 0);
    if (hipGetLastError() != hipSuccess || hipEventRecord(e1, s) != hipSuccess ||
        hipEventSynchronize(e1) != hipSuccess)
      rc = sm_set_error(SM_ERR_HIP, "time_kernel: launch or event failed");
    float el = 0.0f;
    if (rc == SM_OK && hipEventElapsedTime(&el, e0, e1) != hipSuccess)
      rc = sm_set_error(SM_ERR_HIP, "time_kernel: elapsed time");
    if (i > 0) total += el;  // the first launch warms up
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  *ms = total / reps;
  return SM_OK;
}

int time_flip_tunable(sm_mesher *m, int which, int B, int reps, double *ms) {
  SM_CHECK(which == kBlkOrient || which == kBlkMark || which == kBlkApply, SM_ERR_ARG,
           "time_flip_tunable: kernel %d", which);
  SM_CHECK(m->dim == 2 && m->have_tri && m->T > 0, SM_ERR_STATE,
           "time_flip_tunable: needs a 2D handle with a triangulation");
  SM_CHECK(B >= 64 && B <= 1024 && B % 64 == 0 && reps >= 1, SM_ERR_ARG,
           "time_flip_tunable: bad block size or repetition count");
  hipStream_t s = m->stream;
  if (!m->d_tctr) SM_HIP(hipMalloc((void **)&m->d_tctr, 8 * sizeof(int32_t)));
  // a private open round (uncertain predicates only counted): the
  // orientation pass recomputes orientations and clears every claim, so
  // the apply pass that follows finds no claimed candidate and flips
  // nothing; the next flip round recomputes all of it
  int rc = fill<int32_t>(m->d_tctr, 8, 0, s);
  if (rc) return rc;
  const int64_t T = m->T;
  auto launch = [&](int k, int b) {
    const unsigned g = blocks_for(T, b);
    if (k == kBlkOrient)
      hipLaunchKernelGGL(k_fix_orient, dim3(g), dim3(b), 0, s, T, m->d_tri, m->d_adj,
                         m->d_partner, m->d_mark, m->d_p, m->d_ori, m->d_claim, m->d_queue,
                         (const int32_t *)nullptr, m->d_tctr, 0, ActView(), 0);
    else if (k == kBlkMark)
      hipLaunchKernelGGL(k_mark, dim3(g), dim3(b), 0, s, T, m->d_tri, m->d_adj, m->d_ori,
                         m->d_p, m->d_claim, m->d_cand, m->d_queue + T, m->d_tctr, 0, ActView(),
                         0);
    else
      hipLaunchKernelGGL(k_apply, dim3(g), dim3(b), 0, s, T, m->d_tri, m->d_adj, m->d_claim,
                         m->d_cand, m->d_partner, m->d_mark, m->d_tctr, (int32_t *)nullptr,
                         ActView(), 0);
  };
  launch(kBlkOrient, 256);  // orientations of the current points, claims cleared
  hipEvent_t e0, e1;
  SM_HIP(hipEventCreate(&e0));
  SM_HIP(hipEventCreate(&e1));
  launch(which, B);
  SM_HIP(hipEventRecord(e0, s));
  for (int i = 0; i < reps; ++i) launch(which, B);
  SM_HIP(hipEventRecord(e1, s));
  if (which == kBlkMark) launch(kBlkOrient, 256);  // clear the claims again
  SM_LAUNCH_CHECK();
  SM_HIP(hipEventSynchronize(e1));
  SM_HIP(hipStreamSynchronize(s));
  float el = 0.0f;
  SM_HIP(hipEventElapsedTime(&el, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = (double)el / reps;
  return SM_OK;
}

int enqueue_flip_rounds(sm_mesher *m, int nrounds, int32_t *status, const double *pts,
                        bool exact) {
  SM_CHECK(nrounds >= 1 && nrounds <= kMaxBatchRounds, SM_ERR_ARG,
           "flip rounds: batch of %d rounds", nrounds);
  hipStream_t s = m->stream;
  const int64_t T = m->T;
  const int B = 256;
  const unsigned g = blocks_for(T, B);
  const int Bo = m->blk[kBlkOrient], Bm = m->blk[kBlkMark], Ba = m->blk[kBlkApply];
  const unsigned go = blocks_for(T, Bo), gm = blocks_for(T, Bm), ga = blocks_for(T, Ba);
  int32_t *base = m->d_ctr + kCtrRounds;
  int32_t *rc = m->d_ctr + kCtrRoundId;
  const int qx = exact ? 1 : 0;
  ActView av;
  if (m->ndom > 0 && m->d_act) {
    av.tdom = m->d_tdom;
    av.act = m->d_act;
    av.ndom = m->ndom;
    hipLaunchKernelGGL(k_act_begin, dim3(blocks_for(m->ndom, 256)), dim3(256), 0, s, m->d_act,
                       m->ndom);
  }
  hipLaunchKernelGGL(k_batch_begin, dim3(1), dim3(1), 0, s, base, (const int32_t *)rc);
  for (int r = 0; r < nrounds; ++r) {
    int32_t *c = base + 8 * r;
    const int32_t *prev = r == 0 ? nullptr : c - 8;
    int32_t *next = r + 1 < nrounds ? c + 8 : nullptr;
    const int par = r % 2;
    hipLaunchKernelGGL(k_fix_orient, dim3(go), dim3(Bo), 0, s, T, m->d_tri, m->d_adj,
                       m->d_partner, m->d_mark, pts, m->d_ori, m->d_claim, m->d_queue, prev, c,
                       qx, av, par);
    if (exact)
      hipLaunchKernelGGL(k_orient_exact, dim3(8), dim3(64), 0, s, m->d_tri, pts, m->d_ori,
                         m->d_queue, c);
    const int pm = prof_begin(m);
    hipLaunchKernelGGL(k_mark, dim3(gm), dim3(Bm), 0, s, T, m->d_tri, m->d_adj, m->d_ori, pts,
                       m->d_claim, m->d_cand, m->d_queue + T, c, qx, av, par);
    prof_end(m, 7, pm);
    if (exact)
      hipLaunchKernelGGL(k_mark_exact, dim3(8), dim3(64), 0, s, m->d_tri, m->d_adj, pts,
                         m->d_claim, m->d_cand, m->d_queue + T, c);
    hipLaunchKernelGGL(k_apply, dim3(ga), dim3(Ba), 0, s, T, m->d_tri, m->d_adj, m->d_claim,
                       m->d_cand, m->d_partner, m->d_mark, c, next, av, par);
    SM_LAUNCH_CHECK();
  }
  const int32_t *lastc = base + 8 * (nrounds - 1);
  hipLaunchKernelGGL(k_fix_last, dim3(g), dim3(B), 0, s, T, m->d_tri, m->d_adj, m->d_partner,
                     m->d_mark, lastc, (const int32_t *)base, nrounds, status, rc);
  SM_LAUNCH_CHECK();
  return SM_OK;
}

// SMGPU_TRACE_FLIPS=1 prints the outcome of every batch of repair rounds
static bool trace_flips() {
  static const bool on = [] {
    const char *e = std::getenv("SMGPU_TRACE_FLIPS");
    return e && *e && *e != '0';
  }();
  return on;
}

static double wall_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

void account_flips(sm_mesher *m, const int32_t *st) {
  m->stats[1] += st[0];
  m->stats[2] += st[1];
  m->stats[4] += st[5];
}

int flip_repair2d(sm_mesher *m, int max_rounds, bool *need_rebuild) {
  hipStream_t s = m->stream;
  *need_rebuild = false;
  int32_t *dstat = m->d_ctr + kCtrStatus;
  bool ok = false, exact_failed = false;
  int best_inverted = INT32_MAX, stale = 0;
  // A move of less than half an edge length leaves local tangles (a vertex
  // slid across a flat triangle along a straight boundary, or driven almost
  // onto a corner) that the host relocation below resolves in one pass
  // where untangling flips need dozens of rounds: with it available, the
  // flips get two batches.
  const double h0m = m->ndom > 0 ? 1.0 : m->prm.h0;
  const bool large_move = !(m->last_maxdp < 0.5 * h0m);
  // (single-domain handles only: in a batch the failed domains are rebuilt
  // on the host in parallel, which measured cheaper -- 465 against 695 ms
  // per iteration on 512 domains -- than relocation followed by the long
  // Lawson cascades that the ear-clipped stars start)
  const auto relocatable = [&]() {
    return m->ndom == 0 && !large_move && m->have_snap && m->h_ctr[3] > 0 &&
           m->h_ctr[3] <= 4096;
  };
  const double t_start = wall_ms();
  for (int done = 0; done < max_rounds;) {
    const int batch = 4;
    const int pb = prof_begin(m);
    int rc = enqueue_flip_rounds(m, batch, dstat, m->d_p, true);
    if (rc) return rc;
    prof_end(m, 0, pb);
    SM_HIP(hipMemcpyAsync(m->h_ctr, dstat, 8 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    SM_HIP(hipStreamSynchronize(s));
    account_flips(m, m->h_ctr);
    done += m->h_ctr[0];
    if (trace_flips())
      std::fprintf(stderr, "flip_repair2d: rounds %d flips %d last-round-idle %d inverted %d "
                   "exact %d\n", m->h_ctr[0], m->h_ctr[1], m->h_ctr[2], m->h_ctr[3],
                   m->h_ctr[5]);
    if (m->h_ctr[4] > 0) {  // an exact predicate did not fit: rebuild exactly
      m->stats[7] += m->h_ctr[4];
      exact_failed = true;
      break;
    }
    if (m->h_ctr[2]) {
      // the last round flipped nothing: every edge is locally Delaunay
      // unless inverted triangles are left that flips could not untangle
      ok = m->h_ctr[3] == 0;
      break;
    }
    if (done >= 8 && relocatable()) break;
    // inverted triangles that stop decreasing (by at least 5 % per batch)
    // will not be untangled by more rounds
    if (m->h_ctr[3] > 0) {
      if (m->h_ctr[3] < best_inverted - best_inverted / 20) {
        best_inverted = m->h_ctr[3];
        stale = 0;
      } else if (++stale >= 3) {
        break;
      }
    }
  }
  m->rebuild_flagged = false;
  const double t_flips = wall_ms();
  // Relocate one vertex of each inverted triangle on the host, then flip
  // again.  Large moves (the first iterations) go to the kinetic repair
  // below.
  for (int pass = 0; pass < 3 && !ok && !exact_failed && relocatable(); ++pass) {
    int64_t moved = 0;
    const double t_r0 = wall_ms();
    int rc = relocate_inverted(m, &moved);
    if (rc) return rc;
    if (trace_flips())
      std::fprintf(stderr, "flip_repair2d: relocation pass %d moved %lld vertices in %.1f ms\n",
                   pass, (long long)moved, wall_ms() - t_r0);
    if (moved <= 0) break;
    int best = INT32_MAX, best_fl = INT32_MAX, idle = 0;
    for (int done = 0; done < 4 * max_rounds;) {
      rc = enqueue_flip_rounds(m, 4, dstat, m->d_p, true);
      if (rc) return rc;
      SM_HIP(hipMemcpyAsync(m->h_ctr, dstat, 8 * sizeof(int32_t), hipMemcpyDeviceToHost, s));
      SM_HIP(hipStreamSynchronize(s));
      account_flips(m, m->h_ctr);
      done += m->h_ctr[0];
      if (trace_flips())
        std::fprintf(stderr, "flip_repair2d: after relocation: rounds %d flips %d "
                     "last-round-idle %d inverted %d\n", m->h_ctr[0], m->h_ctr[1], m->h_ctr[2],
                     m->h_ctr[3]);
      if (m->h_ctr[4] > 0) {
        exact_failed = true;
        break;
      }
      if (m->h_ctr[2]) {
        ok = m->h_ctr[3] == 0;
        break;
      }
      // stop when neither the inverted triangles nor the flips per batch
      // decrease any more (Lawson flips alone always finish)
      if (m->h_ctr[3] > 0) {
        const bool fewer_inv = m->h_ctr[3] < best - best / 20;
        const bool fewer_flips = m->h_ctr[1] < best_fl - best_fl / 20;
        if (fewer_inv) best = m->h_ctr[3];
        if (fewer_flips) best_fl = m->h_ctr[1];
        if (fewer_inv || fewer_flips) idle = 0;
        else if (++idle >= 3) break;
      }
    }
  }
  const double t_reloc = wall_ms();
  if (trace_flips())
    std::fprintf(stderr, "flip_repair2d: %.1f ms flips, %.1f ms relocation passes, ok %d\n",
                 t_flips - t_start, t_reloc - t_flips, (int)ok);
  if (!ok && !exact_failed) {
    // which domains failed, and whether coincident vertices are the cause
    // (flips and sub-steps cannot separate them; the host rebuild drops
    // the duplicate, as the reference's CGAL insert does)
    int64_t bad_points = 0;
    bool dup = false;
    int rc = flag_failed_domains(m, &bad_points, &dup);
    if (rc) return rc;
    // kinetic repair: restart the failed domains from their last valid
    // state and move their points in sub-steps.  A batch with only a few
    // failed domains rebuilds those on the host straight away (the usual
    // cause there is a vertex projected next to a fixed corner, which the
    // sub-steps cannot get past either)
    // Sub-steps pay off for large moves (the first iterations); after a
    // move of less than half an edge length the failure is a local
    // degeneracy (a vertex driven onto a fixed corner or a straight piece
    // of boundary) and the host rebuild is cheaper.
    const bool kinetic = !dup && m->have_snap && large_move &&
                         (m->ndom == 0 || 8 * bad_points > m->N);
    if (kinetic) ok = kinetic_repair(m);
    // rebuild on the host only the flagged domains (the others were valid
    // and the kinetic repair left them alone)
    if (!ok && m->ndom > 0) m->rebuild_flagged = true;
  }
  if (!ok) {
    *need_rebuild = true;
    return SM_OK;
  }
  return snapshot_valid(m);
}

// Host relocation (relocate2d.cpp) of a vertex of every inverted triangle:
// the whole triangulation for one domain, the failed domains' slot ranges
// for a batch.  *moved: relocations done.
int relocate_inverted(sm_mesher *m, int64_t *moved) {
  hipStream_t s = m->stream;
  *moved = 0;
  std::vector<std::pair<int64_t, int64_t>> ranges;
  if (m->ndom > 0) {
    int64_t bad = 0;
    bool dup = false;
    int rc = flag_failed_domains(m, &bad, &dup);
    if (rc) return rc;
    for (int d = 0; d < m->ndom; ++d)
      if (m->h_dfail[d]) ranges.push_back({m->h_tbase[d], m->h_tbase[d + 1] - m->h_tbase[d]});
  } else {
    ranges.push_back({0, m->T});
  }
  if (ranges.empty()) return SM_OK;
  const int64_t NA = m->NA;
  if ((int64_t)m->h_prel.size() != 4 * NA) m->h_prel.assign(4 * NA, 0.0);
  double *hp = m->h_prel.data(), *hold = hp + 2 * NA;
  if (m->ndom > 0) {
    // points of the failed domains (and their far vertices)
    for (int d = 0; d < m->ndom; ++d) {
      if (!m->h_dfail[d]) continue;
      const int64_t b = m->h_voff[d], n = m->h_voff[d + 1] - b, f = m->N + 4 * d;
      SM_HIP(hipMemcpyAsync(hp + 2 * b, m->d_p + 2 * b, 2 * n * sizeof(double),
                            hipMemcpyDeviceToHost, s));
      SM_HIP(hipMemcpyAsync(hold + 2 * b, m->d_pvalid + 2 * b, 2 * n * sizeof(double),
                            hipMemcpyDeviceToHost, s));
      if (m->nfar == 0) continue;
      SM_HIP(hipMemcpyAsync(hp + 2 * f, m->d_p + 2 * f, 8 * sizeof(double),
                            hipMemcpyDeviceToHost, s));
      SM_HIP(hipMemcpyAsync(hold + 2 * f, m->d_pvalid + 2 * f, 8 * sizeof(double),
                            hipMemcpyDeviceToHost, s));
    }
  } else {
    SM_HIP(hipMemcpyAsync(hp, m->d_p, 2 * NA * sizeof(double), hipMemcpyDeviceToHost, s));
    SM_HIP(hipMemcpyAsync(hold, m->d_pvalid, 2 * NA * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  // all ranges in one buffer: copy down, relocate in parallel (ranges are
  // disjoint and sm_relocate2d only reads the points), copy back
  std::vector<int64_t> off(ranges.size() + 1, 0);
  for (size_t k = 0; k < ranges.size(); ++k) off[k + 1] = off[k] + ranges[k].second;
  std::unique_ptr<int32_t[]> tri(new int32_t[3 * off.back()]), adj(new int32_t[3 * off.back()]);
  for (size_t k = 0; k < ranges.size(); ++k) {
    const auto &r = ranges[k];
    SM_HIP(hipMemcpyAsync(tri.get() + 3 * off[k], m->d_tri + 3 * r.first,
                          3 * r.second * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    SM_HIP(hipMemcpyAsync(adj.get() + 3 * off[k], m->d_adj + 3 * r.first,
                          3 * r.second * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  }
  SM_HIP(hipStreamSynchronize(s));
  std::vector<int64_t> done(ranges.size(), 0);
  std::atomic<int> next{0};
  auto work = [&]() {
    for (int k; (k = next.fetch_add(1)) < (int)ranges.size();) {
      int64_t failed = 0;
      done[k] = sm_relocate2d(hp, hold, tri.get() + 3 * off[k], adj.get() + 3 * off[k],
                              ranges[k].first, ranges[k].second, &failed);
    }
  };
  unsigned nw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  nw = std::min<unsigned>(nw, (unsigned)ranges.size());
  std::vector<std::thread> pool;
  for (unsigned t = 1; t < nw; ++t) pool.emplace_back(work);
  work();
  for (auto &t : pool) t.join();
  for (size_t k = 0; k < ranges.size(); ++k) {
    if (done[k] <= 0) continue;
    *moved += done[k];
    m->stats[8] += done[k];
    const auto &r = ranges[k];
    SM_HIP(hipMemcpyAsync(m->d_tri + 3 * r.first, tri.get() + 3 * off[k],
                          3 * r.second * sizeof(int32_t), hipMemcpyHostToDevice, s));
    SM_HIP(hipMemcpyAsync(m->d_adj + 3 * r.first, adj.get() + 3 * off[k],
                          3 * r.second * sizeof(int32_t), hipMemcpyHostToDevice, s));
  }
  SM_HIP(hipStreamSynchronize(s));
  return SM_OK;
}

// Runs k_domain_fail on the current triangulation; fills m->h_dfail and the
// number of points of the flagged domains.
int flag_failed_domains(sm_mesher *m, int64_t *bad_points, bool *dup) {
  hipStream_t s = m->stream;
  const int nd = std::max(m->ndom, 1);
  if ((int)m->h_dfail.size() != nd) m->h_dfail.assign(nd, 0);
  if (!m->d_dfail) SM_HIP(hipMalloc(&m->d_dfail, (nd + 1) * sizeof(int32_t)));
  int rc = fill<int32_t>(m->d_dfail, nd + 1, 0, s);
  if (rc) return rc;
  const int B = 256;
  hipLaunchKernelGGL(k_domain_fail, dim3(blocks_for(m->T, B)), dim3(B), 0, s, m->T, m->d_tri,
                     m->d_p, m->d_mark, (const int32_t *)(m->d_ctr + kCtrStatus), m->d_vdom,
                     m->d_dfail, m->d_dfail + nd);
  SM_LAUNCH_CHECK();
  std::vector<int32_t> fl(nd + 1);
  SM_HIP(hipMemcpyAsync(fl.data(), m->d_dfail, (nd + 1) * sizeof(int32_t),
                        hipMemcpyDeviceToHost, s));
  SM_HIP(hipStreamSynchronize(s));
  *dup = fl[nd] != 0;
  *bad_points = 0;
  int nwhy[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int d = 0; d < nd; ++d) {
    m->h_dfail[d] = fl[d] != 0;
    if (fl[d]) *bad_points += m->ndom > 0 ? m->h_voff[d + 1] - m->h_voff[d] : m->N;
    nwhy[fl[d] & 7] += 1;
  }
  if (trace_flips()) {
    std::fprintf(stderr, "failed domains by reason (1 flipping, 2 inverted, 4 coincident):");
    for (int k = 1; k < 8; ++k)
      if (nwhy[k]) std::fprintf(stderr, " [%d]=%d", k, nwhy[k]);
    std::fprintf(stderr, "; first:");
    int shown = 0;
    for (int d = 0; d < nd && shown < 12; ++d)
      if (fl[d]) {
        std::fprintf(stderr, " %d(n=%lld,why=%d)", d,
                     (long long)(m->ndom > 0 ? m->h_voff[d + 1] - m->h_voff[d] : m->N), fl[d]);
        ++shown;
      }
    std::fprintf(stderr, "\n");
  }
  return SM_OK;
}

// Records the current triangulation and points as the last valid state.
int snapshot_valid(sm_mesher *m) {
  hipStream_t s = m->stream;
  int32_t *dstat = m->d_ctr + kCtrStatus;
  int32_t *valid = m->h_ctr + kCtrScratchHost;  // pinned staging
  const int32_t v[8] = {1, 0, 1, 0, 0, 0, 0, 0};
  for (int i = 0; i < 8; ++i) valid[i] = v[i];
  SM_HIP(hipMemcpyAsync(dstat, valid, 8 * sizeof(int32_t), hipMemcpyHostToDevice, s));
  int rc = enqueue_snapshot(m, dstat);
  if (rc) return rc;
  SM_HIP(hipStreamSynchronize(s));
  m->have_snap = true;
  return SM_OK;
}

// Batch handles: each domain's triangulated sphere is built on its own
// vertex range (worker threads over the domains) and the pieces are
// concatenated with their vertex and triangle ids offset, so no triangle
// or neighbour link ever joins two domains.  Domain d of `todo` is written
// at slot offset base[k] of tri / adj (its device slots start at
// h_tbase[d]; ids are device ids).  hp holds the points (2 NA doubles).
static int64_t batch_delaunay(sm_mesher *m, const double *hp, int32_t *tri, int32_t *adj,
                              const std::vector<int> &todo, const std::vector<int64_t> &base) {
  std::vector<int64_t> cnt(todo.size(), 0);
  std::atomic<int> next{0};
  const int64_t N = m->N;
  auto work = [&]() {
    std::vector<uint8_t> ins;
    std::vector<double> loc;
    for (int k; (k = next.fetch_add(1)) < (int)todo.size();) {
      const int d = todo[k];
      const int64_t b = m->h_voff[d], n = m->h_voff[d + 1] - b;
      // the domain's points, then its four far vertices if it has them
      // (otherwise its convex hull, ringed by ghost triangles)
      loc.assign(hp + 2 * b, hp + 2 * (b + n));
      const int64_t nf = m->nfar > 0 ? 4 : 0;
      if (nf) loc.insert(loc.end(), hp + 2 * (N + 4 * d), hp + 2 * (N + 4 * d + 4));
      const int64_t na = n + nf, tb = m->h_tbase[d], cap = m->h_tbase[d + 1] - tb;
      int32_t *td = tri + 3 * base[k], *ad = adj + 3 * base[k];
      ins.resize(na);
      int64_t c = sm_host_delaunay2d(loc.data(), na, td, ad, cap, ins.data());
      cnt[k] = c;
      if (c < 0) continue;
      for (int64_t j = 0; j < 3 * c; ++j) {
        const int32_t v = td[j];
        if (v >= 0) td[j] = v < n ? (int32_t)(b + v) : (int32_t)(N + 4 * d + (v - n));
        ad[j] += (int32_t)tb;
      }
      for (int64_t t = c; t < cap; ++t)
        for (int j = 0; j < 3; ++j) {
          td[3 * t + j] = kDeadVertex;
          ad[3 * t + j] = (int32_t)(tb + t);
        }
    }
  };
  unsigned nw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  nw = std::min<unsigned>(nw, (unsigned)std::max<size_t>(todo.size(), 1));
  std::vector<std::thread> pool;
  for (unsigned i = 1; i < nw; ++i) pool.emplace_back(work);
  work();
  for (auto &t : pool) t.join();
  for (size_t k = 0; k < todo.size(); ++k)
    if (cnt[k] < 0) return cnt[k];
  return m->h_tbase[m->ndom];
}

// Batch handles: rebuild only the domains flagged by flag_failed_domains
// from their own points (copied per domain) and upload their slot ranges;
// the other domains keep their (valid) triangulations.
static int rebuild_flagged_domains(sm_mesher *m) {
  m->rebuild_flagged = false;
  const double t0 = wall_ms();
  hipStream_t s = m->stream;
  std::vector<int> todo;
  std::vector<int64_t> base;
  int64_t slots = 0;
  for (int d = 0; d < m->ndom; ++d)
    if (m->h_dfail[d]) todo.push_back(d);
  // largest domains first, so that the worker threads finish together
  std::stable_sort(todo.begin(), todo.end(), [&](int a, int b) {
    return m->h_voff[a + 1] - m->h_voff[a] > m->h_voff[b + 1] - m->h_voff[b];
  });
  for (int d : todo) {
    base.push_back(slots);
    slots += m->h_tbase[d + 1] - m->h_tbase[d];
  }
  if (todo.empty()) return snapshot_valid(m);
  const int64_t NA = m->NA;
  if ((int64_t)m->h_prel.size() != 4 * NA) m->h_prel.assign(4 * NA, 0.0);
  double *hp = m->h_prel.data();
  for (int d : todo) {
    const int64_t b = m->h_voff[d], n = m->h_voff[d + 1] - b, f = m->N + 4 * d;
    SM_HIP(hipMemcpyAsync(hp + 2 * b, m->d_p + 2 * b, 2 * n * sizeof(double),
                          hipMemcpyDeviceToHost, s));
    if (m->nfar > 0)
      SM_HIP(hipMemcpyAsync(hp + 2 * f, m->d_p + 2 * f, 8 * sizeof(double),
                            hipMemcpyDeviceToHost, s));
  }
  SM_HIP(hipStreamSynchronize(s));
  std::unique_ptr<int32_t[]> tri(new int32_t[3 * slots]), adj(new int32_t[3 * slots]);
  int64_t T = batch_delaunay(m, hp, tri.get(), adj.get(), todo, base);
  if (T < 0) return sm_set_error((int)T, "host Delaunay rebuild failed (%lld)", (long long)T);
  for (size_t k = 0; k < todo.size(); ++k) {
    const int d = todo[k];
    const int64_t b = m->h_tbase[d], n = m->h_tbase[d + 1] - b;
    SM_HIP(hipMemcpyAsync(m->d_tri + 3 * b, tri.get() + 3 * base[k], 3 * n * sizeof(int32_t),
                          hipMemcpyHostToDevice, s));
    SM_HIP(hipMemcpyAsync(m->d_adj + 3 * b, adj.get() + 3 * base[k], 3 * n * sizeof(int32_t),
                          hipMemcpyHostToDevice, s));
    m->stats[3] += 1;
  }
  SM_HIP(hipStreamSynchronize(s));
  const double t1 = wall_ms();
  int rc = snapshot_valid(m);
  if (trace_flips())
    std::fprintf(stderr, "rebuild_flagged_domains: %zu domains, %lld slots: %.1f ms rebuild, "
                 "%.1f ms snapshot\n", todo.size(), (long long)slots, t1 - t0, wall_ms() - t1);
  return rc;
}

int host_rebuild2d(sm_mesher *m) {
  if (m->ndom > 0 && m->rebuild_flagged && m->have_tri) return rebuild_flagged_domains(m);
  const int64_t NA = m->NA;
  std::unique_ptr<double[]> hpbuf(new double[2 * NA]);
  double *hp = hpbuf.get();
  SM_HIP(hipMemcpyAsync(hp, m->d_p, NA * 2 * sizeof(double), hipMemcpyDeviceToHost, m->stream));
  SM_HIP(hipStreamSynchronize(m->stream));
  // uninitialised: every slot up to T is written by the triangulation
  std::unique_ptr<int32_t[]> trib(new int32_t[3 * m->capT]), adjb(new int32_t[3 * m->capT]);
  int32_t *tri = trib.get(), *adj = adjb.get();
  int64_t T = 0;
  if (m->ndom == 0) {
    std::vector<uint8_t> ins((size_t)NA);
    T = sm_host_delaunay2d(hp, NA, tri, adj, m->capT, ins.data());
  } else {
    // fixed slot ranges: domain d owns 2 (n_d + 4) slots from h_tbase[d]
    // (2 n_d + 6 triangles inside a far square, 2 n_d - 2 real and ghost
    // triangles over its convex hull; the rest are dead)
    if ((int)m->h_tbase.size() != m->ndom + 1) {
      m->h_tbase.assign(m->ndom + 1, 0);
      std::vector<int32_t> tdom;
      tdom.reserve(m->capT);
      for (int d = 0; d < m->ndom; ++d) {
        const int64_t sz = 2 * (m->h_voff[d + 1] - m->h_voff[d] + 4);
        m->h_tbase[d + 1] = m->h_tbase[d] + sz;
        tdom.insert(tdom.end(), sz, d);
      }
      SM_CHECK(m->h_tbase[m->ndom] <= m->capT, SM_ERR_CAPACITY, "batch slots exceed capacity");
      SM_HIP(hipMemcpy(m->d_tdom, tdom.data(), tdom.size() * sizeof(int32_t),
                       hipMemcpyHostToDevice));
    }
    std::vector<int> todo(m->ndom);
    std::vector<int64_t> base(m->ndom);
    for (int d = 0; d < m->ndom; ++d) {
      todo[d] = d;
      base[d] = m->h_tbase[d];
    }
    T = batch_delaunay(m, hp, tri, adj, todo, base);
  }
  if (T < 0) return sm_set_error((int)T, "host Delaunay rebuild failed (%lld)", (long long)T);
  m->T = T;
  SM_HIP(hipMemcpyAsync(m->d_tri, tri, T * 3 * sizeof(int32_t), hipMemcpyHostToDevice,
                        m->stream));
  SM_HIP(hipMemcpyAsync(m->d_adj, adj, T * 3 * sizeof(int32_t), hipMemcpyHostToDevice,
                        m->stream));
  int rc = fill<int32_t>(m->d_mark, m->capT, 0, m->stream);
  if (rc) return rc;
  SM_HIP(hipStreamSynchronize(m->stream));
  m->stats[3] += 1;
  m->have_tri = true;
  return snapshot_valid(m);
}

}  // namespace smk
