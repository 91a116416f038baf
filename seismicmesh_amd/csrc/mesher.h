// Device state of one force-loop handle (one domain, or a batch of
// independent 2D domains).
#pragma once

#include <algorithm>
#include <vector>

#include "common.h"
#include "delaunay3d.h"
#include "sizing_eval.h"

constexpr int kNumStats = 13;

constexpr double kFarSquare = 100.0;

struct sm_mesher {
  sm_mesher_params prm{};
  int dim = 2;
  int device = 0;
  hipStream_t stream = nullptr;

  // SDF programs, one per level (domain first, then subdomains)
  std::vector<sm_node *> d_tapes;
  std::vector<int> tape_len;
  sm_node *d_tape_all = nullptr;  // concatenated, with offsets below
  std::vector<int> tape_off;
  int *d_tape_off_dev = nullptr, *d_tape_len_dev = nullptr;
  bool deep_sdf = false;  // some program needs the scratch-stack interpreter

  // batch of independent 2D domains (ndom > 0): domain d owns the vertex
  // range [voff[d], voff[d+1]), the SDF program tape_off[d] (tapes above,
  // one per domain instead of one per level), one uniform size h[d] (= its
  // h0) and fixes the first nfix[d] vertices of its range
  int ndom = 0;
  std::vector<int64_t> h_voff;
  std::vector<double> h_dh;
  int32_t *d_vdom = nullptr;                      // (NA) domain of each vertex
  int32_t *d_tdom = nullptr;                      // (capT) domain of each slot
  std::vector<int64_t> h_tbase;                   // (ndom + 1) slot ranges
  int32_t *d_voff = nullptr, *d_dnfix = nullptr;  // (ndom + 1), (ndom)
  double *d_dh = nullptr;                         // (ndom)
  double *d_dscale = nullptr;                     // (ndom) L0 scale per iteration
  // domains whose triangulation the flips could not repair, and whether the
  // next host rebuild may be limited to them
  int32_t *d_dfail = nullptr;
  std::vector<uint8_t> h_dfail;
  // (4 ndom) per-domain activity flags of the flip rounds (flip2d.hip)
  int32_t *d_act = nullptr;
  bool rebuild_flagged = false;
  std::vector<double> h_prel;  // host points for relocations (p, then pvalid)

  // host sizing callback (for sizing given as a Python callable): called
  // with the edge midpoints, fills h; the device path then uses d_hext
  sm_host_sizing_fn host_sizing = nullptr;
  void *host_sizing_ctx = nullptr;
  double *d_mid = nullptr, *d_hext = nullptr;

  // host SDF callback (domains given as Python functions): the centroid
  // filter and the projection run on the host through it (forces.hip)
  sm_host_sdf_fn host_sdf = nullptr;
  void *host_sdf_ctx = nullptr;

  // sizing grid (device copies of axes and values)
  smk::GridView grid{};
  double *d_axis[3] = {nullptr, nullptr, nullptr};
  float *d_values = nullptr;

  // points: N moving points, then (2D batches) nfar fixed far vertices --
  // four per domain, at the corners of a square 2 kFarSquare times the
  // domain's extent -- so each domain's triangulated region is a fixed
  // square instead of the convex hull of its points (whose straight,
  // nearly collinear boundaries would otherwise need long sequential chains
  // of hull flips).  Only triangles with circumcircles wider than
  // ~kFarSquare extents differ from the hull's triangulation: nearly
  // collinear triples on the hull.  Where the initial lattice's outer row
  // lies inside a curved boundary, such flat triangles can pass the
  // centroid filter in the first iteration (circles of 100 - 5000 extents
  // for a row moved by 1e-3 of a spacing), and the reference keeps them; by
  // the next iteration the boundary points are on the boundary and those
  // triangles fail the filter.  A square of 1e6 extents leaves out only
  // triples straight to 1e-6 of a spacing but makes predicates near the far
  // vertices uncertain (the batch's exact pass then runs every round), so
  // 100 it is; NA = N + nfar
  int64_t N = 0, capN = 0, NA = 0;
  int nfar = 0;
  double *d_p = nullptr;

  // triangulation (real + ghost triangles; ghost vertex = -1)
  int64_t T = 0, capT = 0;
  int32_t *d_tri = nullptr, *d_adj = nullptr;
  int32_t *d_ori = nullptr;
  uint64_t *d_claim = nullptr, *d_cand = nullptr;
  int32_t *d_partner = nullptr, *d_mark = nullptr;
  int32_t *d_queue = nullptr;  // unresolved predicate queue
  uint8_t *d_keep = nullptr;

  // unique edges and per-vertex lists
  int64_t E = 0, capE = 0;
  int32_t *d_edges = nullptr;
  void *d_ews = nullptr;
  size_t ews_bytes = 0;
  int32_t *d_sptr = nullptr, *d_scur = nullptr, *d_slist = nullptr;
  void *d_scan_ws = nullptr;

  // forces
  void *d_L = nullptr, *d_h = nullptr, *d_F = nullptr;  // per edge (T = float|double)
  double *d_Lp = nullptr, *d_hp = nullptr;              // exact-order sums
  double *d_part = nullptr;                             // per-block partial sums
  double *d_sums = nullptr;                             // [sumL, sumH, scale]
  uint64_t *d_maxF = nullptr;

  // device counters (layout in smk::kCtr*) and pinned host mirrors
  int32_t *d_ctr = nullptr;
  int32_t *h_ctr = nullptr;
  int32_t *h_status = nullptr;  // written by the captured iteration

  // the FP32 iteration captured as a graph (one launch + one sync each)
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int64_t graph_T = -1;
  bool use_graph = true;
  double *d_psnap = nullptr;  // points at the start of the iteration

  // last valid state: a Delaunay triangulation and the points it is valid
  // for (the start of kinetic repair), plus a staging array for sub-steps
  int32_t *d_tri_s = nullptr, *d_adj_s = nullptr;
  double *d_pvalid = nullptr, *d_pstage = nullptr;
  bool have_snap = false;
  double *h_sums = nullptr;
  uint64_t *h_maxF = nullptr;

  // sliver removal: per-cell flags then ranks, per-vertex winning rank
  int32_t *d_sl_rank = nullptr, *d_sl_win = nullptr, *d_sl_total = nullptr;
  void *d_sl_ws = nullptr;
  int64_t sl_capT = 0, sl_capN = 0;
  int64_t nsl = -1;  // selected slivers, -1 before a selection

  // neural SDF surrogate (single domain, one level): the network decides
  // the centroid filter and the projection test wherever |d_nn| >= band;
  // inside the band the exact program runs.  d_sg: [skipped exact
  // evaluations, exact evaluations, max |d_nn - d| (double bits) over them]
  bool surrogate = false;
  sm_mlp net{};
  double band = 0.0;
  // optional per-region bands: a res^dim grid over [band_lo, band_hi]
  // (sm_mesher_set_surrogate_bands); the band at x is max(band, the value
  // of x's cell), and the largest value outside the grid
  float *d_bands = nullptr;
  int band_res = 0;
  double band_lo[3] = {0, 0, 0}, band_inv[3] = {0, 0, 0};
  double band_max = 0.0;
  float *d_dnn = nullptr;
  int64_t dnn_cap = 0;
  uint64_t *d_sg = nullptr;

  // 3D: the host tetrahedralisation, built by sl_threads threads inserting
  // the points of sl_n slabs along sl_axis into one shared
  // tetrahedralisation (delaunay3d.cpp); its cells are copied to the device
  // and filtered there piece by piece while the host hands out the next.
  int sl_axis = 1, sl_n = 1, sl_threads = 1;
  smdt::DT3 dt3;
  double *h_pts = nullptr;  // pinned staging of the points (N, 3)
  int32_t *h_tet = nullptr;  // pinned staging of the tetrahedra (capT, 4)
  int64_t h_pts_cap = 0, h_tet_cap = 0;

  int32_t *d_tctr = nullptr;  // private counter block of sm_mesher_time_kernel
  // workgroup sizes of the tunable kernels (smk::kBlk*), set by
  // sm_mesher_autotune; 256 until then
  int blk[6] = {256, 256, 256, 256, 256, 256};

  int64_t stats[kNumStats] = {};
  int round = 0;
  bool have_tri = false;
  double last_maxdp = 0.0;

  // optional per-phase timing with HIP events on the handle's stream
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  struct Rec {
    int phase;
    int a, b;  // indices into ev_pool
  };
  std::vector<Rec> recs;
  int ev_next = 0;  // next free event pair in ev_pool
  double prof_ms[SM_NPHASES] = {};
  int64_t prof_n[SM_NPHASES] = {};
};

namespace smk {

// What the force kernels need to treat a batch handle's domains apart
// (all pointers null for a single-domain handle).
struct BatchView {
  const int32_t *vdom = nullptr, *voff = nullptr, *nfix = nullptr;
  const double *h = nullptr;
  double *scale = nullptr;
  const sm_node *tapes = nullptr;
  const int *tape_off = nullptr, *tape_len = nullptr;
  double geps_mult = 0.0, deps_mult = 0.0;
};

// The surrogate's exact-fallback band at a point (see sm_mesher.band*).
struct BandView {
  const float *v = nullptr;
  int res = 0;
  double lo[3] = {0, 0, 0}, inv[3] = {0, 0, 0};
  double band = 0.0, vmax = 0.0;
};

inline BandView band_view(const sm_mesher *m) {
  BandView b;
  b.v = m->d_bands;
  b.res = m->band_res;
  for (int k = 0; k < 3; ++k) {
    b.lo[k] = m->band_lo[k];
    b.inv[k] = m->band_inv[k];
  }
  b.band = m->band;
  b.vmax = std::max(m->band, m->band_max);
  return b;
}

template <int DIM>
__device__ inline double band_at(const BandView &b, const double *x) {
  if (!b.v) return b.band;
  int64_t idx = 0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const double c = (x[d] - b.lo[d]) * b.inv[d];
    if (!(c >= 0.0 && c < (double)b.res)) return b.vmax;  // outside the grid, or NaN
    idx = idx * b.res + (int64_t)c;
  }
  return fmax(b.band, (double)b.v[idx]);
}

inline BatchView batch_view(const sm_mesher *m) {
  BatchView b;
  if (m->ndom == 0) return b;
  b.vdom = m->d_vdom;
  b.voff = m->d_voff;
  b.nfix = m->d_dnfix;
  b.h = m->d_dh;
  b.scale = m->d_dscale;
  b.tapes = m->d_tape_all;
  b.tape_off = m->d_tape_off_dev;
  b.tape_len = m->d_tape_len_dev;
  b.geps_mult = m->prm.geps;
  b.deps_mult = m->prm.deps;
  return b;
}

// tunable kernels: centroid filter, bar forces, gather/update/projection,
// and the flip rounds' orientation, marking and apply kernels
constexpr int kBlkKeep = 0, kBlkForce = 1, kBlkUpdate = 2, kBlkOrient = 3, kBlkMark = 4,
              kBlkApply = 5, kNumBlk = 6;

// d_ctr layout (int32): [4] unique edge count, [16] last round id,
// [24..31] status of the last batch of flip rounds, [32..) one 8-int block
// per flip round of a batch
constexpr int kCtrE = 4;
constexpr int kCtrRoundId = 16;
constexpr int kCtrStatus = 24;
// the 64-bit movement maximum (d_maxF) sits right before the status words,
// so the captured iteration reads both back with one copy
constexpr int kCtrMaxF = kCtrStatus - 2;
static_assert(kCtrMaxF > kCtrRoundId && kCtrMaxF % 2 == 0, "counter layout");
constexpr int kCtrRounds = 32;
constexpr int kMaxBatchRounds = 64;
constexpr int kCtrTotal = kCtrRounds + 8 * kMaxBatchRounds;
constexpr int kCtrScratchHost = 64;  // pinned host staging inside h_ctr

// exact: settle uncertain predicates on the device; otherwise count them
// as unresolved (the batch then reports failure)
int enqueue_flip_rounds(sm_mesher *m, int nrounds, int32_t *status, const double *pts,
                        bool exact);
void account_flips(sm_mesher *m, const int32_t *status);
int enqueue_snapshot(sm_mesher *m, const int32_t *status, double *p_any = nullptr);
int snapshot_valid(sm_mesher *m);
bool kinetic_repair(sm_mesher *m);
// event-timed phases (no-ops unless profiling is enabled)
int prof_begin(sm_mesher *m);
void prof_end(sm_mesher *m, int phase, int begin_idx);
int prof_drain(sm_mesher *m);
// flip2d.hip
int flip_repair2d(sm_mesher *m, int max_rounds, bool *need_rebuild);
int flag_failed_domains(sm_mesher *m, int64_t *bad_points, bool *dup);
int relocate_inverted(sm_mesher *m, int64_t *moved);
int host_rebuild2d(sm_mesher *m);
// average duration of `reps` back-to-back launches of the flip-marking
// kernel on the current triangulation (timing only; no state changes that
// the next flip round does not overwrite)
int time_flip_mark(sm_mesher *m, int reps, double *ms);
// average ms of `reps` launches of tunable kernel `which` (kBlk*) with
// workgroup size B on the current state; the state is left as the next
// iteration expects it (forces.hip: keep / force / update, flip2d.hip:
// orient / mark / apply)
int time_force_tunable(sm_mesher *m, int which, int B, int reps, double *ms);
int time_flip_tunable(sm_mesher *m, int which, int B, int reps, double *ms);
// forces.hip: `sync_edges` reads the edge count back to the host (needed
// by the exact-order sums); otherwise every launch is sized for the
// largest possible edge count and reads the count on the device
int centroid_keep(sm_mesher *m);
// the filter on cells [off, off + count) only (device programs; false when
// this handle filters on the host, which then runs centroid_keep instead)
int centroid_keep_range(sm_mesher *m, int64_t off, int64_t count, bool *done);
int build_edges(sm_mesher *m, bool sync_edges);
int forces_and_update(sm_mesher *m);
// mlp.hip: surrogate inference (points, or centroids of cells)
int mlp_check(const sm_mlp *net);
int mlp_eval_points(const sm_mlp *net, const double *x, int64_t n, float *d, hipStream_t s);
int mlp_eval_cells(const sm_mlp *net, const int32_t *cells, int64_t ncells, const double *p,
                   float *d, hipStream_t s);
}  // namespace smk
