// Sliver removal step on the force-loop handle (3D): selection of kept
// tetrahedra with a dihedral angle out of bounds and the circumsphere
// perturbation of their first vertices.
//
// Follows SeismicMesh/generation/mesh_generator.py:222-272 (the body of the
// sliver_removal loop) and :532-540 (_calc_dihedral_angles):
//   ele  = kept cells with any of the six angles < min or > max (a NaN
//          angle selects nothing, as numpy's comparisons do);
//   g    = circumsphere gradient of (t[ele,0..3]), infinities set to 1,
//          divided by its Euclidean norm sqrt((g0^2 + g1^2) + g2^2);
//   p[t[ele,0]] += scale * g  with numpy's fancy-index semantics: every
//          entry reads the old point, and where several slivers share a
//          first vertex the last one in cell order is the one written.
//
// Layout: the handle's (T, 4) tetrahedra and keep flags; ranks are an
// exclusive scan of the selection flags in cell order (= the order of the
// kept-cell list the reference indexes), so rank order is ele order.
#include <algorithm>

#include "common.h"
#include "mesher.h"
#include "quality3d.h"

using namespace smk;

namespace {

__global__ void k_sliver_flag(int64_t T, const int32_t *tet, const uint8_t *keep, const double *p,
                              double min_dh, double max_dh, int32_t *flag) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= T) return;
  int32_t bad = 0;
  if (keep[c]) {
    int32_t v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = tet[4 * c + k];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      double a = dihedral_angle(p, v, i);
      bad |= (a < min_dh) | (a > max_dh);
    }
  }
  flag[c] = bad;
}

// rank[c] is the exclusive scan of the flags; a cell is selected when the
// next rank is larger.  The last selected cell sharing a first vertex wins.
__global__ void k_sliver_winner(int64_t T, const int32_t *tet, const int32_t *rank,
                                const int32_t *total, int32_t *win) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= T) return;
  const int32_t r = rank[c];
  const int32_t next = c + 1 < T ? rank[c + 1] : *total;
  if (next == r) return;
  atomicMax(&win[tet[4 * c]], r);
}

// New positions of the winning first vertices, from the old points only
// (other threads' cells read those points), staged per vertex.
__global__ void k_sliver_move(int64_t T, const int32_t *tet, const int32_t *rank,
                              const int32_t *total, const int32_t *win, double scale,
                              const double *p, double *stage) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= T) return;
  const int32_t r = rank[c];
  const int32_t next = c + 1 < T ? rank[c + 1] : *total;
  if (next == r) return;
  const int32_t v0 = tet[4 * c];
  if (win[v0] != r) return;
  double q[4][3];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int64_t v = tet[4 * c + k];
#pragma unroll
    for (int d = 0; d < 3; ++d) q[k][d] = p[3 * v + d];
  }
  double g[3];
  circumsphere_grad(q[0], q[1], q[2], q[3], g);
#pragma unroll
  for (int d = 0; d < 3; ++d)
    if (isinf(g[d])) g[d] = 1.0;
  const double nrm = sqrt((g[0] * g[0] + g[1] * g[1]) + g[2] * g[2]);
#pragma unroll
  for (int d = 0; d < 3; ++d) stage[3 * (int64_t)v0 + d] = q[0][d] + scale * (g[d] / nrm);
}

__global__ void k_sliver_commit(int64_t N, const int32_t *win, const double *stage, double *p) {
  int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= N || win[v] < 0) return;
#pragma unroll
  for (int d = 0; d < 3; ++d) p[3 * v + d] = stage[3 * v + d];
}

int ensure_sliver_state(sm_mesher *m) {
  if (m->sl_capT >= m->T + 1 && m->sl_capN >= m->N) return SM_OK;
  if (m->d_sl_rank) (void)hipFree(m->d_sl_rank);
  if (m->d_sl_win) (void)hipFree(m->d_sl_win);
  if (m->d_sl_total) (void)hipFree(m->d_sl_total);
  if (m->d_sl_ws) (void)hipFree(m->d_sl_ws);
  m->d_sl_rank = m->d_sl_win = m->d_sl_total = nullptr;
  m->d_sl_ws = nullptr;
  m->sl_capT = m->sl_capN = 0;
  const int64_t capT = m->capT + 1, capN = std::max<int64_t>(m->capN, 1);
  SM_HIP(hipMalloc((void **)&m->d_sl_rank, capT * sizeof(int32_t)));
  SM_HIP(hipMalloc((void **)&m->d_sl_win, capN * sizeof(int32_t)));
  SM_HIP(hipMalloc((void **)&m->d_sl_total, sizeof(int32_t)));
  SM_HIP(hipMalloc(&m->d_sl_ws, scan_workspace(capT)));
  m->sl_capT = capT;
  m->sl_capN = capN;
  return SM_OK;
}

}  // namespace

extern "C" int sm_mesher_find_slivers(sm_mesher *m, double min_dh, double max_dh,
                                      int64_t *nslivers) {
  SM_CHECK(m && nslivers, SM_ERR_ARG, "mesher_find_slivers: null argument");
  SM_CHECK(m->dim == 3, SM_ERR_STATE, "mesher_find_slivers: the handle meshes 2D domains");
  SM_CHECK(m->have_tri && m->N > 0, SM_ERR_STATE,
           "mesher_find_slivers: no tetrahedralisation (call sm_mesher_triangulate)");
  SM_HIP(hipSetDevice(m->device));
  int rc = ensure_sliver_state(m);
  if (rc) return rc;
  hipStream_t s = m->stream;
  const int64_t T = m->T;
  if (T == 0) {
    m->nsl = 0;
    *nslivers = 0;
    return SM_OK;
  }
  const int B = 256;
  hipLaunchKernelGGL(k_sliver_flag, dim3(blocks_for(T, B)), dim3(B), 0, s, T, m->d_tri,
                     m->d_keep, m->d_p, min_dh, max_dh, m->d_sl_rank);
  SM_LAUNCH_CHECK();
  if ((rc = scan_exclusive(m->d_sl_rank, m->d_sl_rank, T, m->d_sl_total, m->d_sl_ws, s)))
    return rc;
  int32_t total = 0;
  SM_HIP(hipMemcpyAsync(&total, m->d_sl_total, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  SM_HIP(hipStreamSynchronize(s));
  m->nsl = total;
  *nslivers = total;
  return SM_OK;
}

extern "C" int sm_mesher_perturb_slivers(sm_mesher *m, double scale) {
  SM_CHECK(m, SM_ERR_ARG, "mesher_perturb_slivers: null handle");
  SM_CHECK(m->nsl >= 0, SM_ERR_STATE,
           "mesher_perturb_slivers: no sliver selection (call sm_mesher_find_slivers)");
  SM_HIP(hipSetDevice(m->device));
  if (m->nsl == 0) return SM_OK;
  hipStream_t s = m->stream;
  const int64_t T = m->T, N = m->N;
  const int B = 256;
  int rc = fill<int32_t>(m->d_sl_win, N, -1, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_sliver_winner, dim3(blocks_for(T, B)), dim3(B), 0, s, T, m->d_tri,
                     m->d_sl_rank, m->d_sl_total, m->d_sl_win);
  SM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sliver_move, dim3(blocks_for(T, B)), dim3(B), 0, s, T, m->d_tri,
                     m->d_sl_rank, m->d_sl_total, m->d_sl_win, scale, m->d_p, m->d_pstage);
  SM_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_sliver_commit, dim3(blocks_for(N, B)), dim3(B), 0, s, N, m->d_sl_win,
                     m->d_pstage, m->d_p);
  SM_LAUNCH_CHECK();
  SM_HIP(hipStreamSynchronize(s));
  // the points moved: the cells (and selection) describe the old points
  m->nsl = -1;
  return SM_OK;
}
