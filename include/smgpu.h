/* smgpu.h -- C ABI of libsmgpu, the MI355X (gfx950) DistMesh force loop.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and
 * returns an int status (SM_OK = 0, negative on error; sm_last_error()
 * describes the last failure of the calling thread).  No C++ exception
 * crosses this boundary.
 *
 * Device entry points take device pointers (caller-owned) and a HIP stream
 * (passed as void*; NULL = the default stream) and only enqueue work: results
 * are ready when the stream is synchronised.  Host entry points (sm_host_*)
 * run synchronously on the calling thread.  A handle (sm_mesher) is not
 * thread-safe; use one per stream.
 *
 * Reference interfaces replaced (paths relative to the SeismicMesh 3.6.2
 * source root):
 *   sm_unique_edges / sm_cell_edges  <- geometry/cpp/fast_geometry.cpp:49-77
 *                                       (unique_edges) and
 *                                       generation/mesh_generator.py:680-688
 *   sm_sdf_eval                       <- geometry/signed_distance_functions.py
 *                                       (Disk.eval ... Repeat.eval, :283-623),
 *                                       drectangle_fast / dblock_fast
 *                                       (fast_geometry.cpp:104-214)
 *   sm_sizing_eval                    <- sizing/mesh_size_function.py:391-408
 *                                       (RegularGridInterpolator __call__)
 *   sm_mesher_* (the handle)          <- generation/mesh_generator.py:460-527
 *                                       (the force loop: re-triangulation,
 *                                       _remove_triangles_outside :734-738,
 *                                       _compute_forces :691-712 with
 *                                       generation/utils.py:48-68 (dense),
 *                                       update and
 *                                       _project_points_back_newton :762-784)
 *   sm_mesher_find_slivers,
 *   sm_mesher_perturb_slivers         <- generation/mesh_generator.py:222-272,
 *                                       :532-540 (sliver_removal loop body)
 *   sm_mlp_*, sm_gemm_f32,
 *   sm_mesher_set_surrogate           <- no counterpart (neural SDF surrogate)
 *   sm_host_delaunay2d                <- generation/cpp/delaunay_class.cpp
 *                                       :108-121, :232-298 (CGAL insert and
 *                                       finite cells/vertices)
 *   sm_host_delaunay3d                <- generation/cpp/delaunay_class3.cpp
 *                                       :116-129, :276-344
 *   sm_dihedral_angles                <- fast_geometry.cpp:351-452
 *   sm_circumsphere_grad              <- fast_geometry.cpp:580-703
 *   sm_limgrad                        <- sizing/cpp/FastHJ.cpp:63-187
 */
#ifndef SMGPU_H
#define SMGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SM_OK 0
#define SM_ERR_ARG (-1)       /* invalid argument (null pointer, bad size) */
#define SM_ERR_HIP (-2)       /* a HIP runtime call failed */
#define SM_ERR_CAPACITY (-3)  /* an output or workspace buffer is too small */
#define SM_ERR_STATE (-4)     /* handle used in the wrong state */
#define SM_ERR_NUMERIC (-5)   /* divergence: max movement >= 1000 h0 */

/* ------------------------------------------------------------------ misc */
int sm_version(void);
/* Copies the last error message of the calling thread (NUL-terminated). */
int sm_last_error(char *buf, size_t len);
/* Number of usable gfx950 devices (0 when no GPU is present). */
int sm_device_count(int *count);

/* ---------------------------------------------------------- SDF programs
 * A signed distance function is lowered to a postfix program of nodes
 * ("tape").  Primitive nodes push the distance of the current point; CSG
 * nodes pop `nargs` values and push their combination; SM_OP_PUSH_MOD pushes
 * a transformed evaluation point (Repeat) and SM_OP_POP_POINT restores it.
 */
enum {
  SM_OP_DISK = 1,      /* prm: xc, yc, r                                     */
  SM_OP_RECT = 2,      /* prm: x1, x2, y1, y2 (bbox0)                        */
  SM_OP_BALL = 3,      /* prm: xc, yc, zc, r                                 */
  SM_OP_CUBE = 4,      /* prm: x1, x2, y1, y2, z1, z2 (bbox0)               */
  SM_OP_TORUS = 5,     /* prm: r1, r2                                        */
  SM_OP_PRISM = 6,     /* prm: b, h                                          */
  SM_OP_CYLINDER = 7,  /* prm: r, h/2                                        */
  SM_OP_UNION = 16,    /* nargs children; prm[0] = smoothness k (0: sharp)   */
  SM_OP_INTERSECT = 17,
  SM_OP_DIFFERENCE = 18,
  SM_OP_SDIFF_STEP = 19, /* pops b, a; pushes the smooth difference step
                            max(-a, b) + h^2/(4k), h = max(k - |-a - b|, 0) */
  SM_OP_PUSH_MOD = 32, /* prm: period[3]: q = mod(x + p/2, p) - p/2         */
  SM_OP_POP_POINT = 33
};
/* transform flags of a primitive node, applied in this order to the point
 * before the primitive is evaluated (signed_distance_functions.py:89-101) */
enum {
  SM_XF_TRANSLATE = 1, /* x -= xf[0:3]                                       */
  SM_XF_ROT_X = 2,     /* x = Rx_inv x, xf[3:12] row-major (2D: xf[3:7])     */
  SM_XF_ROT_Y = 4,     /* x = Ry_inv x, xf[12:21]                            */
  SM_XF_ROT_Z = 8,     /* x = Rz_inv x, xf[21:30]                            */
  SM_XF_STRETCH = 16   /* v = xf[30:33] (unit), alpha = xf[33]               */
};
typedef struct {
  int32_t op;
  int32_t nargs;
  int32_t flags;
  int32_t dim;
  double prm[8];
  double xf[34];
  double pad[2]; /* pad[0] of the first node: 1.0 when the program needs
                    more than 8 stack values or nests SM_OP_PUSH_MOD (it
                    then runs in the slower scratch-memory interpreter) */
} sm_node; /* 368 bytes */

/* d[i] = sdf(x[i]) in FP64; x is (n, dim) row-major.  `tape` is a device
 * copy of `ntape` nodes. */
int sm_sdf_eval(const sm_node *tape, int ntape, int dim, const double *x, int64_t n,
                double *d, void *stream);

/* ------------------------------------------------- neural SDF surrogate
 * No counterpart in the reference (it evaluates the CSG tape exactly,
 * signed_distance_functions.py:283-623).  A multilayer perceptron
 * in_dim -> width -> ... -> width -> 1 with SiLU activations approximates
 * the signed distance:  x' = (x - shift) * scale  (per axis),
 * a_0 = silu(W_in x' + b_in), a_l = silu(W_l a_{l-1} + b_l) for
 * l = 1..hidden, d = out_scale * (w_out . a_hidden + b_out).
 * The hidden weights are BF16 (the inference kernel multiplies them on the
 * matrix cores with FP32 accumulation; activations are rounded to BF16
 * between layers); the input and output layers stay FP32.  Every pointer
 * is device memory. */
typedef struct {
  int32_t in_dim;      /* 2 or 3                                              */
  int32_t width;       /* 256                                                 */
  int32_t hidden;      /* number of width x width layers (3 for 3->256x4->1)  */
  int32_t reserved;
  const float *w_in;   /* [width][in_dim]                                     */
  const float *b_in;   /* [width]                                             */
  const void *w_hid;   /* bf16 fragments from sm_mlp_pack_hidden, 16-B aligned */
  const float *b_hid;  /* [hidden][width], 16-B aligned                      */
  const float *w_out;  /* [width]                                             */
  float b_out;
  float reserved2;
  double shift[3];
  double scale[3];
  double out_scale;
} sm_mlp;

/* d[i] = mlp(x[i]) (FP32 result), x (n, in_dim) FP64 row-major. */
int sm_mlp_eval(const sm_mlp *net, const double *x, int64_t n, float *d, void *stream);
/* d[t] = mlp(centroid of cell t), cells (ncells, in_dim + 1) int32; a cell
 * with a negative vertex id gets +inf. */
int sm_mlp_eval_centroids(const sm_mlp *net, const int32_t *cells, int64_t ncells,
                          const double *p, float *d, void *stream);

/* Training kernels (FP32).  The forward and backward passes of the
 * network are sequences of these calls on caller-owned buffers.
 *
 * C[m, n] = alpha * sum_k A(m, k) B(k, n) + beta * C[m, n] with
 * A(m, k) = A[m * sam + k * sak], B(k, n) = B[k * sbk + n * sbn] and C
 * row-major with leading dimension ldc.  With nsplit > 1 the k range is cut
 * into nsplit slices whose partial products go to `work`
 * (nsplit * M * N floats) and are summed in slice order, so the result does
 * not depend on scheduling. */
int sm_gemm_f32(int M, int N, int K, float alpha, const float *A, int64_t sam, int64_t sak,
                const float *B, int64_t sbk, int64_t sbn, float beta, float *C, int64_t ldc,
                int nsplit, float *work, void *stream);
/* z[r, c] += b[c]; a[r, c] = silu(z[r, c]) unless a is NULL. */
int sm_mlp_bias_silu(int64_t rows, int cols, float *z, const float *b, float *a, void *stream);
/* dz = da * silu'(z) elementwise. */
int sm_mlp_silu_backward(int64_t count, const float *z, const float *da, float *dz,
                         void *stream);
/* Training inputs: for each of the n samples x (FP64, (n, dim)), rows
 * b * n + i of X (FP32, ((2 dim + 1) n, dim)) hold x' = (x - shift) * scale
 * (b = 0) and x' +- h e_k (b = 2k + 1, 2k + 2) when eik_h > 0; only b = 0
 * when eik_h == 0. */
int sm_mlp_make_inputs(int64_t n, int dim, const double *x, const double *shift,
                       const double *scale, float eik_h, float *X, void *stream);
/* Loss of the surrogate and its gradient with respect to the network
 * outputs y (same row blocks as sm_mlp_make_inputs):
 *   L = mean_i (y_i - t_i)^2 + lam * mean_i' (|g_i| - 1)^2,
 *   g_i,k = (y[(2k+1) n + i] - y[(2k+2) n + i]) / (2 eik_h)
 * (a central-difference eikonal term; lam = 0 drops it).  With trunc > 0
 * the targets are a truncated SDF and samples with |t_i| >= trunc have no
 * eikonal term (i' runs over the others; the mean still divides by n).
 * part[b] receives the loss partial of block b of 256 samples
 * (ceil(n / 256) floats). */
int sm_mlp_sdf_loss(int64_t n, int dim, const float *y, const float *t, float eik_h, float lam,
                    float trunc, float *grad, float *part, void *stream);
/* Adam (Kingma & Ba): m1 = b1 m1 + (1-b1) g, m2 = b2 m2 + (1-b2) g^2,
 * w -= lr (m1 / (1 - b1^step)) / (sqrt(m2 / (1 - b2^step)) + eps). */
int sm_mlp_adam(int64_t count, float *w, const float *g, float *m1, float *m2, float lr,
                float b1, float b2, float eps, int step, void *stream);
/* Hidden weights w (FP32, [hidden][width (out)][width (in)]) rounded to
 * BF16 (nearest even) and stored in the matrix-core fragment order that
 * sm_mlp.w_hid expects (hidden * width * width values, 16-B aligned). */
int sm_mlp_pack_hidden(int hidden, int width, const float *w, void *packed, void *stream);
/* dst (bf16) = round-to-nearest-even(src). */
int sm_f32_to_bf16(int64_t count, const float *src, void *dst, void *stream);

/* --------------------------------------------------------- sizing grids */
typedef struct {
  int32_t dim;        /* 2 or 3; axes are (z, x[, y]) = point columns 0,1,2 */
  int32_t n[3];       /* grid points per axis                               */
  const double *axis[3]; /* device: grid coordinates per axis (ascending)   */
  const float *values;   /* device: n0*n1[*n2] values, C order              */
  double hconst;      /* used when values == NULL: uniform sizing           */
} sm_grid;

/* h[i] = linear interpolation of the grid at x[i] (linear extrapolation
 * outside), FP64 index/weight arithmetic, result in FP32. */
int sm_sizing_eval(const sm_grid *grid, const double *x, int64_t n, float *h,
                   void *stream);

/* ----------------------------------------------------------- unique edges
 * Sorted unique undirected edges: each pair becomes (min, max); the output
 * is lexicographically sorted without duplicates, exactly like
 * std::sort + std::unique in the reference.  Vertex ids must lie in
 * [0, nvert).  `count` is a device int64 receiving the number of edges;
 * `edges` needs room for npairs pairs. */
size_t sm_unique_edges_workspace(int64_t npairs, int64_t nvert);
int sm_unique_edges(const int32_t *pairs, int64_t npairs, int64_t nvert, int32_t *edges,
                    int64_t *count, void *ws, size_t ws_bytes, void *stream);
/* Same, taking the local edges of cells (dim+1 ids each) in the reference's
 * order (0,1),(1,2),(2,0)[,(0,3),(1,3),(2,3)]. */
size_t sm_cell_edges_workspace(int64_t ncells, int dim, int64_t nvert);
int sm_cell_edges(const int32_t *cells, int64_t ncells, int dim, int64_t nvert,
                  int32_t *edges, int64_t *count, void *ws, size_t ws_bytes, void *stream);

/* sm_unique_edges without a workspace argument, allocating nothing: no
 * counterpart in the reference; for callers that cannot allocate device
 * memory (HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The workspace is a library-owned device array of
 * SM_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends), shared by
 * every call on that GPU: calls that overlap on one GPU race.  Inputs whose
 * sm_unique_edges_workspace exceeds it are refused with SM_ERR_CAPACITY
 * before anything is launched. */
#define SM_NOALLOC_SCRATCH_BYTES (64ll << 20)
int sm_unique_edges_noalloc(const int32_t *pairs, int64_t npairs, int64_t nvert,
                            int32_t *edges, int64_t *count, void *stream);

/* ----------------------------------------------------------- the loop
 * A 2D/3D force-loop handle owning device state for one domain:
 * points (FP64), triangulation, edges, per-vertex edge lists, forces.
 */
typedef struct sm_mesher sm_mesher;

/* Host sizing callback: fills h[i] = edge_length(mid[i]) for n midpoints
 * (host arrays, (n, dim) FP64).  Returns 0 on success.  Used when the sizing
 * is a user function that only the host can evaluate. */
typedef int (*sm_host_sizing_fn)(void *ctx, const double *mid, int64_t n, double *h);

typedef struct {
  int32_t dim;            /* 2 (triangles) or 3 (tetrahedra)                 */
  int32_t precision;      /* 64: FP64 forces (bit-exact with the reference's
                             order of operations); 32: FP32 forces            */
  int32_t nfix;           /* the first nfix points are fixed (pfix)          */
  int32_t nlevels;        /* number of level sets (domain + subdomains)      */
  double h0, delta_t, geps, deps, L0mult;
  int32_t far_square;     /* 2D: 1 = triangulate (each domain of a batch)
                             inside a fixed square of four far vertices
                             (100 extents from the points) instead of the
                             convex hull of the points: no sequential
                             hull-flip chains along straight boundaries;
                             hull slivers whose circumcircle reaches that
                             far (kept by the reference where their
                             centroid passes the filter) are not formed.
                             0 = convex hull, as the reference.             */
  int32_t reserved;
} sm_mesher_params;

/* Creates a handle on device `device`.  `tapes`/`tape_lens` give one SDF
 * program per level (host arrays; copied).  `sizing` describes the sizing
 * grid with HOST pointers (copied to the device), or a constant when
 * values == NULL; `host_sizing` (may be NULL) overrides it with a callback. */
int sm_mesher_create(const sm_mesher_params *prm, const sm_node *const *tapes,
                     const int *tape_lens, const sm_grid *sizing,
                     sm_host_sizing_fn host_sizing, void *host_sizing_ctx, int device,
                     sm_mesher **out);
/* A handle meshing `ndom` independent 2D domains at once, each with a
 * uniform size: domain d owns the points [voff[d], voff[d+1]) of the point
 * array (voff[0] = 0), its SDF program tapes[d] (tape_lens[d] nodes), its
 * size h[d] (its h0) and fixes its first nfix[d] points.  Each domain is
 * iterated exactly as its own handle would be (its own triangulation, L0
 * scale, centroid filter and projection), all domains in the same kernel
 * launches.  prm: dim 2, precision 32, delta_t and L0mult as usual; geps
 * and deps are multipliers of each domain's h (0.1 and sqrt(eps) in the
 * reference); h0, nfix and nlevels are ignored.  The max movement that
 * sm_mesher_run reports is in units of each domain's h0.  No counterpart
 * in the reference, which meshes one domain per call; replaces looping
 * generate_mesh (mesh_generator.py:291-529) over many domains. */
/* Host SDF callback, for domains given as plain functions (the reference
 * accepts domain=callable with a bbox keyword, mesh_generator.py:614-618):
 * fills d[i] = sdf_level(x[i]) for n points x (host arrays, (n, dim) FP64);
 * level 0 is the domain, 1.. the subdomains.  Returns 0 on success. */
typedef int (*sm_host_sdf_fn)(void *ctx, int level, const double *x, int64_t n, double *d);
/* From the next triangulation on, the centroid filter
 * (_remove_triangles_outside, mesh_generator.py:734-738) and the projection
 * of every level (_project_points_back_newton, :762-784) call `fn` on the
 * host instead of running the handle's SDF programs: the points and cells
 * are copied to the host, the callback runs there, and the keep flags and
 * projected points are copied back; the rest of the iteration stays on the
 * device.  fn = NULL switches back to the programs.  Single-domain handles
 * without a surrogate only. */
int sm_mesher_set_host_sdf(sm_mesher *m, sm_host_sdf_fn fn, void *ctx);
int sm_mesher_create_batch(const sm_mesher_params *prm, int ndom, const sm_node *const *tapes,
                           const int *tape_lens, const double *h, const int32_t *nfix,
                           const int64_t *voff, int device, sm_mesher **out);
int sm_mesher_destroy(sm_mesher *m);
/* Uploads points (host, (n, dim) FP64); 2D builds the first triangulation
 * (3D re-tetrahedralises at every triangulation step). */
int sm_mesher_set_points(sm_mesher *m, const double *p, int64_t n);
/* Runs `iters` force iterations (each: triangulation repair in 2D or an
 * exact host re-tetrahedralisation in 3D, centroid filter, unique edges,
 * bar forces, update, projection).  `maxdp_out`
 * (host, may be NULL) receives the last iteration's max movement. */
int sm_mesher_run(sm_mesher *m, int iters, double *maxdp_out);
/* FP32 handles run each iteration as one captured HIP graph (default on);
 * disable to launch the kernels one by one with host decisions between
 * stages. */
int sm_mesher_set_graph(sm_mesher *m, int enable);
/* Makes the triangulation of the current points Delaunay and filters cells
 * (the first half of an iteration), without moving points. */
int sm_mesher_triangulate(sm_mesher *m);
/* Reads back sizes of the current state. */
int sm_mesher_sizes(sm_mesher *m, int64_t *npoints, int64_t *ncells, int64_t *nedges);
/* Copies current points / kept cells / unique edges to host buffers. */
int sm_mesher_get_points(sm_mesher *m, double *p, int64_t cap);
int sm_mesher_get_cells(sm_mesher *m, int32_t *cells, int64_t cap);
int sm_mesher_get_edges(sm_mesher *m, int32_t *edges, int64_t cap);
/* 2D: the whole triangulation the handle keeps (real and ghost triangles,
 * ghost vertex -1, dead slots -2, neighbour j opposite vertex j); returns
 * the slot count or an error code.  For checking the device repair. */
int64_t sm_mesher_get_triangulation(sm_mesher *m, int32_t *tri, int32_t *adj, int64_t cap);
/* Statistics since creation: [0] iterations, [1] flip rounds, [2] flips,
 * [3] host rebuilds (every 3D triangulation is one), [4] exact predicate
 * evaluations, [5] speculative (graph) iterations redone because their flips
 * had not converged, [6] kinetic repairs (points moved in sub-steps from the
 * last valid triangulation), [7] predicates the device's exact stage could
 * not settle (each forces a host rebuild), [8] vertices taken out and put
 * back on the host to untangle what flips could not, [9] 3D builds on
 * several threads, [10] points the threads left to the final single-thread
 * pass (their cells were locked by another thread twice), [11] finite
 * tetrahedra of the last 3D build (0 in 2D), [12]
 * microseconds spent in the host 3D tetrahedralisation (wall clock, copies
 * and overlapped filters included). */
int sm_mesher_stats(sm_mesher *m, int64_t *stats, int nstats);
/* 3D handles: build each iteration's tetrahedralisation on nthreads host
 * threads inserting the points of nslabs slabs along `axis` into one shared
 * tetrahedralisation (sm_host_delaunay3d_slabs: the serial build's cells);
 * its cells are copied to the device and filtered there, piece by piece, on
 * the handle's stream while the host still hands out the next pieces.
 * Replaces the per-rank slabs of mesh_generator.py:808-852, :867-877 and
 * decomp/blocker.py:4-111 (axis as the reference's `axis` option). */
int sm_mesher_set_slabs(sm_mesher *m, int axis, int nslabs, int nthreads);
/* Per-phase kernel timing with HIP events recorded on the handle's stream.
 * Phases: 0 flip rounds, 1 centroid filter, 2 unique edges + edge lists,
 * 3 bar lengths and sizing, 4 L0 sums, 5 bar forces, 6 gather + update +
 * projection, 7 the flip-marking kernel alone (inside phase 0), 8 the
 * surrogate network at the cell centroids (inside phase 1).
 * sm_mesher_profile(m, 1) clears and enables; sm_mesher_profile_read fills
 * ms[i] (total) and launches[i]. */
#define SM_NPHASES 9
int sm_mesher_profile(sm_mesher *m, int enable);
int sm_mesher_profile_read(sm_mesher *m, double *ms, int64_t *launches, int nphases);
/* Average duration (ms) of `reps` back-to-back launches of one phase's
 * kernel on the handle's current state, timed with HIP events on the
 * handle's stream: a roofline measurement of the kernel alone (no
 * counterpart in the reference).  Only phase 7 (the flip-marking kernel of
 * a 2D handle) is supported; its launches leave nothing that the next flip
 * round does not recompute. */
int sm_mesher_time_kernel(sm_mesher *m, int phase, int reps, double *ms);
/* Autotuner (no counterpart in the reference): times the handle's tunable
 * kernels -- [0] centroid filter, [1] bar forces, [2] gather / update /
 * projection, [3] flip orientation, [4] flip marking, [5] flip apply (2D
 * only) -- with workgroups of 64 ... 1024 threads on the current state
 * (`reps` back-to-back launches each) and keeps the fastest size per
 * kernel (the default 256 unless another is 3 % faster).  Needs one
 * iteration run first; leaves the state as the next iteration expects it.
 * chosen[k] / ms[k] (either may be NULL) receive the sizes and the best
 * per-launch times for k < nk. */
int sm_mesher_autotune(sm_mesher *m, int reps, int32_t *chosen, double *ms, int nk);
/* Sets the six workgroup sizes (same order; multiples of 64 in [64, 1024])
 * directly, e.g. to reproduce a tuned configuration. */
int sm_mesher_set_workgroups(sm_mesher *m, const int32_t *sizes, int nk);
/* Device pointers of the state, for zero-copy access from the host side. */
int sm_mesher_device_points(sm_mesher *m, double **p);
/* Stream used by the handle (hipStream_t as void*). */
int sm_mesher_stream(sm_mesher *m, void **stream);

/* ------------------------------------------------------- sliver removal (3D)
 * On the handle's current kept tetrahedra (after sm_mesher_triangulate):
 * selects the cells with a dihedral angle < min_dh or > max_dh (radians)
 * and returns their number. */
int sm_mesher_find_slivers(sm_mesher *m, double min_dh, double max_dh, int64_t *nslivers);
/* Moves the first vertex of each selected cell by scale * g / |g|, g the
 * circumsphere gradient (infinite components set to 1); where several
 * selected cells share a first vertex the last one in cell order is
 * applied.  The selection is consumed; triangulate again before the next. */
int sm_mesher_perturb_slivers(sm_mesher *m, double scale);

/* Neural surrogate for the handle's SDF (no counterpart in the reference;
 * single-domain handles without subdomains).  From the next iteration on
 * the centroid filter and the projection test first evaluate `net` (the
 * struct is copied; the weights it points to must stay alive): a cell
 * centroid with d_nn <= -band is kept and one with d_nn >= band dropped
 * without the exact program, a point with d_nn <= -band is not projected;
 * everything else goes through the exact program as before.  The result
 * is the exact path's whenever |d_nn - d| < band - geps at every skipped
 * evaluation.  net = NULL switches the surrogate off. */
int sm_mesher_set_surrogate(sm_mesher *m, const sm_mlp *net, double band);
/* Per-region bands (no counterpart in the reference): bands (res^dim FP32,
 * row-major over axes 0..dim-1, a host array) give each cell of a regular
 * grid over [lo, hi] its own exact-fallback band, e.g. from the network's
 * largest error measured in that cell; the band at x is then
 * max(band, bands[cell of x]), and the largest of them outside the grid.
 * bands = NULL goes back to the single band. */
int sm_mesher_set_surrogate_bands(sm_mesher *m, int res, const double *lo, const double *hi,
                                  const float *bands);
/* Exact evaluations skipped and made since the surrogate was set, and the
 * largest |d_nn - d| over the evaluations made (the network's error in
 * the band, judged by the exact program). */
int sm_mesher_surrogate_stats(sm_mesher *m, int64_t *skipped, int64_t *exact,
                              double *max_err);

/* ------------------------------------------------------- host triangulation */
/* Delaunay triangulation of the convex hull of p (n, 2), exact predicates.
 * Writes real and ghost triangles (ghost vertex = -1), neighbour j opposite
 * vertex j; returns the triangle count (>= 0) or an error code.  inserted[i]
 * is 0 for exact duplicates that were skipped. */
int64_t sm_host_delaunay2d(const double *p, int64_t n, int32_t *tri, int32_t *adj,
                           int64_t cap, uint8_t *inserted);

/* Delaunay tetrahedralisation of the convex hull of p (n, 3), exact
 * predicates with CGAL's symbolic perturbation, so that cospherical points
 * get one tetrahedralisation whatever the insertion order (replaces
 * delaunay_class3.cpp:116-129, :276-344).  Writes the finite tetrahedra,
 * positively oriented, canonical (lowest id first, orientation kept) and
 * sorted; returns their count or an error code.  inserted[i] is 0 for exact
 * duplicates (of equal points the lowest id is kept). */
int64_t sm_host_delaunay3d(const double *p, int64_t n, int32_t *tet, int64_t cap,
                           uint8_t *inserted);

/* The same tetrahedralisation built by nthreads threads at once: the points
 * are split into nslabs equal-width slabs along `axis` (the decomposition of
 * generation/utils.py:28-45 (make_init_points), mesh_generator.py:867-877
 * and decomp/blocker.py:4-111, whose slabs the reference meshes on separate
 * MPI ranks), and the threads insert the slabs' points into one shared
 * tetrahedralisation, locking the cells they touch.  The cells are exactly
 * those of sm_host_delaunay3d, in the same order: the perturbed predicates
 * make the tetrahedralisation a function of the point set, and both
 * functions return it canonical (lowest id first, orientation kept) and
 * sorted. */
int64_t sm_host_delaunay3d_slabs(const double *p, int64_t n, int axis, int nslabs, int nthreads,
                                 int32_t *tet, int64_t cap, uint8_t *inserted);
/* 1 when no point of p (n, 3) lies inside the circumsphere of any of the
 * `count` positively oriented tetrahedra (the perturbed test of
 * sm_host_delaunay3d, so a tetrahedron of it passes), 0 otherwise; an error
 * code for bad arguments.  The check a slab build applies to uncertified
 * cells that the force loop keeps. */
int sm_host_delaunay3d_check(const double *p, int64_t n, const int32_t *tet, int64_t count,
                             int nthreads);

/* ------------------------------------------------------- 3D quality kernels */
/* angles: (6*ncells) dihedral angles in radians, FP64 (fast_geometry.cpp
 * edge table {{2,3},{1,3},{1,2},{0,3},{0,2},{0,1}}). */
int sm_dihedral_angles(const double *p, const int32_t *cells, int64_t ncells,
                       double *angles, void *stream);
/* grad: (k, 3) gradient of the circumradius w.r.t. p0 for k tetrahedra. */
int sm_circumsphere_grad(const double *p0, const double *p1, const double *p2,
                         const double *p3, int64_t k, double *grad, void *stream);

/* ------------------------------------------------------- gradient limiting */
/* Host-callable, device-executed: ffun (nz*nx*ny, Fortran order, FP64) is
 * limited in place so that neighbouring values differ by at most elen*dfdx
 * (the fixed point of FastHJ.cpp's active-set relaxation). */
int sm_limgrad(double *ffun_dev, int nz, int nx, int ny, double elen, double dfdx,
               void *stream);

#ifdef __cplusplus
}
#endif
#endif /* SMGPU_H */
