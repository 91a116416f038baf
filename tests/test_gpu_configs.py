"""Every benchmark workload on the GPU against the oracle.

* BP2004-shaped 2D (1911 x 5395 grid) at full size, end to end: vertex count
  within 1 % and min / mean cell quality within 0.01 of the oracle's mesh.
* Many independent 2D CSG domains in one batch handle: per domain, the kept
  cells and sorted unique edges equal the oracle's Delaunay + filter on the
  same (jittered) points, and each FP32 update is within 1e-5 of the
  oracle's FP64 update, over two iterations.
* The 76-primitive 3D CSG at h = 0.08: the same per-iteration checks, and the
  surrogate-steered loop gives the exact loop's points.
* The EAGE-shaped 3D workload (grid coarsened 10x, same physical sizes):
  one iteration's cells, edges and update.

Inputs that are Delaunay-degenerate (the initial lattices) are jittered by
1e-3 h so that the triangulation being compared is unique."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    from seismicmesh_amd import _lib
    _lib.require_gpu()


def _jitter(p, h, seed, nfix=0):
    """Points moved by up to 1e-3 h so that the Delaunay triangulation is
    unique.  (The oracle triangulates with scipy's Qhull, which is not
    exact: on lattices moved by only 1e-6 h it returns a few cells whose
    spheres hold another point, where the library's exact predicates do
    not; and it breaks the ties of cospherical points, such as the corners
    of a box, its own way.)  Comparisons of single iterations pass the
    fixed points to the oracle by index and jitter them too (nfix = 0); an
    end-to-end comparison keeps the first nfix points on the fixed-point
    coordinates, since the reference fixes, every iteration, the point
    closest to each of them (mesh_generator.py:471-475)."""
    rng = np.random.default_rng(seed)
    q = p + rng.uniform(-1e-3, 1e-3, p.shape) * h
    q[:nfix] = p[:nfix]
    return q


def _cellset(t):
    return {tuple(sorted(c)) for c in np.asarray(t).tolist()}


def _same_cells(a, b, what):
    """Set equality with a short message (pytest's own diff of two large
    sets takes minutes)."""
    A, B = _cellset(a), _cellset(b)
    if A != B:
        raise AssertionError("%s: %d cells only on the device, %d only in the oracle (of %d)"
                             % (what, len(A - B), len(B - A), len(B)))


def _same_edges(a, b, what):
    if not np.array_equal(a, b):
        raise AssertionError("%s: edge lists differ (%d device, %d oracle edges)"
                             % (what, len(a), len(b)))


def _oracle_state(domain, h, p):
    from oracle import distmesh as odm
    opts = dict(odm.GEN_DEFAULTS)
    opts.update(verbose=0, points=p)
    return odm.setup(domain, h, opts)


def _check_iterations(m, s, p, nfix, iters, what):
    """Runs `iters` iterations of handle m from points p, checking kept cells,
    edges and the update against the oracle state s each time."""
    import time
    from oracle import distmesh as odm

    def say(msg):
        print("  %s iteration %d: %s (%.1f s)" % (what, it, msg, time.time() - t0), flush=True)

    for it in range(iters):
        t0 = time.time()
        t_all = odm.delaunay(p)
        say("oracle Delaunay")
        m.triangulate()
        say("device triangulation")
        t = odm.remove_triangles_outside(p, t_all, s["fd"], s["geps"])
        _same_cells(m.cells(), t, "%s: cells, iteration %d" % (what, it))
        _same_edges(m.edges(), odm.get_edges(t), "%s: edges, iteration %d" % (what, it))
        say("cells and edges compared")
        m.run(1)
        say("device iteration")
        q_ref, _, _ = odm.force_step(p, t_all, s, list(range(nfix)))
        q = m.points()
        dp, dp_ref = q - p, q_ref - p
        scale = np.abs(dp_ref).max()
        assert scale > 0
        err = np.abs(dp - dp_ref).max() / scale
        print("%s iteration %d: %d points, %d cells, update error %.2e of the largest move"
              % (what, it, len(p), len(t), err))
        assert err <= 1e-5, "%s: update, iteration %d" % (what, it)
        p = q


def test_bp2004_full_size_end_to_end_vs_oracle():
    """The full BP2004-shaped workload (1911 x 5395 grid, 43 k vertices, 50
    iterations, termination) against the oracle's generate_mesh from the
    same jittered initial points: in FP64 the meshes are the same (every
    iteration is bit-exact and the triangulations of points in general
    position are unique); in FP32 they agree on the vertex count within 1 %
    and on mean quality within 0.01, and the minimum quality -- a single
    element of a run that diverges from the FP64 one at the first flip whose
    decision the 1e-5 update differences change -- within 0.03 (0.016
    measured)."""
    import time

    import seismicmesh_amd as sm
    from oracle import distmesh as odm
    from oracle import geometry as og
    from oracle.sizing import SizeFunction, build_interpolant
    from seismicmesh_amd import generation, mesh_utils as mu, workloads
    t0 = time.time()
    domain, ef, info = workloads.bp2004()
    grid = ef.cell_size
    st = generation.prepare(domain, ef, verbose=0)
    p0 = _jitter(st["p"], st["h0"], 7, st["nfix"])
    oef = SizeFunction(info["bbox"], build_interpolant(grid.values.astype(np.float64), grid.bbox),
                       ef.hmin)
    q, s = odm.generate_mesh(og.Rectangle(info["bbox"]), oef, verbose=0, points=p0)
    qb = og.simp_qual(q, s)
    print("BP2004 oracle: %d vertices, %d cells, quality min %.4f mean %.4f (%.0f s)"
          % (len(q), len(s), qb.min(), qb.mean(), time.time() - t0), flush=True)
    for prec in (64, 32):
        p, t = sm.generate_mesh(domain, ef, verbose=0, precision=prec, points=p0)
        qa = mu.simp_qual(p, t)
        print("BP2004 device FP%d: %d vertices, %d cells, quality min %.4f mean %.4f"
              % (prec, len(p), len(t), qa.min(), qa.mean()), flush=True)
        if prec == 64:
            assert len(p) == len(q) and len(t) == len(s)
            assert np.array_equal(p, q)
            _same_cells(t, s, "BP2004 FP64 end to end")
        else:
            assert abs(len(p) - len(q)) <= 0.01 * len(q)
            assert abs(qa.mean() - qb.mean()) <= 0.01
            assert abs(qa.min() - qb.min()) <= 0.03


def _check_sphere(tri, adj, npts):
    """A valid triangulated sphere: every point in a real triangle, links
    mutual, no two triangles sharing two edges."""
    live = tri[:, 0] != -2
    idx = np.nonzero(live)[0]
    t, a = tri[idx], adj[idx]
    real = t[(t >= 0).all(1)]
    present = np.zeros(npts, bool)
    present[real.ravel()] = True
    assert present.all(), "points in no real triangle: %s" % np.nonzero(~present)[0][:10]
    assert ((a[:, 0] != a[:, 1]) & (a[:, 1] != a[:, 2]) & (a[:, 0] != a[:, 2])).all(), \
        "two triangles share two edges"
    for k in range(3):
        nb = a[:, k]
        assert (adj[nb] == idx[:, None]).any(1).all(), "links not mutual"


def test_bp2004_device_repair_keeps_collinear_hull_points():
    """BP2004, iteration 1 from jittered points with the fixed corners in
    place: boundary points projected onto the box's sides lie exactly on the
    line through a fixed corner, and the device repair must keep that
    corner in the triangulation (collinear hull points are all vertices, as
    in the reference's CGAL triangulation): a valid triangulated sphere
    whose kept cells equal the oracle's."""
    from oracle import distmesh as odm
    from oracle import geometry as og
    from oracle.sizing import SizeFunction, build_interpolant
    from seismicmesh_amd import generation, workloads
    domain, ef, info = workloads.bp2004()
    st = generation.prepare(domain, ef, verbose=0)
    p0 = _jitter(st["p"], st["h0"], 7, st["nfix"])
    st["p"] = p0
    m = generation.make_mesher(st)
    m.set_points(p0)
    m.run(1)
    m.triangulate()
    p1 = m.points()
    tri, adj = m.triangulation()
    _check_sphere(tri, adj, len(p1))
    grid = ef.cell_size
    oef = SizeFunction(info["bbox"], build_interpolant(grid.values.astype(np.float64), grid.bbox),
                       ef.hmin)
    s = _oracle_state(og.Rectangle(info["bbox"]), oef, p0)
    t = odm.remove_triangles_outside(p1, odm.delaunay(p1), s["fd"], s["geps"])
    _same_cells(m.cells(), t, "BP2004 iteration 1")
    m.close()


def _far_square_misses(p, t_oracle, t_dev):
    """Cells the oracle keeps and a far-square triangulation leaves out must
    be triangles with an edge on the convex hull whose circumcircle holds a
    far vertex (100 extents out); a far-square triangulation keeps no other
    cell than the oracle.  Returns the number of such cells."""
    A = {tuple(sorted(c)) for c in np.asarray(t_oracle).tolist()}
    B = {tuple(sorted(c)) for c in np.asarray(t_dev).tolist()}
    assert B <= A, "%d cells only in the far-square triangulation" % len(B - A)
    lo, hi = p.min(0), p.max(0)
    ext, c = (hi - lo).max(), (lo + hi) / 2
    far = c + 100 * ext * np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]])
    from scipy.spatial import ConvexHull
    hull = {tuple(sorted(e)) for e in ConvexHull(p).simplices.tolist()}
    for cell in A - B:
        assert any(tuple(sorted((cell[i], cell[j]))) in hull for i, j in ((0, 1), (1, 2), (0, 2))), \
            "left-out cell %s has no hull edge" % (cell,)
        a, b, q = (p[i] for i in cell)
        # circumcentre
        d = 2 * (a[0] * (b[1] - q[1]) + b[0] * (q[1] - a[1]) + q[0] * (a[1] - b[1]))
        ux = ((a @ a) * (b[1] - q[1]) + (b @ b) * (q[1] - a[1]) + (q @ q) * (a[1] - b[1])) / d
        uy = ((a @ a) * (q[0] - b[0]) + (b @ b) * (a[0] - q[0]) + (q @ q) * (b[0] - a[0])) / d
        r2 = (a[0] - ux) ** 2 + (a[1] - uy) ** 2
        assert (((far - [ux, uy]) ** 2).sum(1) < r2).any(), \
            "left-out cell %s: no far vertex in its circle" % (cell,)
    return len(A - B)


@pytest.mark.parametrize("far_square", [False, True], ids=["hull", "far_square"])
def test_many_domains_batch_vs_oracle(far_square):
    """Per domain, from jittered initial lattices: over each domain's convex
    hull (the default, as the reference) kept cells and edges equal the
    oracle's every iteration; inside a far square the kept cells are the
    oracle's but for hull triangles whose circumcircle reaches a far vertex
    (flat triangles of nearly collinear outer-row points, whose centroids
    can pass the filter before the boundary rows are projected).  Updates
    agree within 1e-5 of the oracle's step from the same points whenever
    the cells agree."""
    from oracle import distmesh as odm
    from oracle import geometry as og
    from seismicmesh_amd import batch, workloads
    doms, hs = workloads.many_domains(n=6, nmax=2.0e4)
    odoms, ohs = workloads.many_domains(n=6, nmax=2.0e4, geom=og)
    assert hs == ohs
    sts = batch.prepare_batch(doms, hs)
    for k, st in enumerate(sts):
        st["p"] = _jitter(st["p"], st["h0"], 100 + k)
    bm = batch.make_batch_mesher(sts, far_square=far_square)
    p0 = np.vstack([st["p"] for st in sts])
    bm.set_points(p0)
    ps = [st["p"] for st in sts]
    missed = 0
    for it in range(3):
        bm.triangulate()
        parts = bm.split(bm.points(), bm.cells())
        be = bm.edges()
        same = []
        for d, (st, od) in enumerate(zip(sts, odoms)):
            s = _oracle_state(od, hs[d], ps[d])
            t = odm.remove_triangles_outside(ps[d], odm.delaunay(ps[d]), s["fd"], s["geps"])
            if far_square:
                k = _far_square_misses(ps[d], t, parts[d][1])
                missed += k
                if k:
                    same.append(False)
                    continue
            _same_cells(parts[d][1], t, "domain %d, iteration %d" % (d, it))
            b, e = bm.voff[d], bm.voff[d + 1]
            mine = be[(be[:, 0] >= b) & (be[:, 0] < e)] - b
            _same_edges(mine, odm.get_edges(t), "domain %d, iteration %d" % (d, it))
            same.append(True)
        bm.run(1)
        pb = bm.points()
        for d, (st, od) in enumerate(zip(sts, odoms)):
            if not same[d]:
                ps[d] = pb[bm.voff[d]:bm.voff[d + 1]].copy()
                continue
            s = _oracle_state(od, hs[d], ps[d])
            q_ref, _, _ = odm.force_step(ps[d], odm.delaunay(ps[d]), s, list(range(st["nfix"])))
            b, e = bm.voff[d], bm.voff[d + 1]
            dp, dp_ref = pb[b:e] - ps[d], q_ref - ps[d]
            scale = np.abs(dp_ref).max()
            assert scale > 0
            err = np.abs(dp - dp_ref).max() / scale
            print("batch domain %d (%d points) iteration %d: update error %.2e"
                  % (d, len(ps[d]), it, err))
            assert err <= 1e-5, "domain %d, iteration %d" % (d, it)
            ps[d] = pb[b:e].copy()
    print("far square: %s, hull cells left out %d" % (far_square, missed))
    bm.close()


def test_csg3d_reduced_vs_oracle():
    from oracle import geometry as og
    from seismicmesh_amd import generation, workloads
    dom, _ = workloads.csg3d()
    odom, _ = workloads.csg3d(geom=og)
    h = 0.08
    st = generation.prepare(dom, h, verbose=0, precision=32)
    p0 = _jitter(st["p"], h, 3)
    st["p"] = p0
    m = generation.make_mesher(st)
    m.set_points(p0)
    s = _oracle_state(odom, h, p0)
    _check_iterations(m, s, p0, st["nfix"], 2, "csg3d h=0.08")
    m.close()


def test_csg3d_surrogate_matches_exact():
    """The surrogate-steered 3D loop on the 76-primitive CSG gives the exact
    loop's points while skipping exact SDF evaluations."""
    from seismicmesh_amd import generation, workloads
    from seismicmesh_amd.neural import NeuralSDF
    dom, bbox = workloads.csg3d()
    h = 0.08
    net = NeuralSDF(3, bbox, hidden=3, seed=0, omega=8.0)
    net.train(dom, nsamples=200000, steps=1500, batch=4096, lr=2e-3, lr_final=1e-4, eik=0.1,
              near=0.05, seed=0)
    bands, lo, hi, err = net.calibrate_bands(dom, 2.0 * h, res=16, nval=400000)
    band = float(bands.max())
    print("csg3d surrogate trained: max error %.4f, bands %.4f .. %.4f"
          % (err, bands.min(), band), flush=True)
    st = generation.prepare(dom, h, verbose=0, precision=32)
    a, b = generation.make_mesher(st), generation.make_mesher(st)
    b.set_surrogate(net, 2.0 * h, bands=(bands, lo, hi))
    a.set_points(st["p"])
    b.set_points(st["p"])
    for _ in range(3):
        a.run(1)
        b.run(1)
        assert np.array_equal(a.points(), b.points())
    sg = b.surrogate_stats()
    frac = sg["skipped"] / (sg["skipped"] + sg["exact"])
    print("csg3d surrogate: max error on the check %.4f, band %.4f: %.1f%% of exact "
          "evaluations skipped, max |d_nn - d| in the band %.4f"
          % (err, band, 100 * frac, sg["max_err"]))
    assert sg["max_err"] < band
    assert frac > 0.0
    a.close()
    b.close()


def test_eage3d_reduced_vs_oracle():
    from oracle import geometry as og
    from oracle.sizing import build_interpolant
    from seismicmesh_amd import generation, workloads
    domain, ef, info = workloads.eage3d(scale=0.1)
    st = generation.prepare(domain, ef, verbose=0, precision=32)
    p0 = _jitter(st["p"], st["h0"], 4)
    st["p"] = p0
    m = generation.make_mesher(st)
    m.set_points(p0)
    grid = ef.cell_size
    fh = build_interpolant(grid.values.astype(np.float64), grid.bbox)
    fd = og.Cube(info["bbox"]).eval
    s = dict(fd=fd, fh=fh, h0=st["h0"], dim=3, L0mult=1 + 0.4 / 4, geps=st["geps"],
             deps=st["deps"], pfix=p0[:st["nfix"]], levels=[fd], delta_t=0.3)
    _check_iterations(m, s, p0, st["nfix"], 1, "eage3d scale=0.1")
    m.close()


def test_far_square_with_surrogate_keeps_no_far_cell():
    """With a far square, the cells of the four far vertices are never kept,
    even when a surrogate (here an untrained network, whose sign far outside
    its box is arbitrary) decides the centroid test there."""
    from seismicmesh_amd import generation, geometry as geo
    from seismicmesh_amd.neural import NeuralSDF
    dom = geo.Disk([0.0, 0.0], 1.0)
    h = 0.1
    net = NeuralSDF(2, (-1.0, 1.0, -1.0, 1.0), hidden=3, seed=5, omega=8.0)
    st = generation.prepare(dom, h, verbose=0, precision=32)
    m = generation.make_mesher(st, far_square=True)
    m.set_surrogate(net, 1e-9)
    m.set_points(st["p"])
    for _ in range(3):
        m.run(1)
        m.triangulate()
        c = m.cells()
        assert len(c) > 0 and c.max() < len(st["p"]) and c.min() >= 0
    m.close()


def test_unique_edges_recursive_scan():
    """More than 1,048,576 vertices: the offset scans take the recursive
    path (over 1,024 tiles)."""
    import torch
    from oracle import geometry as og
    from seismicmesh_amd import _lib
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(9)
    nvert = 1_200_000
    pairs = rng.integers(0, nvert, (3_000_000, 2)).astype(np.int32)
    want = og.unique_edges(pairs)
    pt = torch.from_numpy(pairs).to(dev)
    out = torch.empty((len(pairs), 2), dtype=torch.int32, device=dev)
    cnt = torch.zeros(1, dtype=torch.int64, device=dev)
    wsb = _lib.lib.sm_unique_edges_workspace(len(pairs), nvert)
    ws = torch.empty(int(wsb), dtype=torch.uint8, device=dev)
    _lib.check(_lib.lib.sm_unique_edges(pt.data_ptr(), len(pairs), nvert, out.data_ptr(),
                                        cnt.data_ptr(), ws.data_ptr(), int(wsb), None),
               "sm_unique_edges")
    torch.cuda.synchronize()
    n = int(cnt.item())
    assert n == len(want)
    assert np.array_equal(out[:n].cpu().numpy(), want)
