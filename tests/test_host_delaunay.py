"""Host exact Delaunay (C ABI sm_host_delaunay2d/3d) against the exact
brute-force oracle and Qhull (CPU)."""
import numpy as np
import pytest
from scipy.spatial import Delaunay

from oracle.predicates import is_delaunay_2d, orient2d


def _dt(p):
    from seismicmesh_amd import _lib
    p = np.ascontiguousarray(p, dtype=np.float64)
    cap = 2 * len(p) + 16
    tri = np.zeros((cap, 3), np.int32)
    adj = np.zeros((cap, 3), np.int32)
    ins = np.zeros(len(p), np.uint8)
    m = _lib.lib.sm_host_delaunay2d(p.ctypes.data, len(p), tri.ctypes.data, adj.ctypes.data,
                                    cap, ins.ctypes.data)
    assert m >= 0, _lib.last_error()
    return tri[:m], adj[:m], ins


def _check_links(tri, adj):
    for t in range(len(tri)):
        for k in range(3):
            o = adj[t, k]
            a, b = tri[t, (k + 1) % 3], tri[t, (k + 2) % 3]
            assert any(tri[o, (j + 1) % 3] == b and tri[o, (j + 2) % 3] == a and adj[o, j] == t
                       for j in range(3))


@pytest.mark.parametrize("seed", range(6))
def test_random_general_position(seed):
    rng = np.random.default_rng(seed)
    p = rng.random((int(rng.integers(3, 60)), 2))
    tri, adj, ins = _dt(p)
    _check_links(tri, adj)
    real = tri[(tri >= 0).all(1)]
    assert is_delaunay_2d(p, real)
    want = {tuple(sorted(s)) for s in Delaunay(p).simplices.tolist()}
    assert {tuple(sorted(s)) for s in real.tolist()} == want
    # ghosts: one per convex-hull edge, Euler count of the triangulated sphere
    assert len(tri) == 2 * len(p) - 2


def test_degenerate_lattice_and_duplicates():
    g = np.stack(np.meshgrid(np.arange(6.0), np.arange(5.0)), -1).reshape(-1, 2)
    p = np.vstack([g, g[:3]])  # exact duplicates are skipped
    tri, adj, ins = _dt(p)
    _check_links(tri, adj)
    assert ins.sum() == len(g)
    real = tri[(tri >= 0).all(1)]
    assert is_delaunay_2d(g, real)
    for t in real:
        assert orient2d(p[t[0]], p[t[1]], p[t[2]]) > 0
    # a triangulation of the 6 x 5 lattice has 2 * (5 * 4) cells
    assert len(real) == 40


def test_collinear_and_tiny_inputs():
    assert _dt(np.array([[0.0, 0.0], [1.0, 1.0]]))[0].shape[0] == 0
    line = np.array([[0.0, 0.0], [1.0, 1.0], [2.0, 2.0], [3.0, 3.0]])
    assert _dt(line)[0].shape[0] == 0


def test_large_matches_qhull():
    rng = np.random.default_rng(7)
    p = rng.random((20000, 2))
    tri, _, _ = _dt(p)
    real = tri[(tri >= 0).all(1)]
    want = {tuple(sorted(s)) for s in Delaunay(p).simplices.tolist()}
    assert {tuple(sorted(s)) for s in real.tolist()} == want


# ----------------------------------------------------------------- 3D
def _tets(p):
    from seismicmesh_amd import _lib
    p = np.ascontiguousarray(p, dtype=np.float64)
    n = len(p)
    cap = 10 * n + 64
    t = np.zeros((cap, 4), np.int32)
    ins = np.zeros(n, np.uint8)
    m = _lib.lib.sm_host_delaunay3d(p.ctypes.data, n, t.ctypes.data, cap, ins.ctypes.data)
    assert m >= 0, m
    return t[:m], ins


@pytest.mark.parametrize("seed", range(4))
def test_3d_empty_spheres_exact(seed):
    # brute force against the exact (rational) predicates on small inputs
    from oracle.predicates import insphere, orient3d
    rng = np.random.default_rng(seed)
    p = rng.random((int(rng.integers(5, 30)), 3))
    t, ins = _tets(p)
    assert ins.all()
    for tt in t:
        a, b, c, d = p[tt]
        assert orient3d(a, b, c, d) > 0  # positively oriented
        for k in range(len(p)):
            if k not in tt:
                assert insphere(a, b, c, d, p[k]) <= 0
    # the union of the cells is the convex hull: volumes add up
    from scipy.spatial import ConvexHull
    vol = np.abs(np.einsum("ij,ij->i", np.cross(p[t[:, 1]] - p[t[:, 0]], p[t[:, 2]] - p[t[:, 0]]),
                           p[t[:, 3]] - p[t[:, 0]])).sum() / 6
    assert abs(vol - ConvexHull(p).volume) < 1e-12


def test_3d_matches_qhull_and_skips_duplicates():
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(7)
    p = rng.random((3000, 3))
    t, _ = _tets(p)
    want = {tuple(sorted(x)) for x in Delaunay(p).simplices.tolist()}
    assert {tuple(sorted(x)) for x in t.tolist()} == want
    q = np.vstack((p[:50], p[:10]))
    t2, ins = _tets(q)
    assert ins[:50].all() and not ins[50:].any()
    assert t2.max() < 50


# ------------------------------------------------- 3D: uniqueness and slabs
def _tets_slabs(p, nslabs, axis=1, threads=4):
    """The tetrahedralisation built by `threads` threads inserting the points
    of `nslabs` slabs along `axis` into one shared triangulation."""
    from seismicmesh_amd import _lib
    p = np.ascontiguousarray(p, dtype=np.float64)
    n = len(p)
    cap = 12 * n + 64
    t = np.zeros((cap, 4), np.int32)
    ins = np.zeros(n, np.uint8)
    m = _lib.lib.sm_host_delaunay3d_slabs(p.ctypes.data, n, axis, nslabs, threads, t.ctypes.data,
                                          cap, ins.ctypes.data)
    assert m >= 0, _lib.last_error()
    return t[:m], ins


def _canon_ok(p, t):
    # lowest id first, second-lowest second, positively oriented, sorted
    from oracle.predicates import orient3d
    assert (t[:, 0] < t[:, 1]).all() and (t[:, 1] < t[:, 2]).all() and (t[:, 1] < t[:, 3]).all()
    key = [tuple(x) for x in t.tolist()]
    assert key == sorted(key)
    for tt in t[:: max(1, len(t) // 300)]:
        assert orient3d(*p[tt]) > 0


def _circumradius(p, t):
    a = p[t[:, 0]]
    u, v, w = p[t[:, 1]] - a, p[t[:, 2]] - a, p[t[:, 3]] - a
    num = ((u * u).sum(1)[:, None] * np.cross(v, w) + (v * v).sum(1)[:, None] * np.cross(w, u)
           + (w * w).sum(1)[:, None] * np.cross(u, v))
    den = 2 * np.einsum("ij,ij->i", u, np.cross(v, w))
    return np.linalg.norm(num / den[:, None], axis=1)


def test_3d_lattice_unique_whatever_the_order():
    """Cospherical lattice points: the symbolic perturbation makes the
    tetrahedralisation a function of the point set, so any insertion order
    gives the same cells; they tile the box exactly."""
    g = np.stack(np.meshgrid(np.arange(7.0), np.arange(6.0), np.arange(5.0), indexing="ij"),
                 -1).reshape(-1, 3)
    t0, ins = _tets(g)
    assert ins.all()
    _canon_ok(g, t0)
    vol = np.einsum("ij,ij->i", np.cross(g[t0[:, 1]] - g[t0[:, 0]], g[t0[:, 2]] - g[t0[:, 0]]),
                    g[t0[:, 3]] - g[t0[:, 0]]) / 6
    assert vol.min() > 0 and abs(vol.sum() - 6 * 5 * 4) < 1e-9
    rng = np.random.default_rng(3)
    for _ in range(3):
        perm = rng.permutation(len(g))
        t1, _ = _tets(g[perm])
        t1 = perm[t1]
        assert {tuple(sorted(x)) for x in t1.tolist()} == {tuple(sorted(x)) for x in t0.tolist()}


def test_3d_perturbed_empty_spheres_on_degenerate_points():
    """On points with many cospherical subsets every cell is empty of the
    other points in the exact sense (insphere <= 0), and the cells fill the
    convex hull."""
    from scipy.spatial import ConvexHull
    from oracle.predicates import insphere
    g = np.stack(np.meshgrid(np.arange(3.0), np.arange(3.0), np.arange(3.0), indexing="ij"),
                 -1).reshape(-1, 3)
    p = np.vstack([g, [[0.5, 0.5, 0.5], [1.5, 1.0, 0.5]]])
    t, ins = _tets(p)
    assert ins.all()
    for tt in t:
        for k in range(len(p)):
            if k not in tt:
                assert insphere(*p[tt], p[k]) <= 0
    vol = np.einsum("ij,ij->i", np.cross(p[t[:, 1]] - p[t[:, 0]], p[t[:, 2]] - p[t[:, 0]]),
                    p[t[:, 3]] - p[t[:, 0]]) / 6
    assert abs(vol.sum() - ConvexHull(p).volume) < 1e-12


@pytest.mark.parametrize("kind", ["lattice", "jittered", "random", "clustered"])
def test_3d_threads_equal_serial_build(kind):
    """Several threads inserting the points of slabs into one shared
    tetrahedralisation return the serial build's cells, in the same order,
    for any slab count, axis and thread count -- on cospherical lattices,
    jittered lattices, random and strongly clustered points."""
    rng = np.random.default_rng(11)
    if kind == "random":
        p = rng.random((30000, 3)) * [3.0, 2.0, 1.0]
    elif kind == "clustered":
        c = rng.random((20, 3))
        p = (c[rng.integers(0, 20, 30000)] + rng.normal(0, 0.02, (30000, 3)))
    else:
        h = 0.06
        g = np.mgrid[0:2.0 + h / 2:h, 0:1.5 + h / 2:h, 0:0.8 + h / 2:h].reshape(3, -1).T
        p = g if kind == "lattice" else g + rng.uniform(-0.2, 0.2, g.shape) * h
    serial, ins0 = _tets(p)
    _canon_ok(p, serial)
    for axis, nslabs, threads in [(0, 4, 4), (1, 7, 3), (2, 16, 8), (1, 1, 8)]:
        t, ins = _tets_slabs(p, nslabs, axis=axis, threads=threads)
        assert np.array_equal(ins, ins0)
        assert np.array_equal(t, serial), (axis, nslabs, threads)


def test_3d_threads_duplicates_keep_lowest_id():
    rng = np.random.default_rng(5)
    p = rng.random((20000, 3))
    q = np.vstack([p[15000:], p[:15000], p[:400]])  # duplicates of ids 5000.. at the end
    t_s, ins_s = _tets(q)
    t_p, ins_p = _tets_slabs(q, 6, axis=0, threads=6)
    assert np.array_equal(ins_s, ins_p)
    assert ins_s[:20000].all() and not ins_s[20000:].any()
    assert t_s.max() < 20000 and np.array_equal(t_s, t_p)


def test_3d_convex_hull_cells_match_qhull_on_a_jittered_lattice():
    """Flat cells along the hull faces of a lattice (which the reference's
    CGAL triangulation has, and whose centroids lie inside a box domain that
    the lattice does not reach) are produced too: every cell equals Qhull's
    on a lattice moved by 1e-3 of its spacing."""
    from scipy.spatial import Delaunay
    h = 1.0
    g = np.mgrid[0:9.0:h, 0:7.0:h, 0:4.0:h].reshape(3, -1).T
    p = g + np.random.default_rng(2).uniform(-1e-3, 1e-3, g.shape) * h
    t, ins = _tets(p)
    assert ins.all()
    want = {tuple(sorted(x)) for x in Delaunay(p).simplices.tolist()}
    got = {tuple(sorted(x)) for x in t.tolist()}
    assert got == want
    vol = np.einsum("ij,ij->i", np.cross(p[t[:, 1]] - p[t[:, 0]], p[t[:, 2]] - p[t[:, 0]]),
                    p[t[:, 3]] - p[t[:, 0]]) / 6
    assert vol.min() > 0
    # the hull's flat cells are there: some have a circumradius far above h
    assert _circumradius(p, t).max() > 50 * h


def test_3d_check_accepts_delaunay_cells_and_rejects_others():
    """sm_host_delaunay3d_check passes the cells of the tetrahedralisation
    and fails cells that are not Delaunay."""
    from seismicmesh_amd import _lib
    rng = np.random.default_rng(4)
    p = rng.random((3000, 3)) * [2.0, 1.0, 1.0]

    def check(cells):
        c = np.ascontiguousarray(cells, dtype=np.int32).reshape(-1, 4)
        r = _lib.lib.sm_host_delaunay3d_check(p.ctypes.data, len(p), c.ctypes.data, len(c), 4)
        assert r >= 0, _lib.last_error()
        return r

    serial, _ = _tets(p)
    assert check(serial) == 1
    assert check(serial[::7]) == 1
    # a cell of the tetrahedralisation with one vertex swapped for a far point
    from oracle.predicates import orient3d
    bad = serial[0].copy()
    bad[3] = np.argmax(np.linalg.norm(p - p[bad].mean(0), axis=1))
    if orient3d(*p[bad]) < 0:
        bad[[2, 3]] = bad[[3, 2]]
    assert check(bad) == 0
