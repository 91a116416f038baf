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
