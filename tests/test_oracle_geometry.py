"""Pins of the oracle's geometry layer: the reference's compiled helpers,
closed forms and the reference tests' known answers (CPU)."""
import math

import numpy as np
import pytest

from oracle import geometry as og


def test_known_answer_cube_tets():
    # tests/test_geometry.py:44-56 of the reference: a cube split into 12 tets
    pts = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1],
                    [1, -1, 1], [1, 1, 1], [-1, 1, 1], [0, 0, 0]], dtype=float)
    cells = np.array([[3, 4, 0, 8], [3, 7, 8, 2], [6, 8, 7, 2], [6, 5, 8, 2], [6, 5, 7, 8],
                      [8, 7, 4, 5], [3, 7, 4, 8], [3, 8, 1, 2], [8, 5, 1, 2], [8, 4, 1, 5],
                      [8, 4, 0, 1], [3, 8, 0, 1]])
    assert len(og.get_facets(cells)) == 48
    assert len(og.get_boundary_facets(cells)) == 12
    assert len(og.get_boundary_vertices(cells, dim=3)) == 8
    assert np.sum(og.simp_vol(pts, cells)) == 8.0


def test_unique_edges_brute_force():
    rng = np.random.default_rng(1)
    for n in (0, 1, 7, 200):
        e = rng.integers(0, 12, (n, 2)).astype(np.int32)
        got = og.unique_edges(e)
        want = sorted({(min(a, b), max(a, b)) for a, b in e.tolist()})
        assert got.tolist() == [list(w) for w in want]


def test_unique_edges_vs_reference(fast_geometry_ref):
    rng = np.random.default_rng(2)
    for n in (1, 10, 5000):
        e = rng.integers(0, 300, (n, 2)).astype(np.int32)
        assert np.array_equal(og.unique_edges(e), fast_geometry_ref.unique_edges(e))


def test_box_sdfs_vs_reference(fast_geometry_ref):
    rng = np.random.default_rng(3)
    p2 = rng.uniform(-2, 2, (1000, 2))
    p3 = rng.uniform(-2, 2, (1000, 3))
    assert np.array_equal(og.drectangle(p2, -1.0, 0.5, -0.3, 1.2),
                          fast_geometry_ref.drectangle_fast(p2, -1.0, 0.5, -0.3, 1.2))
    assert np.array_equal(og.dblock(p3, -1.0, 0.5, -0.3, 1.2, 0.0, 0.1),
                          fast_geometry_ref.dblock_fast(p3, -1.0, 0.5, -0.3, 1.2, 0.0, 0.1))


def _tets(n, seed=4):
    rng = np.random.default_rng(seed)
    p = rng.uniform(-1, 1, (4 * n, 3))
    t = np.arange(4 * n).reshape(n, 4)
    return p, t


def test_dihedral_closed_forms():
    # regular tetrahedron: every dihedral angle is arccos(1/3)
    reg = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]], dtype=float)
    a = og.calc_dihedral_angles(reg, np.array([[0, 1, 2, 3]]))
    assert np.allclose(a, math.acos(1.0 / 3.0), atol=1e-14)
    # right-corner tetrahedron: the three edges at the corner have 90 degrees
    rc = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=float)
    a = og.calc_dihedral_angles(rc, np.array([[0, 1, 2, 3]])).ravel()
    # angles 3, 4, 5 are along edges (0,3), (0,2), (0,1)
    assert np.allclose(a[3:], math.pi / 2, atol=1e-14)
    assert np.allclose(a[:3], math.acos(1 / math.sqrt(3)), atol=1e-14)


def test_dihedral_vs_reference(fast_geometry_ref):
    p, t = _tets(500)
    ref = fast_geometry_ref.calc_dihedral_angles(p, t.astype(np.int32))
    got = og.calc_dihedral_angles(p, t)
    assert np.array_equal(np.isnan(ref), np.isnan(got))
    assert np.allclose(got, ref, rtol=0, atol=1e-12, equal_nan=True)


def _circumradius2(p0, p1, p2, p3):
    # independent: solve 2 (pi - p3) . c = |pi - p3|^2 for the centre
    A = np.stack([p0 - p3, p1 - p3, p2 - p3])
    b = (A * A).sum(1)
    c = np.linalg.solve(2 * A, b)
    return float(c @ c)


def test_circumsphere_grad_finite_difference():
    rng = np.random.default_rng(5)
    for _ in range(50):
        q = rng.uniform(-1, 1, (4, 3))
        g = og.calc_circumsphere_grad(q[0:1], q[1:2], q[2:3], q[3:4])[0]
        fd = np.zeros(3)
        h = 1e-6
        for k in range(3):
            a = q.copy()
            b = q.copy()
            a[0, k] += h
            b[0, k] -= h
            fd[k] = (_circumradius2(*a) - _circumradius2(*b)) / (2 * h)
        assert np.allclose(g, fd, rtol=1e-5, atol=1e-6 * max(1.0, np.abs(fd).max()))


def test_circumsphere_grad_vs_reference(fast_geometry_ref):
    p, t = _tets(300, seed=6)
    args = [p[t[:, k]] for k in range(4)]
    assert np.array_equal(og.calc_circumsphere_grad(*args),
                          fast_geometry_ref.calc_circumsphere_grad(*args))


def test_sdf_closed_forms():
    x = np.array([[0.0, 0.0], [2.0, 0.0], [0.0, -0.5]])
    assert np.allclose(og.Disk([0, 0], 1).eval(x), [-1.0, 1.0, -0.5])
    r = og.Rectangle((0.0, 1.0, 0.0, 2.0))
    assert np.allclose(r.eval(np.array([[0.5, 1.0], [2.0, 1.0], [0.5, -1.0]])), [-0.5, 1.0, 1.0])
    b = og.Ball([0, 0, 0], 2.0)
    assert np.allclose(b.eval(np.array([[0, 0, 0], [3, 0, 0.0]])), [-2.0, 1.0])
    tor = og.Torus(1.0, 0.25)
    assert np.allclose(tor.eval(np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0]])), [-0.25, 0.75])
    cyl = og.Cylinder(h=2.0, r=0.5)
    assert np.allclose(cyl.eval(np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0, 2.0, 0]])),
                       [-0.5, 0.5, 1.0])
    # CSG: union = min, difference = max(d0, -d1), smooth union <= min
    d1, d2 = og.Disk([0, 0], 1), og.Disk([1, 0], 1)
    u = og.Union([d1, d2])
    pts = np.random.default_rng(0).uniform(-2, 3, (100, 2))
    assert np.array_equal(u.eval(pts), np.minimum(d1.eval(pts), d2.eval(pts)))
    df = og.Difference([d1, d2])
    assert np.array_equal(df.eval(pts), np.maximum(d1.eval(pts), -d2.eval(pts)))
    su = og.Union([d1, d2], smoothness=0.1)
    assert np.all(su.eval(pts) <= u.eval(pts) + 1e-15)


def test_simp_qual_closed_forms():
    eq = np.array([[0, 0], [1, 0], [0.5, math.sqrt(3) / 2]])
    assert np.allclose(og.simp_qual(eq, np.array([[0, 1, 2]])), 1.0)
    rt = np.array([[0, 0], [1, 0], [0, 1.0]])
    # right isosceles: 2 r / R = 2 (2 - sqrt 2) / 2 / (sqrt 2 / 2)
    assert np.allclose(og.simp_qual(rt, np.array([[0, 1, 2]])), 2 * (1 - 1 / math.sqrt(2)) / (math.sqrt(2) / 2))


def test_boundary_vertices_match_the_oracle():
    """The package's boundary vertices of tetrahedra (facets keyed as one
    integer each) equal the oracle's (unique rows of the sorted facets) on
    random cell sets, including shared and repeated facets."""
    from seismicmesh_amd import mesh_utils as mu
    rng = np.random.default_rng(3)
    for _ in range(20):
        n = int(rng.integers(5, 3000))
        t = rng.integers(0, n, (int(rng.integers(1, 5000)), 4)).astype(np.int32)
        s = np.sort(t, 1)
        t = t[(s[:, 1:] != s[:, :-1]).all(1)]
        if len(t) == 0:
            continue
        want = np.unique(og.get_boundary_facets(t).reshape(-1))
        assert np.array_equal(mu.get_boundary_vertices(t, 3), want)


def test_used_ids_match_numpy_unique():
    """fix_mesh's compaction of the vertex ids (by counting) returns exactly
    np.unique(t, return_inverse=True): values, dtypes and shapes."""
    from seismicmesh_amd import mesh_utils as mu
    rng = np.random.default_rng(2)
    for trial in range(30):
        shape = (int(rng.integers(0, 3000)), int(rng.integers(1, 5)))
        t = rng.integers(0, int(rng.integers(1, 10000)), shape).astype([np.int32, np.int64][trial % 2])
        a_p, _, a_j = np.unique(t, return_index=True, return_inverse=True)
        b_p, b_j = mu._used_ids(t)
        assert a_p.dtype == b_p.dtype and np.array_equal(a_p, b_p)
        assert a_j.dtype == b_j.dtype and a_j.shape == b_j.shape and np.array_equal(a_j, b_j)
