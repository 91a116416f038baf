"""Pins of the oracle functions that the reference's compiled helpers do not
reach: Prism, Repeat, sharp and smooth Intersection / Difference, the
rotation, stretch and translation set-up, and the sliver-removal loop (CPU).

Each check is a geometric fact or a known answer that a plausible mistake in
the oracle (a swapped axis, a wrong sign, a transposed rotation, a fold in
the wrong order, a wrong step update) fails; none retypes the oracle's
formula."""
import math

import numpy as np
import pytest

from oracle import distmesh as odm
from oracle import geometry as og


# ---------------------------------------------------------------- Prism
# signed_distance_functions.py:546-566: a triangular prism whose cross
# section (z = const) is the equilateral triangle with inradius b/2 centred
# at the origin, one edge at y = -b/2 and the opposite vertex at (0, b); it
# spans z in [-h, h].
def test_prism_faces_vertices_and_axes():
    b, h = 0.8, 0.5
    pr = og.Prism(b, h)
    x = np.array([
        [0.0, 0.0, 0.0],           # centre: nearest faces are the triangle's sides
        [0.0, -b / 2, 0.0],        # on the bottom side
        [0.0, b, 0.0],             # on the apex edge
        [0.0, 0.0, h],             # on the top face
        [0.0, 0.0, -h],            # on the bottom face
        [0.0, -b / 2 - 0.3, 0.0],  # 0.3 outside the bottom side
        [0.0, 0.0, h + 0.2],       # 0.2 above the top face
    ])
    want = [-min(b / 2, h), 0.0, 0.0, 0.0, 0.0, 0.3, 0.2]
    assert np.allclose(pr.eval(x), want, atol=1e-6)
    # the two slanted sides: outward normals (+-sqrt(3)/2, 1/2) at distance
    # b/2 from the axis (the reference rounds sqrt(3)/2 to 0.866025)
    for sx in (1.0, -1.0):
        n = np.array([sx * math.sqrt(3) / 2, 0.5, 0.0])
        for s in (-0.1, 0.0, 0.25):
            assert abs(pr.eval(((b / 2 + s) * n)[None])[0] - s) < 1e-6
    # symmetric in x and z, not in y
    rng = np.random.default_rng(0)
    q = rng.uniform(-1, 1, (200, 3))
    assert np.array_equal(pr.eval(q), pr.eval(q * [-1, 1, 1]))
    assert np.array_equal(pr.eval(q), pr.eval(q * [1, 1, -1]))
    assert not np.allclose(pr.eval(q), pr.eval(q * [1, -1, 1]))
    # inside <=> inside the triangle and |z| < h (brute force half-planes)
    inside = (q[:, 1] > -b / 2) & (math.sqrt(3) / 2 * np.abs(q[:, 0]) + 0.5 * q[:, 1] < b / 2) \
        & (np.abs(q[:, 2]) < h)
    d = pr.eval(q)
    sure = np.abs(d) > 1e-5
    assert np.array_equal((d < 0)[sure], inside[sure])


# ---------------------------------------------------------------- Repeat
def _cube_interior_distance(x, bbox):
    # exact distance to the nearest face for points inside the box
    lo, hi = np.array(bbox[::2]), np.array(bbox[1::2])
    return np.minimum(x - lo, hi - x).min(1)


def test_repeat_is_periodic_copy_clipped_to_parent():
    # signed_distance_functions.py:283-297: the domain repeated with the
    # given period and clipped (max) by the parent box
    r, P = 0.15, np.array([0.5, 0.4, 0.6])
    bbox = (-1.0, 1.0, -1.0, 1.0, -1.0, 1.0)
    rep = og.Repeat(bbox, og.Ball([0.0, 0.0, 0.0], r), P)
    rng = np.random.default_rng(1)
    x = rng.uniform(-0.95, 0.95, (400, 3))
    # brute force: nearest lattice copy of the ball
    ks = np.array(np.meshgrid(*[np.arange(-4, 5)] * 3, indexing="ij")).reshape(3, -1).T
    centres = ks * P
    dball = np.min(np.linalg.norm(x[:, None, :] - centres[None], axis=2), axis=1) - r
    want = np.maximum(dball, -_cube_interior_distance(x, bbox))
    assert np.allclose(rep.eval(x), want, atol=1e-12)
    # a lattice point inside the box is a ball centre
    assert np.isclose(rep.eval(np.array([[P[0], -P[1], 0.0]]))[0], -r)
    # outside the parent box the box decides
    far = np.array([[3.0, 0.0, 0.0]])
    assert np.isclose(rep.eval(far)[0], 2.0)


# ----------------------------------------------------- sharp / smooth CSG
def _disks():
    return og.Disk([0.0, 0.0], 1.0), og.Disk([1.0, 0.0], 1.0)


def _disk_d(x, c, r):
    return np.hypot(x[:, 0] - c[0], x[:, 1] - c[1]) - r


def test_sharp_intersection_is_the_lens():
    a, b = _disks()
    rng = np.random.default_rng(2)
    x = rng.uniform(-2, 3, (500, 2))
    d = og.Intersection([a, b]).eval(x)
    da, db = _disk_d(x, (0, 0), 1), _disk_d(x, (1, 0), 1)
    inside = (da < 0) & (db < 0)
    assert np.array_equal(d < 0, inside)
    # the lens's corner points (0.5, +-sqrt(3)/2) lie on the boundary
    c = np.array([[0.5, math.sqrt(3) / 2], [0.5, -math.sqrt(3) / 2]])
    assert np.allclose(og.Intersection([a, b]).eval(c), 0.0, atol=1e-15)
    # three sets: the intersection is inside all of them
    c3 = og.Disk([0.5, 0.8], 0.6)
    d3 = og.Intersection([a, b, c3]).eval(x)
    assert np.array_equal(d3 < 0, inside & (_disk_d(x, (0.5, 0.8), 0.6) < 0))


def test_smooth_ops_blend_only_inside_k():
    # smooth union / intersection / difference (signed_distance_functions.py
    # :332-334, :372-374, :412-414): equal to the sharp operation where the
    # two distances differ by at least k, and offset by k/4 where they are
    # equal (the blend's quadratic at its centre)
    a, b = _disks()
    k = 0.2
    rng = np.random.default_rng(3)
    x = rng.uniform(-2, 3, (2000, 2))
    da, db = _disk_d(x, (0, 0), 1), _disk_d(x, (1, 0), 1)
    far = np.abs(da - db) >= k
    u = og.Union([a, b], smoothness=k).eval(x)
    i = og.Intersection([a, b], smoothness=k).eval(x)
    assert np.allclose(u[far], np.minimum(da, db)[far], atol=1e-14)
    assert np.allclose(i[far], np.maximum(da, db)[far], atol=1e-14)
    assert np.all(u <= np.minimum(da, db) + 1e-15)
    assert np.all(i >= np.maximum(da, db) - 1e-15)
    # on the bisector x = 1/2 the two distances are equal
    y = np.linspace(-2, 2, 41)
    mid = np.column_stack((np.full_like(y, 0.5), y))
    dm = _disk_d(mid, (0, 0), 1)
    assert np.allclose(og.Union([a, b], smoothness=k).eval(mid), dm - k / 4, atol=1e-14)
    assert np.allclose(og.Intersection([a, b], smoothness=k).eval(mid), dm + k / 4, atol=1e-14)
    # smooth difference a - b: sharp max(da, -db) where |da + db| >= k, and
    # +k/4 where da = -db, i.e. on the ellipse with foci (0,0), (1,0) and
    # major semi-axis 1, whose axis points are (1.5, 0) and (-0.5, 0)
    dif = og.Difference([a, b], smoothness=k).eval(x)
    farm = np.abs(da + db) >= k
    assert np.allclose(dif[farm], np.maximum(da, -db)[farm], atol=1e-14)
    z = np.array([[1.5, 0.0], [-0.5, 0.0]])
    assert np.allclose(og.Difference([a, b], smoothness=k).eval(z), [0.5 + k / 4, -0.5 + k / 4],
                       atol=1e-14)


def test_smooth_difference_of_three_folds_from_the_back():
    # Difference([A, B, C], k) folds the reversed list (:423):
    # sdiff(sdiff(C, B), A) = A minus (B minus C); away from the blend zones
    # that is max(dA, -max(dB, -dC))
    A = og.Disk([0.0, 0.0], 1.5)
    B = og.Disk([0.5, 0.0], 1.0)
    C = og.Disk([0.7, 0.0], 0.4)
    k = 1e-3
    rng = np.random.default_rng(4)
    x = rng.uniform(-2, 2, (3000, 2))
    da, db, dc = _disk_d(x, (0, 0), 1.5), _disk_d(x, (0.5, 0), 1.0), _disk_d(x, (0.7, 0), 0.4)
    inner = np.maximum(db, -dc)
    want = np.maximum(da, -inner)
    away = (np.abs(db + dc) > k) & (np.abs(da + inner) > k)
    got = og.Difference([A, B, C], smoothness=k).eval(x)
    assert np.allclose(got[away], want[away], atol=1e-14)
    # C's disk lies inside B: its points are kept (A - (B - C)), not cut
    assert og.Difference([A, B, C], smoothness=k).eval(np.array([[0.7, 0.0]]))[0] < 0
    # the sharp difference with three sets cuts both
    assert og.Difference([A, B, C]).eval(np.array([[0.7, 0.0]]))[0] > 0


# ------------------------------------------------- rotation / stretch / translate
def _box2(x, lo, hi):
    # the reference's rectangle distance (minus the smallest slab distance)
    return -np.minimum(np.minimum(x[:, 0] - lo[0], hi[0] - x[:, 0]),
                       np.minimum(x[:, 1] - lo[1], hi[1] - x[:, 1]))


def test_rotation_quarter_turn_2d():
    # Rectangle [0,2]x[0,1] rotated by +pi/2 about the origin is [-1,0]x[0,2]
    r = og.Rectangle((0.0, 2.0, 0.0, 1.0), rotate=(math.pi / 2, 0, 0))
    rng = np.random.default_rng(5)
    x = rng.uniform(-1.5, 0.5, (300, 2)) * [1, 2]
    want = _box2(x, (-1.0, 0.0), (0.0, 2.0))
    assert np.allclose(r.eval(x), want, atol=1e-12)
    assert np.allclose(r.bbox, (-1.0, 0.0, 0.0, 2.0), atol=1e-12)


def _box3(x, lo, hi):
    s = np.concatenate([x - lo, hi - x], axis=1)
    return -s.min(1)


def test_rotation_quarter_turns_3d_order():
    # A quarter turn about x maps (x, y, z) -> (x, -z, y); about y
    # (x, y, z) -> (z, y, -x); about z (x, y, z) -> (-y, x, z).  The
    # reference undoes them in the order x, y, z on the query point
    # (signed_distance_functions.py:89-101), i.e. the object is rotated by
    # R = Rx Ry Rz.  Box [0,1]x[0,2]x[0,3]:
    lo, hi = np.zeros(3), np.array([1.0, 2.0, 3.0])
    box = (0.0, 1.0, 0.0, 2.0, 0.0, 3.0)
    Rx = np.array([[1, 0, 0], [0, 0, -1], [0, 1, 0]], dtype=float)
    Ry = np.array([[0, 0, 1], [0, 1, 0], [-1, 0, 0]], dtype=float)
    Rz = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]], dtype=float)
    rng = np.random.default_rng(6)
    x = rng.uniform(-3.5, 3.5, (400, 3))
    cases = [((math.pi / 2, 0, 0), Rx), ((0, math.pi / 2, 0), Ry), ((0, 0, math.pi / 2), Rz),
             ((math.pi / 2, math.pi / 2, 0), Rx @ Ry),
             ((math.pi / 2, 0, math.pi / 2), Rx @ Rz),
             ((math.pi / 2, math.pi / 2, math.pi / 2), Rx @ Ry @ Rz)]
    for rot, R in cases:
        c = og.Cube(box, rotate=rot)
        # the rotated object's distance at x is the box's at R^T x
        want = _box3(x @ R, lo, hi)
        assert np.allclose(c.eval(x), want, atol=1e-12), rot


def test_stretch_and_translate():
    # stretch [2, 0] turns the unit disk into the ellipse x^2/4 + y^2 = 1
    e = og.Disk([0.0, 0.0], 1.0, stretch=[2.0, 0.0])
    th = np.linspace(0, 2 * math.pi, 50)
    ring = np.column_stack((2 * np.cos(th), np.sin(th)))
    assert np.allclose(e.eval(ring), 0.0, atol=1e-12)
    assert np.isclose(e.eval(np.array([[1.0, 0.0]]))[0], -0.5)
    assert np.allclose(e.bbox, (-2.0, 2.0, -1.0, 1.0))
    # along a diagonal direction: stretch by |v| along v/|v|
    v = np.array([1.0, 1.0]) * 1.5
    e2 = og.Disk([0.0, 0.0], 1.0, stretch=list(v))
    u = v / np.linalg.norm(v)
    w = np.array([-u[1], u[0]])
    pts = np.array([np.linalg.norm(v) * u, w, -w])
    assert np.allclose(e2.eval(pts), 0.0, atol=1e-12)
    # translation moves the shape and its bbox
    t = og.Ball([0.0, 0.0, 0.0], 0.5, translate=[1.0, -2.0, 0.5])
    assert np.isclose(t.eval(np.array([[1.0, -2.0, 0.5]]))[0], -0.5)
    assert np.isclose(t.eval(np.array([[1.0, -2.0, 2.0]]))[0], 1.0)
    assert np.allclose(t.bbox, (0.5, 1.5, -2.5, -1.5, 0.0, 1.0))


# ---------------------------------------------------------- sliver removal
def _sliver_fixture():
    # a unit-ish tetrahedron plus a flat one (a sliver: four nearly
    # co-circular points) sharing no vertex with it
    p = np.array([
        [0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0],
        [3.0, 0.0, 0.0], [4.0, 0.0, 0.05], [4.0, 1.0, 0.0], [3.0, 1.0, 0.05],
    ])
    t = np.array([[0, 1, 2, 3], [4, 5, 6, 7]])
    vol = og.simp_vol(p, t)
    t[vol < 0, :2] = t[vol < 0, 1::-1]
    return p, t


def _circumradius(q):
    A = np.stack([q[0] - q[3], q[1] - q[3], q[2] - q[3]])
    c = np.linalg.solve(2 * A, (A * A).sum(1))
    return float(np.linalg.norm(c))


def _fd_direction(q):
    g = np.zeros(3)
    h = 1e-7
    for k in range(3):
        a, b = q.copy(), q.copy()
        a[0, k] += h
        b[0, k] -= h
        g[k] = (_circumradius(a) ** 2 - _circumradius(b) ** 2) / (2 * h)
    return g / np.linalg.norm(g)


@pytest.mark.parametrize("gamma", [1.0, 2.0])
def test_sliver_removal_step_control(gamma):
    # mesh_generator.py:204-286 with the triangulation held fixed: the
    # first vertex of each sliver moves by step * h0 along the normalised
    # circumradius gradient; step starts at 0.1, is divided by gamma when
    # the sliver count drops (the first pass: from infinity), and divided by
    # 0.8 when the count repeats
    p0, t = _sliver_fixture()
    seen = []

    def fixed(p):
        seen.append(p.copy())
        return t.copy()

    dom = og.Cube((-1.0, 6.0, -1.0, 2.0, -1.0, 2.0))
    h0 = 0.5
    odm.sliver_removal(p0, dom, h0, triangulate=fixed, max_iter=4, gamma=gamma,
                       min_dh_angle_bound=10.0, verbose=0)
    assert len(seen) == 4
    steps = [0.1 / gamma, 0.1 / gamma / 0.8, 0.1 / gamma / 0.8 / 0.8]
    for k in range(3):
        dp = seen[k + 1] - seen[k]
        moved = np.nonzero(np.abs(dp).sum(1))[0]
        assert moved.tolist() == [t[1, 0]]
        q = seen[k][t[1]]
        want = steps[k] * h0 * _fd_direction(q)
        assert np.allclose(dp[t[1, 0]], want, rtol=1e-5, atol=1e-9)


def test_sliver_removal_returns_when_clean():
    p0, t = _sliver_fixture()
    good = np.array([[0, 1, 2, 3]])
    dom = og.Cube((-1.0, 6.0, -1.0, 2.0, -1.0, 2.0))
    p, c = odm.sliver_removal(p0[:4], dom, 0.5, triangulate=lambda p: good.copy(), verbose=0)
    # nothing moves; fix_mesh only renumbers the points (sorted unique rows)
    assert np.array_equal(np.unique(p, axis=0), np.unique(p0[:4], axis=0)) and len(c) == 1
    assert np.array_equal(np.sort(p[c[0]], axis=0), np.sort(p0[:4], axis=0))


def test_smooth_difference_sliver_known_answer():
    # reference tests/test_smooth_sets.py:8-22: generate_mesh + sliver_removal
    # on Ball - Cube with smoothness 0.2, h = 0.1: 9004 +- 100 cells
    cube1 = og.Cube((-0.5, 0.5, -0.5, 0.5, -0.5, 0.5))
    ball1 = og.Ball((0.0, 0.0, 0.5), 0.85)
    domain = og.Difference([ball1, cube1], smoothness=0.20)
    p, t = odm.generate_mesh(domain, 0.10, verbose=0)
    p, t = odm.sliver_removal(p, domain, 0.10, verbose=0)
    assert abs(t.shape[0] - 9004) < 100
    # and no cell has a dihedral angle below 10 degrees
    ele, _ = odm.calc_slivers(p, t, 10 * math.pi / 180, math.pi)
    assert len(ele) == 0
