"""Known answers of the reference's own test suite, run through this
package's API on the GPU, plus domains given as plain Python functions.

Each test cites the reference test it mirrors; the asserted numbers and
tolerances are the reference's."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    from seismicmesh_amd import _lib
    _lib.require_gpu()


def _quiet(fn, *a, **k):
    import contextlib
    import io
    with contextlib.redirect_stdout(io.StringIO()):
        return fn(*a, **k)


def _slivers(p, t, lo=10.0, hi=180.0):
    from oracle import geometry as og
    dh = og.calc_dihedral_angles(p, t)
    out = np.argwhere((dh[:, 0] < lo * math.pi / 180) | (dh[:, 0] > hi * math.pi / 180))
    return np.unique(np.floor(out / 6).astype(int))


# ------------------------------------------------------ callable domains
def test_callable_domain_matches_geometry_object():
    """A domain given as a plain function (evaluated on the host through the
    library's callback) gives the same FP64 iterations, bit for bit, as the
    same disk given as a geometry object (evaluated on the device)."""
    from seismicmesh_amd import generation, geometry as geo
    h = 0.06
    st = generation.prepare(geo.Disk([0.0, 0.0], 1.0), h, verbose=0)
    fn = generation.prepare(lambda x: np.sqrt(x[:, 0] ** 2 + x[:, 1] ** 2) - 1.0, h, verbose=0,
                            bbox=(-1.0, 1.0, -1.0, 1.0))
    assert np.array_equal(st["p"], fn["p"])
    a, b = generation.make_mesher(st), generation.make_mesher(fn)
    a.set_points(st["p"])
    b.set_points(fn["p"])
    for it in range(8):
        a.run(1)
        b.run(1)
        assert np.array_equal(a.points(), b.points()), "iteration %d" % it
    a.triangulate()
    b.triangulate()
    assert np.array_equal(a.cells(), b.cells())
    a.close()
    b.close()


def test_callable_domain_vs_oracle_fp64():
    from oracle import distmesh as odm
    from seismicmesh_amd import generation

    def sdf(x):
        return np.maximum(np.abs(x[:, 0]) - 0.8, np.abs(x[:, 1]) - 0.5)

    rng = np.random.default_rng(2)
    g = np.stack(np.meshgrid(np.linspace(-0.85, 0.85, 40), np.linspace(-0.55, 0.55, 28)),
                 -1).reshape(-1, 2)
    p0 = g + rng.uniform(-0.004, 0.004, g.shape)
    bbox = (-0.8, 0.8, -0.5, 0.5)
    st = generation.prepare(sdf, 0.05, verbose=0, bbox=bbox, points=p0)
    m = generation.make_mesher(st)
    m.set_points(p0)
    opts = dict(odm.GEN_DEFAULTS)
    opts.update(verbose=0, points=p0, bbox=bbox)
    s = odm.setup(sdf, 0.05, opts)
    p = p0.copy()
    for it in range(6):
        m.run(1)
        p, _, _ = odm.force_step(p, odm.delaunay(p), s, [])
        assert np.array_equal(m.points(), p), "iteration %d" % it
    m.close()


def test_3dmesher_sdf_callable_known_answer():
    # reference tests/test_3dmesher_SDF.py:8-50: a cylinder given as a
    # function with a bbox, callable sizing, 100 iterations, then
    # sliver_removal: volume 6.28 +- hmin
    import seismicmesh_amd as sm
    from seismicmesh_amd import mesh_utils as mu
    hmin = 0.10
    bbox = (-1.0, 1.0, -1.0, 1.0, -1.0, 1.0)

    def cylinder(p):
        r, z = np.sqrt(p[:, 0] ** 2 + p[:, 1] ** 2), p[:, 2]
        d1, d2, d3 = r - 1.0, z - 1.0, -z - 1.0
        d4, d5 = np.sqrt(d1 ** 2 + d2 ** 2), np.sqrt(d1 ** 2 + d3 ** 2)
        d = np.maximum.reduce([d1, d2, d3])
        ix = (d1 > 0) * (d2 > 0)
        d[ix] = d4[ix]
        ix = (d1 > 0) * (d3 > 0)
        d[ix] = d5[ix]
        return d

    def EF(p):
        return np.array([hmin] * len(p))

    p, t = sm.generate_mesh(bbox=bbox, domain=cylinder, h0=hmin, edge_length=EF, max_iter=100,
                            verbose=0)
    p, t = sm.sliver_removal(points=p, domain=cylinder, edge_length=EF, h0=hmin, bbox=bbox,
                             verbose=0)
    vol = mu.simp_vol(p, t).sum()
    print("callable cylinder: %d vertices, %d cells, volume %.4f" % (len(p), len(t), vol))
    assert abs(vol - 6.28) <= hmin


# ------------------------------------------------------------ 2D answers
def _r0m_sizing():
    import seismicmesh_amd as sm
    bbox = (-10000.0, 0.0, 0.0, 10000.0)
    vs = np.zeros((200, 200))
    vs[0:150, :] = 1000
    ef = _quiet(sm.get_sizing_function_from_segy, None, bbox=bbox, grade=0.0, grad=0.0, wl=5,
                freq=2.0, hmin=75, hmax=10e6, velocity_data=vs, nz=200, nx=200)
    return sm.Rectangle(bbox), ef


@pytest.mark.parametrize("r0m_is_h0,answer", [(False, (15460, 30464)), (True, (10566, 20771))])
def test_r0m_values_known_answer(r0m_is_h0, answer):
    # reference tests/test_2dmesher_r0m_values.py:14-125
    import seismicmesh_amd as sm
    rect, ef = _r0m_sizing()
    assert ef.eval((-5000, 5000)) == 100
    assert ef.eval((-1, 5000)) == 150
    p, t = _quiet(sm.generate_mesh, rect, ef, h0=125, perform_checks=True, r0m_is_h0=r0m_is_h0)
    print("r0m_is_h0=%s: %d vertices, %d cells" % (r0m_is_h0, len(p), len(t)))
    assert np.allclose([len(p), len(t)], answer, atol=100)


@pytest.mark.parametrize("h", list(np.logspace(-1.0, -2.0, num=5)))
def test_min_quality_all_sizes(h):
    # reference tests/test_2d_min_qual.py:7-31, each of its five sizes
    import seismicmesh_amd as sm
    rect = sm.Rectangle((0.0, 1.0, 0.0, 1.0))
    disk0 = sm.Disk([0.0, 0.0], 0.25)
    quarter = sm.Intersection([sm.Difference([rect, disk0]), sm.Disk([0.0, 0.0], 1.0)])
    p, t = sm.generate_mesh(domain=quarter, edge_length=lambda x: h + 0.1 * np.abs(disk0.eval(x)),
                            h0=h, verbose=0)
    q = sm.mesh_utils.simp_qual(p, t)
    print("h=%.4f: %d vertices, min quality %.4f" % (h, len(p), q.min()))
    assert np.amin(q) > 0.60


# ------------------------------------------------------------ 3D answers
def test_pfix_3d_known_answer():
    # reference tests/test_pfix.py:8-28
    import seismicmesh_amd as sm
    hmin = 0.05
    bbox = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    pfix = np.linspace((0.0, 0.0, 0.0), (1.0, 0.0, 1.0), int(np.sqrt(2) / hmin))
    pfix = np.vstack((pfix, sm.geometry.corners(bbox)))
    p, t = sm.generate_mesh(domain=sm.Cube(bbox), edge_length=hmin, pfix=pfix, verbose=0)
    for f in pfix:
        d = ((p - f) ** 2).sum(1)
        assert np.isclose(np.min(d), 0.0)


def test_smooth_difference_sliver_known_answer():
    # reference tests/test_smooth_sets.py:8-22
    import seismicmesh_amd as sm
    cube1 = sm.Cube((-0.5, 0.5, -0.5, 0.5, -0.5, 0.5))
    ball1 = sm.Ball((0.0, 0.0, 0.5), 0.85)
    domain = sm.Difference([ball1, cube1], smoothness=0.20)
    p, t = sm.generate_mesh(domain=domain, edge_length=0.10, verbose=0)
    p, t = sm.sliver_removal(points=p, domain=domain, edge_length=0.10, verbose=0)
    print("smooth difference: %d cells (reference 9004 +- 100)" % len(t))
    assert np.abs(t.shape[0] - 9004) < 100


@pytest.mark.parametrize("h", list(np.logspace(-1.0, -1.3, num=5)))
def test_3d_sliver_cylinder_all_sizes(h):
    # reference tests/test_3d_sliver.py:40-47, 67-74
    import seismicmesh_amd as sm
    cyl = sm.Cylinder(h=1.0, r=1.0)
    p, t = sm.generate_mesh(domain=cyl, edge_length=h, verbose=0)
    p, t = sm.sliver_removal(points=p, domain=cyl, edge_length=h, verbose=0)
    assert len(_slivers(p, t)) == 0


@pytest.mark.parametrize("h", list(np.logspace(-1.0, -1.2, num=5)))
def test_3d_sliver_l_shape_all_sizes(h):
    # reference tests/test_3d_sliver.py:50-64, 77-82
    import seismicmesh_amd as sm
    tol = h / 10
    cube0 = sm.Cube((-1.0, 1.0, -1.0, 1.0, -1.0, tol))
    cube1 = sm.Cube((-1.0, 1.0, 0.0, 1.0, -tol, 1.0))
    cube2 = sm.Cube((-tol, 1.0, -1.0, 1.0, -tol, 1.0))
    domain = sm.Union([cube0, cube1, cube2])
    p, t = sm.generate_mesh(domain=domain, edge_length=h, verbose=0)
    p, t = sm.sliver_removal(domain=domain, points=p, edge_length=h, verbose=0)
    assert len(_slivers(p, t)) == 0


@pytest.mark.parametrize("h", list(np.logspace(-1.0, -1.3, num=5)))
def test_3d_sliver_box_refined_all_sizes(h):
    # reference tests/test_3d_sliver.py:20-37, 85-91 (callable sizing)
    import seismicmesh_amd as sm
    cube = sm.Cube((-1.0, 1.0, -1.0, 1.0, -1.0, 1.0))

    def edge_length(x):
        return h + 0.1 * np.sqrt(x[:, 0] ** 2 + x[:, 1] ** 2 + x[:, 2] ** 2)

    p, t = sm.generate_mesh(domain=cube, h0=h, edge_length=edge_length, verbose=0)
    p, t = sm.sliver_removal(domain=cube, points=p, h0=h, edge_length=edge_length, verbose=0)
    assert len(_slivers(p, t)) == 0
