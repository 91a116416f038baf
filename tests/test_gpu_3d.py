"""The 3D force loop and sliver removal on the GPU against the oracle.

3D FP64 iterations are compared to 1e-12 (absolute, unit-sized domains)
rather than bit for bit: the reference sums L**3 and h**3, and numpy's
power on x86-64 runs a SIMD pow (AVX-512 SVML where the CPU has it) that is
not correctly rounded, so its cubes depend on the host CPU; the device
computes correctly rounded cubes.  Edges and cells are compared exactly."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    from seismicmesh_amd import _lib
    _lib.require_gpu()


def _jittered_ball(n_side=16, seed=0, r=0.97):
    rng = np.random.default_rng(seed)
    ax = np.linspace(-1, 1, n_side)
    g = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
    g = g + rng.uniform(-0.3, 0.3, g.shape) * (2.0 / n_side)
    return g[np.linalg.norm(g, axis=1) < r]


def _oracle_state(domain, h, **kw):
    from oracle import distmesh as odm
    opts = dict(odm.GEN_DEFAULTS)
    opts.update(kw)
    opts["verbose"] = 0
    return odm.setup(domain, h, opts)


def _mesher(domain, h, p, precision=64):
    from seismicmesh_amd import generation
    st = generation.prepare(domain, h, verbose=0, precision=precision, points=p)
    st["nfix"] = 0
    m = generation.make_mesher(st)
    m.set_points(p)
    return m


def test_3d_iterations_match_oracle():
    from oracle import distmesh as odm
    from oracle import geometry as og
    from seismicmesh_amd import geometry as geo
    h = 0.13
    p0 = _jittered_ball()
    m = _mesher(geo.Ball([0, 0, 0], 1.0), h, p0)
    s = _oracle_state(og.Ball([0, 0, 0], 1.0), h, points=p0)
    p = p0.copy()
    for it in range(6):
        t_all = odm.delaunay(p)
        m.triangulate()
        # the exact host tetrahedralisation keeps the same cells and edges
        got = {tuple(sorted(c)) for c in m.cells().tolist()}
        t = odm.remove_triangles_outside(p, t_all, s["fd"], s["geps"])
        assert got == {tuple(sorted(c)) for c in t.tolist()}, "iteration %d" % it
        assert np.array_equal(m.edges(), odm.get_edges(t)), "iteration %d" % it
        m.run(1)
        p_ref, _, _ = odm.force_step(p, t_all, s, [])
        q = m.points()
        assert np.abs(q - p_ref).max() <= 1e-12, "iteration %d" % it
        p = q  # continue both from the device state


def test_3d_fp32_update_within_tolerance():
    from oracle import distmesh as odm
    from oracle import geometry as og
    from seismicmesh_amd import geometry as geo
    h = 0.13
    p0 = _jittered_ball(seed=2)
    m = _mesher(geo.Ball([0, 0, 0], 1.0), h, p0, precision=32)
    m.run(1)
    s = _oracle_state(og.Ball([0, 0, 0], 1.0), h, points=p0)
    q, _, _ = odm.force_step(p0, odm.delaunay(p0), s, [])
    dp_gpu, dp_ref = m.points() - p0, q - p0
    scale = np.abs(dp_ref).max()
    assert scale > 0
    assert np.abs(dp_gpu - dp_ref).max() <= 1e-5 * scale


def test_sliver_step_matches_numpy():
    # one selection + perturbation on the device against the reference's
    # numpy statements (mesh_generator.py:222-272) on the same cells
    from oracle import distmesh as odm
    from oracle import geometry as og
    from seismicmesh_amd import _lib, geometry as geo
    import ctypes
    h = 0.13
    p0 = _jittered_ball(seed=5)
    m = _mesher(geo.Ball([0, 0, 0], 1.0), h, p0)
    m.triangulate()
    t = m.cells().astype(np.int64)
    p = m.points()
    min_dh, max_dh = 15 * math.pi / 180, 165 * math.pi / 180
    ele, _ = odm.calc_slivers(p, t, min_dh, max_dh)
    n = ctypes.c_int64()
    _lib.check(_lib.lib.sm_mesher_find_slivers(m.h, min_dh, max_dh, ctypes.byref(n)), "find")
    assert n.value == len(ele) and len(ele) > 0
    perturb = og.calc_circumsphere_grad(p[t[ele, 0]], p[t[ele, 1]], p[t[ele, 2]], p[t[ele, 3]])
    perturb[np.isinf(perturb)] = 1.0
    perturb /= (np.sum(np.abs(perturb) ** 2, axis=-1) ** 0.5)[:, None]
    want = p.copy()
    want[t[ele, 0]] += 0.1 * h * perturb
    _lib.check(_lib.lib.sm_mesher_perturb_slivers(m.h, 0.1 * h), "perturb")
    got = m.points()
    moved = np.unique(t[ele, 0])
    assert np.array_equal(np.nonzero((got != p).any(1))[0], moved)
    assert np.abs(got - want).max() <= 1e-13


def test_3d_ball_volume_and_orientation():
    import seismicmesh_amd as sm
    from seismicmesh_amd import mesh_utils as mu
    h = 0.1
    p, t = sm.generate_mesh(sm.Ball([0, 0, 0], 1.0), h, verbose=0, max_iter=30)
    vol = mu.simp_vol(p, t)
    # cells come out positively oriented; the final boundary projection
    # (mesh_generator.py:741-759) may flatten a few boundary cells
    assert (vol > 0).mean() > 0.999
    assert abs(vol.sum() - 4.0 / 3.0 * math.pi) < 0.15


def test_3d_sliver_cylinder():
    # reference tests/test_3d_sliver.py:67-74 at its coarsest size
    import seismicmesh_amd as sm
    from oracle import geometry as og
    h = 0.1
    cyl = sm.Cylinder(h=1.0, r=1.0)
    p, t = sm.generate_mesh(domain=cyl, edge_length=h, verbose=0)
    p, t = sm.sliver_removal(points=p, domain=cyl, edge_length=h, verbose=0)
    dh = og.calc_dihedral_angles(p, t)
    bad = np.argwhere((dh[:, 0] < 10 * math.pi / 180) | (dh[:, 0] > math.pi))
    assert len(bad) == 0
    # Cylinder(h, r) spans z in [-h/2, h/2]: volume pi r^2 h
    from seismicmesh_amd import mesh_utils as mu
    assert abs(mu.simp_vol(p, t).sum() - math.pi) < h


@pytest.mark.parametrize("axis,nslabs,threads", [(1, 16, 8), (0, 5, 5), (2, 3, 2)])
def test_slab_build_iterations_identical_to_serial(axis, nslabs, threads):
    """The 3D loop with its tetrahedralisation built by several threads
    (slabs inserted into one shared triangulation, each piece of cells
    copied and filtered while the next is handed out) moves the points
    exactly as the serial build does, from the degenerate initial lattice
    on."""
    from seismicmesh_amd import generation, geometry as geo
    dom = geo.Difference([geo.Union([geo.Ball([0.0, 0.0, 0.0], 0.8),
                                     geo.Cube((-0.2, 1.0, -0.3, 0.3, -0.3, 0.3))]),
                          geo.Ball([0.3, 0.3, 0.3], 0.3)])
    h = 0.05
    st = generation.prepare(dom, h, verbose=0, precision=32)
    a = generation.make_mesher(st, slabs=1)
    b = generation.make_mesher(st, slabs=1)
    b.set_slabs(nslabs, axis=axis, threads=threads)
    a.set_points(st["p"])
    b.set_points(st["p"])
    for it in range(4):
        a.run(1)
        b.run(1)
        assert np.array_equal(a.points(), b.points()), "iteration %d" % it
    a.triangulate()
    b.triangulate()
    assert np.array_equal(a.cells(), b.cells())
    s = b.stats()
    print("%d points, slabs %d axis %d threads %d: %d parallel builds, %d points retried, "
          "host %.1f ms per build" % (len(st["p"]), nslabs, axis, threads, s["parallel_builds"],
                                      s["retried_points"], s["host_tet_us"] / 1e3 /
                                      s["host_rebuilds"]))
    assert s["parallel_builds"] == s["host_rebuilds"] >= 5
    a.close()
    b.close()
