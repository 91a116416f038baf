"""Device kernels through the C ABI against the oracle (GPU)."""
import ctypes
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    from seismicmesh_amd import _lib
    _lib.require_gpu()
    return torch.device("cuda", 0)


def _t(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


SHAPES_2D = [
    ("disk", lambda m: m.Disk([0.1, -0.2], 1.3)),
    ("rect", lambda m: m.Rectangle((-1.0, 1.0, -0.5, 0.7))),
    ("diff", lambda m: m.Difference([m.Rectangle((0.0, 1.0, 0.0, 1.0)), m.Disk([0, 0], 0.25)])),
    ("quarter", lambda m: m.Intersection([m.Difference([m.Rectangle((0.0, 1.0, 0.0, 1.0)),
                                                        m.Disk([0, 0], 0.25)]),
                                          m.Disk([0, 0], 1.0)])),
    ("smooth_union", lambda m: m.Union([m.Disk([0, 0], .5), m.Disk([.6, 0], .4)], 0.1)),
    ("smooth_inter", lambda m: m.Intersection([m.Disk([0, 0], .5), m.Disk([.3, 0], .4)], 0.1)),
    ("smooth_diff", lambda m: m.Difference([m.Disk([0, 0], .8), m.Disk([.3, 0], .4),
                                            m.Disk([-.4, 0], .2)], 0.15)),
    ("translated", lambda m: m.Disk([0, 0], 0.5, translate=[0.2, -0.1])),
]
SHAPES_3D = [
    ("ball", lambda m: m.Ball([0.1, 0.2, -0.1], 0.7)),
    ("cube", lambda m: m.Cube((-1.0, 0.5, -0.3, 1.2, 0.0, 0.4))),
    ("torus", lambda m: m.Torus(0.8, 0.2)),
    ("prism", lambda m: m.Prism(0.6, 0.4)),
    ("cylinder", lambda m: m.Cylinder(h=1.2, r=0.5)),
    ("csg", lambda m: m.Difference([m.Union([m.Ball([0, 0, 0], 0.8), m.Cube((0., 1., 0., 1., 0., 1.))]),
                                    m.Cylinder(h=3.0, r=0.2)])),
    ("smooth_diff3", lambda m: m.Difference([m.Ball((0.0, 0.0, 0.5), 0.85),
                                             m.Cube((-.5, .5, -.5, .5, -.5, .5))], smoothness=0.2)),
    ("repeat", lambda m: m.Repeat((-1.0, 1.0, -1.0, 1.0, -1.0, 1.0), m.Ball([0, 0, 0], 0.2),
                                  [0.5, 0.5, 0.5])),
]


def _eval_mode(tape, dim, x, deep, dev):
    import torch
    from seismicmesh_amd import _lib
    from seismicmesh_amd import geometry as geo
    arr = geo.tape_array(tape)
    arr[0].pad[0] = 1.0 if deep else 0.0
    tt = torch.frombuffer(bytearray(bytes(arr)), dtype=torch.uint8).to(dev)
    xt = _t(x, dev)
    d = torch.empty(len(x), dtype=torch.float64, device=dev)
    _lib.check(_lib.lib.sm_sdf_eval(tt.data_ptr(), len(arr), dim, xt.data_ptr(), len(x),
                                    d.data_ptr(), None), "sm_sdf_eval")
    torch.cuda.synchronize()
    return d.cpu().numpy()


@pytest.mark.parametrize("name,make", SHAPES_2D + SHAPES_3D, ids=[s[0] for s in SHAPES_2D + SHAPES_3D])
def test_sdf_bit_exact(dev, name, make):
    from oracle import geometry as og
    from seismicmesh_amd import geometry as geo
    a, b = make(geo), make(og)
    rng = np.random.default_rng(0)
    x = rng.uniform(-2, 2, (20000, a.dim))
    ref = b.eval(x)
    assert np.array_equal(a.eval(x), ref), name
    # both interpreters: registers (default) and the scratch-stack one
    for deep in (False, True):
        assert np.array_equal(_eval_mode(a.tape(), a.dim, x, deep, dev), ref), (name, deep)


def test_sdf_rotated_within_rounding(dev):
    from oracle import geometry as og
    from seismicmesh_amd import geometry as geo
    rng = np.random.default_rng(1)
    x = rng.uniform(-2, 2, (5000, 2))
    a = geo.Rectangle((-1.0, 1.0, -0.5, 0.5), rotate=[0.4, 0, 0])
    b = og.Rectangle((-1.0, 1.0, -0.5, 0.5), rotate=[0.4, 0, 0])
    # the reference rotates with a BLAS product; only rounding may differ
    assert np.allclose(a.eval(x), b.eval(x), rtol=0, atol=1e-14)
    x3 = rng.uniform(-2, 2, (5000, 3))
    a = geo.Cube((-1.0, 1.0, -0.5, 0.5, -0.2, 0.3), rotate=[0.3, 0.2, 0.1], stretch=[1.0, 2.0, 1.0])
    b = og.Cube((-1.0, 1.0, -0.5, 0.5, -0.2, 0.3), rotate=[0.3, 0.2, 0.1], stretch=[1.0, 2.0, 1.0])
    assert np.allclose(a.eval(x3), b.eval(x3), rtol=0, atol=1e-13)


def test_sizing_vs_scipy(dev):
    from oracle.sizing import build_interpolant
    from seismicmesh_amd.sizing import GridSizing
    rng = np.random.default_rng(2)
    bbox = (-12000.0, 0.0, 0.0, 67000.0)
    vals = rng.uniform(75, 500, (191, 540)).astype(np.float32)
    g = GridSizing(vals, bbox)
    rgi = build_interpolant(vals.astype(np.float64), bbox)
    x = np.column_stack([rng.uniform(-12500, 500, 20000), rng.uniform(-1000, 68000, 20000)])
    got = g(x)
    want = rgi(x)
    # the FP64 interpolation in scipy's order, rounded to FP32 once
    assert np.array_equal(got, want.astype(np.float32))
    # 3D trilinear
    vals3 = rng.uniform(150, 300, (21, 34, 34)).astype(np.float32)
    b3 = (-4200.0, 0.0, 0.0, 13520.0, 0.0, 13520.0)
    g3 = GridSizing(vals3, b3)
    rgi3 = build_interpolant(vals3.astype(np.float64), b3)
    x3 = np.column_stack([rng.uniform(-4300, 100, 5000), rng.uniform(-100, 13600, 5000),
                          rng.uniform(0, 13520, 5000)])
    assert np.array_equal(g3(x3), rgi3(x3).astype(np.float32))


def _unique_edges_dev(pairs, nvert, dev):
    import torch
    from seismicmesh_amd import _lib
    n = len(pairs)
    pt = _t(pairs.astype(np.int32), dev) if n else torch.zeros((1, 2), dtype=torch.int32, device=dev)
    out = torch.empty((max(n, 1), 2), dtype=torch.int32, device=dev)
    cnt = torch.zeros(1, dtype=torch.int64, device=dev)
    wsb = _lib.lib.sm_unique_edges_workspace(n, nvert)
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    _lib.check(_lib.lib.sm_unique_edges(pt.data_ptr(), n, nvert, out.data_ptr(), cnt.data_ptr(),
                                        ws.data_ptr(), wsb, None), "sm_unique_edges")
    torch.cuda.synchronize()
    e = int(cnt.item())
    return out[:e].cpu().numpy()


@pytest.mark.parametrize("case", ["empty", "single", "random", "dense_rows", "huge_row",
                                  "many_vertices", "wave_rows", "mesh"])
def test_unique_edges_bit_exact(dev, case):
    from oracle import distmesh as odm
    from oracle import geometry as og
    rng = np.random.default_rng(3)
    if case == "empty":
        pairs, nv = np.zeros((0, 2), np.int32), 10
    elif case == "single":
        pairs, nv = np.array([[5, 2]], np.int32), 6
    elif case == "random":
        nv = 5000
        pairs = rng.integers(0, nv, (40000, 2)).astype(np.int32)
    elif case == "dense_rows":  # buckets longer than the register sort
        nv = 50
        pairs = rng.integers(0, nv, (6000, 2)).astype(np.int32)
    elif case == "huge_row":  # one bucket beyond the LDS tile
        nv = 20000
        pairs = np.column_stack([np.zeros(12000, np.int32), rng.integers(1, nv, 12000)]).astype(np.int32)
    elif case == "many_vertices":  # prefix sums past the one-workgroup scan
        nv = 300000
        pairs = rng.integers(0, nv, (400000, 2)).astype(np.int32)
    elif case == "wave_rows":  # buckets of 17..64 ids (the one-wave sort)
        nv = 3000
        a = np.repeat(np.arange(nv, dtype=np.int32), 40)
        pairs = np.column_stack([a, rng.integers(0, nv, len(a))]).astype(np.int32)
    else:
        p = rng.random((3000, 2))
        t = odm.delaunay(p)
        pairs = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]]).astype(np.int32)
        nv = len(p)
    got = _unique_edges_dev(pairs, nv, dev)
    assert np.array_equal(got, og.unique_edges(pairs).reshape(-1, 2))


def test_cell_edges_bit_exact(dev):
    import torch
    from oracle import distmesh as odm
    from seismicmesh_amd import _lib
    rng = np.random.default_rng(4)
    for dim in (2, 3):
        p = rng.random((2000, dim))
        t = odm.delaunay(p).astype(np.int32)
        tt = _t(t, dev)
        npairs = len(t) * (3 if dim == 2 else 6)
        out = torch.empty((npairs, 2), dtype=torch.int32, device=dev)
        cnt = torch.zeros(1, dtype=torch.int64, device=dev)
        wsb = _lib.lib.sm_cell_edges_workspace(len(t), dim, len(p))
        ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
        _lib.check(_lib.lib.sm_cell_edges(tt.data_ptr(), len(t), dim, len(p), out.data_ptr(),
                                          cnt.data_ptr(), ws.data_ptr(), wsb, None), "cell_edges")
        torch.cuda.synchronize()
        got = out[:int(cnt.item())].cpu().numpy()
        assert np.array_equal(got, odm.get_edges(t))


def test_tet_edges_in_every_cell_order(dev):
    """Tetrahedra count their edges by runs of equal lowest id (one atomic
    per run of lanes): cells sorted by lowest id (the host
    tetrahedralisation's order, long runs), reversed, shuffled, with their
    vertices permuted, and with half of them masked out of a 3D handle's
    keep flags all give the reference's sorted unique edges."""
    import ctypes
    import torch
    from oracle import distmesh as odm
    from seismicmesh_amd import _lib
    rng = np.random.default_rng(8)
    p = np.ascontiguousarray(rng.random((30000, 3)))
    cap = 8 * len(p) + 64
    tet = np.zeros((cap, 4), np.int32)
    ins = np.zeros(len(p), np.uint8)
    m = _lib.lib.sm_host_delaunay3d(p.ctypes.data_as(ctypes.c_void_p), len(p),
                                    tet.ctypes.data_as(ctypes.c_void_p), cap,
                                    ins.ctypes.data_as(ctypes.c_void_p))
    assert m > 0
    t0 = tet[:m]
    assert (np.diff(t0[:, 0]) >= 0).all()  # sorted by lowest id
    orders = {"sorted": t0, "reversed": t0[::-1], "shuffled": t0[rng.permutation(m)],
              "permuted vertices": t0[:, rng.permutation(4)],
              "half": t0[rng.random(m) < 0.5]}
    for name, t in orders.items():
        t = np.ascontiguousarray(t, dtype=np.int32)
        tt = _t(t, dev)
        out = torch.empty((6 * len(t), 2), dtype=torch.int32, device=dev)
        cnt = torch.zeros(1, dtype=torch.int64, device=dev)
        wsb = _lib.lib.sm_cell_edges_workspace(len(t), 3, len(p))
        ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
        _lib.check(_lib.lib.sm_cell_edges(tt.data_ptr(), len(t), 3, len(p), out.data_ptr(),
                                          cnt.data_ptr(), ws.data_ptr(), wsb, None), "cell_edges")
        torch.cuda.synchronize()
        got = out[:int(cnt.item())].cpu().numpy()
        assert np.array_equal(got, odm.get_edges(t)), name


def test_dihedral_and_circumsphere_grad(dev):
    import torch
    from oracle import geometry as og
    from seismicmesh_amd import _lib
    rng = np.random.default_rng(5)
    n = 4000
    p = rng.uniform(-1, 1, (4 * n, 3))
    t = np.arange(4 * n, dtype=np.int32).reshape(n, 4)
    pt, tt = _t(p, dev), _t(t, dev)
    ang = torch.empty(6 * n, dtype=torch.float64, device=dev)
    _lib.check(_lib.lib.sm_dihedral_angles(pt.data_ptr(), tt.data_ptr(), n, ang.data_ptr(), None),
               "dihedral")
    ref = og.calc_dihedral_angles(p, t).ravel()
    got = ang.cpu().numpy()
    assert np.allclose(got, ref, rtol=0, atol=1e-12, equal_nan=True)
    q = [_t(p[t[:, k]], dev) for k in range(4)]
    g = torch.empty((n, 3), dtype=torch.float64, device=dev)
    _lib.check(_lib.lib.sm_circumsphere_grad(q[0].data_ptr(), q[1].data_ptr(), q[2].data_ptr(),
                                             q[3].data_ptr(), n, g.data_ptr(), None), "cgrad")
    ref = og.calc_circumsphere_grad(*[p[t[:, k]] for k in range(4)])
    got = g.cpu().numpy()
    # same quantity (gradient of the squared circumradius), different
    # association of the arithmetic: compare directions and magnitudes
    ok = np.isfinite(ref).all(1) & (np.abs(ref).max(1) < 1e8)
    assert np.allclose(got[ok], ref[ok], rtol=1e-7, atol=1e-9 * np.abs(ref[ok]).max())


def test_limgrad_device(dev):
    from oracle import sizing as osz
    from seismicmesh_amd.sizing import limgrad
    rng = np.random.default_rng(6)
    for shape in [(30, 41), (9, 11, 13)]:
        f0 = rng.uniform(75, 600, shape)
        got = limgrad(f0, 6.28, 0.15)
        dims = shape if len(shape) == 3 else (shape[0], shape[1], 1)
        ref = osz.limgrad(dims, 6.28, 0.15, 10000, f0.flatten("F")).reshape(shape, order="F")
        # exact fixed point on the device; the reference stops within ftol
        ftol = f0.min() * math.sqrt(1e-9)
        assert np.all(got <= ref + 1e-9)
        assert np.all(ref - got <= 2 * ftol * max(shape))
