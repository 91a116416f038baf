"""Velocity-model readers and mesh linting on the CPU.

Fixtures: tests/golden/testing.segy and tests/golden/test3D.bin are the
reference test suite's own data files (tests/testing.segy, tests/test3D.bin),
copied unchanged."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_ibm_float_known_values():
    from seismicmesh_amd.sizing import _ibm_to_float32
    # textbook IBM hexadecimal floats: 0xC276A000 = -118.625,
    # 0x42640000 = 100.0, 0x41100000 = 1.0, 0x3F800000 = 0.03125
    w = np.array([0xC276A000, 0x42640000, 0x41100000, 0x3F800000, 0x0], dtype=np.uint32)
    assert _ibm_to_float32(w).tolist() == [-118.625, 100.0, 1.0, 0.03125, 0.0]


def test_read_segy_fixture():
    from seismicmesh_amd import read_velocity_model
    vp, nz, nx, ny = read_velocity_model(os.path.join(GOLDEN, "testing.segy"))
    assert (nz, nx, ny) == (10, 10, 0) and vp.shape == (10, 10)
    # IBM-float traces 1500 ... 7000 m/s increasing with the sample index;
    # the grid is flipped so the last sample is row 0
    assert vp.min() == 1500.0 and vp.max() == 7000.0
    assert (vp[0] == 7000.0).all() and (vp[-1] == 1500.0).all()
    assert (np.diff(vp[::-1, 0]) > 0).all()
    # every column (trace) is the same profile
    assert (vp == vp[:, :1]).all()


def test_read_binary_fixture_axes_order():
    from seismicmesh_amd import read_velocity_model
    fn = os.path.join(GOLDEN, "test3D.bin")
    vp, nz, nx, ny = read_velocity_model(fn, nz=20, nx=10, ny=10, byte_order="little",
                                         axes_order=(2, 0, 1), axes_order_sort="F",
                                         dtype="float32")
    assert vp.shape == (20, 10, 10)
    raw = np.fromfile(fn, dtype="<f4")
    # axes_order (2, 0, 1): the file's axes are (nx, ny, nz) in Fortran order
    direct = raw.reshape(10, 10, 20, order="F").transpose(2, 0, 1)[::-1]
    assert np.array_equal(vp, direct)
    with pytest.raises(ValueError):
        read_velocity_model(fn, nz=20, nx=10, ny=None, byte_order="little",
                            axes_order=(2, 0, 1), axes_order_sort="F", dtype="float32")
    with pytest.raises(ValueError):
        read_velocity_model(fn, nz=20, nx=10, ny=10, byte_order="middle",
                            axes_order=(2, 0, 1), axes_order_sort="F", dtype="float32")


def test_domain_pad_shapes_and_styles():
    from seismicmesh_amd.sizing import _build_domain_pad
    vp = np.arange(12.0).reshape(3, 4) + 1500
    cell = vp / 10
    bbox = (-300.0, 0.0, 0.0, 400.0)
    for style in ("edge", "constant", "linear_ramp"):
        c2, v2, bb = _build_domain_pad(cell, vp, bbox, {"domain_pad": 200.0, "pad_style": style})
        # dz = dx = 100: two cells at the bottom (z min) and on both sides of x
        assert c2.shape == (5, 8) and v2.shape == (5, 8)
        assert bb == (-500.0, 0.0, -200.0, 600.0)
        assert np.array_equal(c2[2:, 2:6], cell)
    with pytest.raises(ValueError):
        _build_domain_pad(cell, vp, bbox, {"domain_pad": 200.0, "pad_style": "mirror"})


def test_linter_finds_overlaps():
    from seismicmesh_amd import mesh_utils as mu
    p = np.array([[0, 0], [1, 0], [0, 1], [1, 1], [0.5, 0.5]], float)
    t = np.array([[0, 1, 4], [1, 3, 4], [3, 2, 4], [2, 0, 4]])
    assert mu.do_any_overlap(p, t) == []
    # a folded triangle whose centroid lies in its neighbour
    t2 = np.vstack((t, [[0, 1, 3]]))
    inter = mu.do_any_overlap(p, t2)
    assert (4, 1) in inter or (1, 4) in inter
    q, s = mu.linter(p, t2)
    assert len(s) < len(t2)
