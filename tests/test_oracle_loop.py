"""Pins of the oracle's force loop against the reference tests' known
answers (CPU)."""
import contextlib
import io

import numpy as np

from oracle import distmesh as odm
from oracle import geometry as og
from oracle import sizing as osz
from oracle.predicates import is_delaunay_2d


def test_verbose_byte_counts():
    # reference tests/test_verbose.py:13-22: stdout sizes for verbose 0, 1, 2
    # pin the message formats and the 49 iterations of max_iter=50
    square = og.Rectangle((0.0, 1.0, 0.0, 1.0))
    for verbosity, size in zip([0, 1, 2], [0, 192, 6014]):
        buf = io.StringIO()
        with contextlib.redirect_stdout(buf):
            odm.generate_mesh(square, 0.1, verbose=verbosity)
        assert len(buf.getvalue()) == size


def test_disk_initial_points():
    s = odm.setup(og.Disk([0, 0], 1), 0.02, dict(odm.GEN_DEFAULTS))
    assert len(s["p"]) == 7879


def test_sdf_disk_counts():
    # reference tests/test_2dmesher_SDF.py:11-29: (63, 93) +- 10, area ~ pi
    disk = og.Disk([0.0, 0.0], 1)
    p, t = odm.generate_mesh(disk, lambda x: 0.2 - disk.eval(x) * 0.15,
                             bbox=(-1.0, 1.0, -1.0, 1.0), h0=0.2, max_iter=100, verbose=0)
    assert np.allclose([len(p), len(t)], [63, 93], atol=10)
    assert np.allclose(np.sum(og.simp_vol(p, t)), 3.14, atol=0.2)


def test_vs_water_counts():
    # reference tests/test_2dmesher_vs_water.py:43-53: (5616, 10955) +- 100
    bbox = (-10000.0, 0.0, 0.0, 10000.0)
    vs = np.zeros((200, 200))
    vs[0:150, :] = 1000
    vs[vs < 1e-3] = 1500
    ef, _ = osz.sizing_from_velocity(vs, bbox, hmin=10, hmax=10e6, wl=5, freq=2.0)
    ef.hmin = None
    p, t = odm.generate_mesh(og.Rectangle(bbox), ef, h0=125, verbose=0)
    assert np.allclose([len(p), len(t)], [5616, 10955], atol=100)


def test_force_step_invariants():
    # forces are pairwise opposite: with no fixed points and no projection
    # the total force is zero; bars are only pushed apart (F >= 0)
    rng = np.random.default_rng(0)
    p = rng.uniform(-0.5, 0.5, (300, 2))
    t = odm.delaunay(p)
    Ftot, parts = odm.compute_forces(p, t, lambda x: np.full(len(x), 0.05), 0.05, 1.2,
                                     return_parts=True)
    assert np.all(parts["F"] >= 0)
    assert np.allclose(Ftot.sum(0), 0.0, atol=1e-12)
    # uniform sizing: L0 = h L0mult (sum L^2 / (E h^2))^(1/2) in closed form
    L = parts["L"]
    assert np.allclose(parts["L0"], 1.2 * np.sqrt(np.mean(L ** 2)))


def test_qhull_stand_in_is_delaunay():
    rng = np.random.default_rng(1)
    p = rng.uniform(0, 1, (40, 2))
    assert is_delaunay_2d(p, odm.delaunay(p))
