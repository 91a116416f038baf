"""Pins of the oracle's sizing layer (CPU)."""
import itertools
import os

import numpy as np
import pytest

from oracle import sizing as osz


def _l1_transform(f0, step):
    """Brute force: min_j f0_j + step * |i - j|_1 over the whole grid."""
    idx = np.array(list(itertools.product(*[range(n) for n in f0.shape])))
    vals = f0.reshape(-1)
    out = np.empty(len(idx))
    for k, i in enumerate(idx):
        out[k] = np.min(vals + step * np.abs(idx - i).sum(1))
    return out.reshape(f0.shape)


@pytest.mark.parametrize("shape", [(5, 7), (4, 3, 6), (1, 9)])
def test_limgrad_fixed_point(shape):
    rng = np.random.default_rng(0)
    f0 = rng.uniform(50, 200, shape)
    dims = shape if len(shape) == 3 else (shape[0], shape[1], 1)
    got = osz.limgrad(dims, 6.0, 0.5, 10000, f0.flatten("F")).reshape(shape, order="F")
    want = _l1_transform(f0, 3.0)
    # the reference updates only when the excess exceeds ftol = min(f) * sqrt(1e-9)
    ftol = f0.min() * np.sqrt(1e-9)
    assert np.all(got >= want - 1e-9)
    assert np.all(got <= want + 2 * ftol * max(shape))


def test_limgrad_vs_reference(fasthj_ref):
    rng = np.random.default_rng(1)
    for shape in [(6, 9, 1), (5, 4, 3), (12, 10, 1)]:
        f0 = rng.uniform(75, 500, shape)
        ff = f0.flatten("F")
        ref = fasthj_ref.limgrad(np.array(shape, dtype=np.int32), 6.28, 0.15, 10000, ff).ravel()
        got = osz.limgrad(shape, 6.28, 0.15, 10000, ff)
        assert np.array_equal(got, ref)


def test_interpolant_known_values():
    # reference tests/test_2dmesher_vs_water.py:14-43: water (v = 0 -> 1500)
    # under a 1000 m/s layer, wl = 5, freq = 2: h = v / 10
    bbox = (-10000.0, 0.0, 0.0, 10000.0)
    vs = np.zeros((200, 200))
    vs[0:150, :] = 1000
    vs[vs < 1e-3] = 1500
    ef, cell = osz.sizing_from_velocity(vs, bbox, hmin=10, hmax=10e6, wl=5, freq=2.0)
    assert ef.eval((-5000, 5000)) == 100
    assert ef.eval((-1, 5000)) == 150


def test_interpolant_is_linear_with_extrapolation():
    bbox = (0.0, 1.0, 0.0, 2.0)
    z, x = osz.grid_vectors(bbox, (11, 21))
    f = 3.0 * z[:, None].astype(float) - 2.0 * x[None, :].astype(float) + 1.0
    rgi = osz.build_interpolant(f, bbox)
    q = np.array([[0.33, 1.1], [-0.5, 0.3], [1.7, 2.9]])
    assert np.allclose(rgi(q), 3 * q[:, 0] - 2 * q[:, 1] + 1, atol=1e-6)


_GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_oracle_segy_domain_extension_known_answer():
    # reference tests/test_2dmesher_domain_extension.py:15-60 asserts
    # (9428, 18525) for the constant pad with atol=100: SEG-Y decoding,
    # gradient + wavelength sizing, gradation, padding and the whole loop
    import contextlib
    import io
    from oracle import distmesh as odm, geometry as og, sizing as osz
    bbox = (-10e3, 0.0, 0.0, 10e3)
    ef, _ = osz.sizing_from_segy(os.path.join(_GOLDEN, "testing.segy"), bbox, hmin=100,
                                 hmax=10e6, wl=5, freq=5.0, grad=50.0, grade=0.005,
                                 stencil_size=100, domain_pad_m=1e3, pad_style="constant")
    with contextlib.redirect_stdout(io.StringIO()):
        p, t = odm.generate_mesh(og.Rectangle(bbox), ef, h0=100, verbose=0)
    assert abs(len(p) - 9428) <= 100 and abs(len(t) - 18525) <= 100


def test_product_readers_match_oracle():
    from oracle import sizing as osz
    from seismicmesh_amd import sizing as S
    fn = os.path.join(_GOLDEN, "testing.segy")
    a = osz.read_segy(fn)
    b = S.read_velocity_model(fn)
    assert np.array_equal(a[0], b[0]) and a[1:] == b[1:]
    vp = a[0]
    assert np.array_equal(osz.gradient_sizing(vp, 50.0, 100), S._gradient_sizing(vp, 50.0, 100))
    cell = vp / 25.0
    for style in ("edge", "constant", "linear_ramp"):
        c1, bb1 = osz.domain_pad(cell, (-10e3, 0.0, 0.0, 10e3), 1e3, style)
        c2, _, bb2 = S._build_domain_pad(cell, vp, (-10e3, 0.0, 0.0, 10e3),
                                         {"domain_pad": 1e3, "pad_style": style})
        assert np.array_equal(c1, c2) and bb1 == bb2
