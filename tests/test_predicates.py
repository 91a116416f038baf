"""The filtered + exact geometric predicates (seismicmesh_amd/csrc/predicates.h)
against exact rational arithmetic (oracle/predicates.py), on the CPU.

The header is compiled into a small test library with g++.  Inputs cover
general position, exactly degenerate configurations (cocircular and
cospherical lattice points, collinear and coplanar points), one-ulp
perturbations of those, and far-from-origin coordinates whose differences
are not exact (so the exact stage needs the rounding tails)."""
import ctypes
import os
import subprocess
from fractions import Fraction

import numpy as np
import pytest

from oracle import predicates as P

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "seismicmesh_amd", "csrc")


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("pred") / "libpredh.so")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-shared", "-fPIC",
                           "-I", CSRC, "-I", os.path.join(HERE, "..", "include"),
                           os.path.join(HERE, "native", "pred_harness.cpp"), "-o", out])
    L = ctypes.CDLL(out)
    for name in ("h_orient2d_exact", "h_incircle_exact96", "h_incircle_h", "h_orient3d_h",
                 "h_insphere_h", "h_insphere_exact160", "h_orient2d_if_exact",
                 "h_incircle_if_exact", "h_orient2d_fix", "h_incircle_fix", "h_orient3d_fix",
                 "h_insphere_fix"):
        getattr(L, name).restype = ctypes.c_int
    return L


def _p(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(ctypes.c_void_p)


UNRES = 0x7FFFFFFF


def _sign(x):
    return (x > 0) - (x < 0)


def _cases2d(rng):
    out = []
    for _ in range(200):
        out.append(rng.random((4, 2)))
    # cocircular: lattice squares and rectangles, scaled and shifted
    for h in (1.0, 0.1, 75.0, 1.0 / 3.0):
        for off in (0.0, 0.37, 12345.678):
            for _ in range(20):
                i, j = rng.integers(-50, 50, 2)
                w, v = rng.integers(1, 4, 2)
                sq = np.array([[i, j], [i + w, j], [i + w, j + v], [i, j + v]], float) * h + off
                out.append(sq[rng.permutation(4)])
    # one-ulp perturbations of cocircular quadruples
    for _ in range(200):
        i, j = rng.integers(-50, 50, 2)
        sq = np.array([[i, j], [i + 1, j], [i + 1, j + 1], [i, j + 1]], float) * 0.1 + 7.3
        k, c = rng.integers(0, 4), rng.integers(0, 2)
        sq[k, c] = np.nextafter(sq[k, c], np.inf if rng.random() < 0.5 else -np.inf)
        out.append(sq)
    # far from the origin: differences carry rounding tails
    for _ in range(200):
        base = rng.random(2) * 1e6
        q = base + rng.random((4, 2)) * 1e-6
        out.append(q)
    return out


def test_orient2d_and_incircle_exact(lib):
    rng = np.random.default_rng(0)
    for q in _cases2d(rng):
        a, b, c, d = q
        o = P.orient2d(a, b, c)
        assert lib.h_orient2d_exact(_p(a), _p(b), _p(c)) == o
        if o == 0:
            continue
        if o < 0:
            a, b = b, a
        want = P.incircle(a, b, c, d)
        got96 = lib.h_incircle_exact96(_p(a), _p(b), _p(c), _p(d))
        assert got96 in (want, UNRES)
        assert lib.h_incircle_h(_p(a), _p(b), _p(c), _p(d)) == want


def test_incircle_exact96_settles_exact_differences(lib):
    # nearby points have exact coordinate differences (Sterbenz): the
    # device-sized stage must settle every such case, degenerate or not
    rng = np.random.default_rng(1)
    n = 0
    for q in _cases2d(rng)[:560]:
        a, b, c, d = q
        o = P.orient2d(a, b, c)
        if o == 0:
            continue
        if o < 0:
            a, b = b, a
        exact_diffs = all(Fraction(float(x)) - Fraction(float(y)) == Fraction(float(x - y))
                          for x, y in ((a[0], d[0]), (a[1], d[1]), (b[0], d[0]), (b[1], d[1]),
                                       (c[0], d[0]), (c[1], d[1])))
        if exact_diffs:
            assert lib.h_incircle_exact96(_p(a), _p(b), _p(c), _p(d)) == P.incircle(a, b, c, d)
            n += 1
    assert n > 300


def test_orient3d_and_insphere(lib):
    rng = np.random.default_rng(2)
    cases = [rng.random((5, 3)) for _ in range(150)]
    for h in (1.0, 0.1, 1.0 / 3.0):
        for _ in range(40):
            i, j, k = rng.integers(-20, 20, 3)
            cube = np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)], float)
            pts = (cube + [i, j, k]) * h + 0.25
            cases.append(pts[rng.permutation(8)[:5]])  # cospherical (cube corners)
    for _ in range(100):
        base = rng.random(3) * 1e5
        cases.append(base + rng.random((5, 3)) * 1e-5)
    for q in cases:
        a, b, c, d, e = q
        o = P.orient3d(a, b, c, d)
        assert lib.h_orient3d_h(_p(a), _p(b), _p(c), _p(d)) == o
        if o == 0:
            continue
        if o < 0:
            b, c = c, b
        want = P.insphere(a, b, c, d, e)
        assert lib.h_insphere_h(_p(a), _p(b), _p(c), _p(d), _p(e)) == want
        got = lib.h_insphere_exact160(_p(a), _p(b), _p(c), _p(d), _p(e))
        assert got in (want, UNRES)


def test_register_stage_is_exact_when_it_answers(lib):
    # the expansion-free stage either answers exactly or defers; on lattice
    # points (multiples of 0.5, the staggered initial grid) it must answer
    rng = np.random.default_rng(3)
    answered = 0
    for q in _cases2d(rng):
        a, b, c, d = q
        o = lib.h_orient2d_if_exact(_p(a), _p(b), _p(c))
        assert o in (P.orient2d(a, b, c), UNRES)
        if P.orient2d(a, b, c) <= 0:
            continue
        s = lib.h_incircle_if_exact(_p(a), _p(b), _p(c), _p(d))
        assert s in (P.incircle(a, b, c, d), UNRES)
        answered += s != UNRES
    assert answered > 100
    for _ in range(300):
        i, j = rng.integers(-400, 400, 2)
        w, v = rng.integers(1, 3, 2)
        sq = (np.array([[i, j], [i + w, j], [i + w, j + v], [i, j + v]], float) * 75.0
              + [-12000.0 + 37.5 * (j % 2), 0.0])
        a, b, c, d = sq
        assert lib.h_orient2d_if_exact(_p(a), _p(b), _p(c)) == P.orient2d(a, b, c) == 1
        assert lib.h_incircle_if_exact(_p(a), _p(b), _p(c), _p(d)) == 0


def test_fixed_point_stage(lib):
    """The 128/256-bit integer stage answers exactly or defers.  Cases that
    a slip in it would get wrong: degenerate quads (sign 0) next to
    one-ulp perturbations (sign decided by the last bit), points far from
    the origin (differences of 50+ significant bits: the 256-bit sums and
    carries), mixed signs around the axes, subnormals and zeros, and inputs
    it must refuse (exponent spans too wide, inf/nan)."""
    rng = np.random.default_rng(4)
    cases = _cases2d(rng)
    for _ in range(300):
        base = rng.standard_normal(2) * 10.0 ** rng.integers(-3, 8)
        span = abs(base).max() * 10.0 ** -rng.integers(1, 15)
        cases.append(base + rng.standard_normal((4, 2)) * span)
    for _ in range(200):  # around the axes, zeros and subnormals included
        q = rng.standard_normal((4, 2)) * 1e-3
        q[rng.integers(0, 4), rng.integers(0, 2)] = 0.0
        if rng.random() < 0.3:
            q[rng.integers(0, 4), rng.integers(0, 2)] = 5e-324 * rng.integers(1, 1000)
        cases.append(q)
    for _ in range(200):  # cocircular on a scaled, shifted lattice, then nudged
        i, j = rng.integers(-1000, 1000, 2)
        w, v = rng.integers(1, 50, 2)
        sq = np.array([[i, j], [i + w, j], [i + w, j + v], [i, j + v]], float) * 0.75 + [
            -12000.0, 3000.0]
        if rng.random() < 0.5:
            k, c = rng.integers(0, 4), rng.integers(0, 2)
            sq[k, c] = np.nextafter(sq[k, c], np.inf if rng.random() < 0.5 else -np.inf)
        cases.append(sq[rng.permutation(4)])
    answered = {"o": 0, "i": 0, "i0": 0}
    for q in cases:
        a, b, c, d = q
        o = P.orient2d(a, b, c)
        got = lib.h_orient2d_fix(_p(a), _p(b), _p(c))
        assert got in (o, UNRES), (q, o, got)
        answered["o"] += got != UNRES
        if o == 0:
            continue
        if o < 0:
            a, b = b, a
        want = P.incircle(a, b, c, d)
        got = lib.h_incircle_fix(_p(a), _p(b), _p(c), _p(d))
        assert got in (want, UNRES), (q, want, got)
        answered["i"] += got != UNRES
        answered["i0"] += got == 0
        # the answer does not depend on which vertex of the triangle comes first
        got2 = lib.h_incircle_fix(_p(b), _p(c), _p(a), _p(d))
        assert got2 == got
    assert answered["o"] > 0.8 * len(cases) and answered["i"] > 0.7 * len(cases), answered
    assert answered["i0"] > 100, answered
    # refusals: a difference of 2^62 units or more, and non-finite input
    a, b, c, d = np.array([[1.0, 0.0], [0.0, 1.0], [-1.0, 0.0], [0.0, 1e-30]])
    assert lib.h_incircle_fix(_p(a), _p(b), _p(c), _p(d)) == UNRES
    assert lib.h_orient2d_fix(_p(a), _p(b), _p(d)) == UNRES
    assert lib.h_orient2d_fix(_p(a), _p(b), _p(np.array([np.inf, 0.0]))) == UNRES
    assert lib.h_incircle_fix(_p(a), _p(b), _p(c), _p(np.array([np.nan, 0.0]))) == UNRES


def test_fixed_point_stage_3d(lib):
    """The 256/384-bit integer stage of orient3d and insphere answers
    exactly or defers: cospherical cube corners of lattices (sign 0) at
    the spacings of the 3D workloads, one-ulp perturbations of them (sign
    decided by the last bit), points far from the origin, mixed signs
    around the axes with zeros, and refusals (exponent spans too wide,
    non-finite input).  A slip in a carry, a sign or a minor's order fails
    the zeros or the perturbed cases."""
    rng = np.random.default_rng(5)
    cube = np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)], float)
    cases = [rng.random((5, 3)) for _ in range(100)]
    for h in (1.0, 0.1, 0.015, 150.0, 1.0 / 3.0):
        for off in (0.0, 0.25, -4200.0):
            for _ in range(20):
                i, j, k = rng.integers(-60, 60, 3)
                pts = (cube + [i, j, k]) * h + off
                q = pts[rng.permutation(8)[:5]]
                cases.append(q.copy())
                k2, c2 = rng.integers(0, 5), rng.integers(0, 3)
                q[k2, c2] = np.nextafter(q[k2, c2], np.inf if rng.random() < 0.5 else -np.inf)
                cases.append(q)
    for _ in range(100):
        base = rng.random(3) * 1e5
        cases.append(base + rng.random((5, 3)) * 1e-5)
    for _ in range(100):
        q = rng.standard_normal((5, 3)) * 1e-2
        q[rng.integers(0, 5), rng.integers(0, 3)] = 0.0
        cases.append(q)
    answered = {"o": 0, "i": 0, "i0": 0}
    for q in cases:
        a, b, c, d, e = q
        o = P.orient3d(a, b, c, d)
        got = lib.h_orient3d_fix(_p(a), _p(b), _p(c), _p(d))
        assert got in (o, UNRES), (q, o, got)
        answered["o"] += got != UNRES
        if o == 0:
            continue
        if o < 0:
            b, c = c, b
        want = P.insphere(a, b, c, d, e)
        got = lib.h_insphere_fix(_p(a), _p(b), _p(c), _p(d), _p(e))
        assert got in (want, UNRES), (q, want, got)
        answered["i"] += got != UNRES
        answered["i0"] += got == 0
        # an even permutation of the cell leaves the answer
        got2 = lib.h_insphere_fix(_p(b), _p(c), _p(a), _p(d), _p(e))
        assert got2 == got
    assert answered["o"] > 0.8 * len(cases) and answered["i"] > 0.6 * len(cases), answered
    assert answered["i0"] > 50, answered
    a, b, c, d, e = np.array([[1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0], [0, 0, 0], [1e-30, 0, 0]])
    assert lib.h_insphere_fix(_p(a), _p(b), _p(c), _p(d), _p(e)) == UNRES
    assert lib.h_orient3d_fix(_p(a), _p(b), _p(c), _p(e)) == UNRES
    assert lib.h_orient3d_fix(_p(a), _p(b), _p(c), _p(np.array([np.inf, 0, 0]))) == UNRES
