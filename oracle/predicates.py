"""Exact geometric predicates in rational arithmetic (test infrastructure).

Used by tests as an independent, exact checker of Delaunay connectivity on
small inputs: every float64 is converted to a Fraction, so signs are exact.
"""
from fractions import Fraction

import numpy as np


def _F(x):
    return Fraction(float(x))


def orient2d(a, b, c):
    """Sign of det[[bx-ax, by-ay], [cx-ax, cy-ay]] (> 0: counter-clockwise)."""
    ax, ay = _F(a[0]), _F(a[1])
    d = (_F(b[0]) - ax) * (_F(c[1]) - ay) - (_F(b[1]) - ay) * (_F(c[0]) - ax)
    return (d > 0) - (d < 0)


def incircle(a, b, c, d):
    """> 0 when d lies strictly inside the circle through the CCW triangle abc."""
    dx, dy = _F(d[0]), _F(d[1])
    rows = []
    for p in (a, b, c):
        x = _F(p[0]) - dx
        y = _F(p[1]) - dy
        rows.append((x, y, x * x + y * y))
    (a0, a1, a2), (b0, b1, b2), (c0, c1, c2) = rows
    det = (a0 * (b1 * c2 - b2 * c1) - a1 * (b0 * c2 - b2 * c0) + a2 * (b0 * c1 - b1 * c0))
    return (det > 0) - (det < 0)


def orient3d(a, b, c, d):
    """Sign of the volume det of (b-a, c-a, d-a) (> 0: positive tet)."""
    A = [[_F(p[k]) - _F(a[k]) for k in range(3)] for p in (b, c, d)]
    det = (A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1])
           - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0])
           + A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]))
    return (det > 0) - (det < 0)


def insphere(a, b, c, d, e):
    """> 0 when e lies strictly inside the sphere through the positive tet abcd."""
    rows = []
    for p in (a, b, c, d):
        v = [_F(p[k]) - _F(e[k]) for k in range(3)]
        rows.append(v + [v[0] * v[0] + v[1] * v[1] + v[2] * v[2]])
    M = rows

    def det3(m):
        return (m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1])
                - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0])
                + m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]))

    det = Fraction(0)
    for j in range(4):
        minor = [[M[i][k] for k in range(4) if k != j] for i in range(1, 4)]
        det += (-1) ** j * M[0][j] * det3(minor)
    # expanding along the first row of the 4x4 lifted matrix gives the
    # negative of the standard insphere determinant orientation
    det = -det
    return (det > 0) - (det < 0)


def is_delaunay_2d(p, t):
    """Brute force: every cell CCW and no point strictly inside any circumcircle."""
    p = np.asarray(p, dtype=float)
    for tri in np.asarray(t):
        a, b, c = p[tri[0]], p[tri[1]], p[tri[2]]
        if orient2d(a, b, c) <= 0:
            return False
        for k in range(len(p)):
            if k in tri:
                continue
            if incircle(a, b, c, p[k]) > 0:
                return False
    return True
