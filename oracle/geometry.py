"""Oracle restatement of the reference's geometry layer (FP64 numpy).

Test infrastructure only: imported by tests/, the smoke check and the bench's
CPU baseline, never by the product.

Cited reference files (relative to the reference root):
  SeismicMesh/geometry/signed_distance_functions.py (SDFs, CSG, transforms)
  SeismicMesh/geometry/cpp/fast_geometry.cpp         (box SDFs, unique edges,
                                                      dihedral angles,
                                                      circumsphere gradient)
  SeismicMesh/geometry/utils.py                      (simp_vol, simp_qual,
                                                      fix_mesh, boundary sets)
"""
import itertools
from functools import reduce

import numpy as np


# --------------------------------------------------------------------------
# helpers (signed_distance_functions.py:70-101)
# --------------------------------------------------------------------------
def _length(x):
    # signed_distance_functions.py:70-71
    return np.sum(np.abs(x) ** 2, axis=-1) ** (1.0 / 2)


def corners(bbox):
    # signed_distance_functions.py:74-78
    mins = bbox[::2]
    maxs = bbox[1::2]
    return np.array(list(itertools.product(*zip(mins, maxs))))


def _gather_corners(domains):
    # signed_distance_functions.py:81-86
    cs = [d.corners for d in domains if d.corners is not None]
    return None if len(cs) == 0 else np.concatenate(cs)


def _manipulate(obj, x):
    # signed_distance_functions.py:89-101: translate, Rx, Ry, Rz, stretch
    if obj.translation_vec is not None:
        x = x - obj.translation_vec
    if obj.rotation[0] != 0.0:
        x = np.dot(obj.Rx_inv, x.T).T
    if obj.dim == 3:
        if obj.rotation[1] != 0.0:
            x = np.dot(obj.Ry_inv, x.T).T
        if obj.rotation[2] != 0.0:
            x = np.dot(obj.Rz_inv, x.T).T
    if obj.v is not None:
        xt = x.T
        vx = np.multiply.outer(np.dot(obj.v, xt), obj.v)
        x = vx / obj.alpha + (x - vx)
    return x


def _build_stretch(obj):
    # signed_distance_functions.py:130-159
    if obj.v is not None:
        assert len(obj.v) == obj.dim
        obj.v = np.array(obj.v, dtype=float)
        obj.alpha = np.sqrt(np.dot(obj.v, obj.v))
        obj.v /= obj.alpha
        c = corners(obj.bbox)
        vx = np.multiply.outer(np.dot(obj.v, c.T), obj.v)
        tmp = (vx * obj.alpha + (c - vx)).T
        obj.bbox = tuple(
            v for k in range(obj.dim) for v in (np.min(tmp[k]), np.max(tmp[k]))
        )
        obj.corners = corners(obj.bbox)
    return obj


def _build_rotation(obj):
    # signed_distance_functions.py:162-254
    r = obj.rotation
    if obj.dim == 2:
        obj.R = np.array([[np.cos(r[0]), -np.sin(r[0])], [np.sin(r[0]), np.cos(r[0])]])
        obj.Rx_inv = obj.R.T
        if r[0] != 0.0:
            c = corners(obj.bbox)
            tmp = np.dot(obj.R, c.T).T
            obj.bbox = (
                np.min(tmp[:, 0]), np.max(tmp[:, 0]),
                np.min(tmp[:, 1]), np.max(tmp[:, 1]),
            )
            obj.corners = corners(obj.bbox)
    else:
        obj.Rx = np.array([[1, 0, 0], [0, np.cos(r[0]), -np.sin(r[0])],
                           [0, np.sin(r[0]), np.cos(r[0])]])
        obj.Ry = np.array([[np.cos(r[1]), 0, np.sin(r[1])], [0, 1, 0],
                           [-np.sin(r[1]), 0, np.cos(r[1])]])
        obj.Rz = np.array([[np.cos(r[2]), -np.sin(r[2]), 0],
                           [np.sin(r[2]), np.cos(r[2]), 0], [0, 0, 1]])
        obj.Rx_inv, obj.Ry_inv, obj.Rz_inv = obj.Rx.T, obj.Ry.T, obj.Rz.T
        if reduce(lambda a, b: a or b, r):
            boxes = []
            for k, R in enumerate((obj.Rx, obj.Ry, obj.Rz)):
                if r[k] != 0.0:
                    c = corners(obj.bbox)
                    t = np.dot(R, c.T).T
                    boxes.append(tuple(v for j in range(3)
                                       for v in (np.min(t[:, j]), np.max(t[:, j]))))
                else:
                    boxes.append(obj.bbox)
            obj.bbox = tuple(
                min(b[i] for b in boxes) if i % 2 == 0 else max(b[i] for b in boxes)
                for i in range(6)
            )
            obj.corners = corners(obj.bbox)
    return obj


def _build_translation(obj):
    # signed_distance_functions.py:257-280
    v = obj.translation_vec
    if v is not None:
        assert len(v) == obj.dim
        obj.bbox = tuple(b + v[i // 2] for i, b in enumerate(obj.bbox))
    return obj


def _configure(obj):
    # signed_distance_functions.py:104-108
    return _build_translation(_build_rotation(_build_stretch(obj)))


# --------------------------------------------------------------------------
# box SDFs (fast_geometry.cpp:104-134 and 165-185)
# --------------------------------------------------------------------------
def drectangle(p, x1, x2, y1, y2):
    s1 = -y1 + p[:, 1]
    s2 = y2 - p[:, 1]
    s3 = -x1 + p[:, 0]
    s4 = x2 - p[:, 0]
    return -np.minimum(np.minimum(np.minimum(s1, s2), s3), s4)


def dblock(p, x1, x2, y1, y2, z1, z2):
    s1 = -z1 + p[:, 2]
    s2 = z2 - p[:, 2]
    s3 = -y1 + p[:, 1]
    s4 = y2 - p[:, 1]
    s5 = -x1 + p[:, 0]
    s6 = x2 - p[:, 0]
    t = np.minimum(np.minimum(np.minimum(np.minimum(s1, s2), s3), s4), s5)
    return -np.minimum(t, s6)


# --------------------------------------------------------------------------
# primitives (signed_distance_functions.py:429-593)
# --------------------------------------------------------------------------
class _Prim:
    def _setup(self, rotate, stretch, translate):
        self.rotation = list(rotate) if rotate is not None else [0.0, 0.0, 0.0]
        self.v = None if stretch is None else np.array(stretch, dtype=float)
        self.translation_vec = None if translate is None else np.array(translate, dtype=float)
        _configure(self)


class Disk(_Prim):
    def __init__(self, x0, r, rotate=(0, 0, 0), stretch=None, translate=None):
        self.dim = 2
        self.xc, self.yc, self.r = x0[0], x0[1], r
        self.bbox = (x0[0] - r, x0[0] + r, x0[1] - r, x0[1] + r)
        self._setup(rotate, stretch, translate)
        self.corners = None

    def eval(self, x):
        x = _manipulate(self, x)
        return np.sqrt(((x - np.array([self.xc, self.yc])) ** 2).sum(-1)) - self.r


class Ball(_Prim):
    def __init__(self, x0, r, rotate=(0, 0, 0), stretch=None, translate=None):
        self.dim = 3
        self.xc, self.yc, self.zc, self.r = x0[0], x0[1], x0[2], r
        self.bbox = (x0[0] - r, x0[0] + r, x0[1] - r, x0[1] + r, x0[2] - r, x0[2] + r)
        self._setup(rotate, stretch, translate)
        self.corners = None

    def eval(self, x):
        x = _manipulate(self, x)
        return np.sqrt((x[:, 0] - self.xc) ** 2 + (x[:, 1] - self.yc) ** 2
                       + (x[:, 2] - self.zc) ** 2) - self.r


class Rectangle(_Prim):
    def __init__(self, bbox, rotate=(0.0, 0.0, 0.0), stretch=None, translate=None):
        self.dim = 2
        self.corners = corners(bbox)
        self.bbox0 = bbox
        self.bbox = bbox
        self._setup(rotate, stretch, translate)

    def eval(self, x):
        x = _manipulate(self, x)
        b = self.bbox0
        return drectangle(x, b[0], b[1], b[2], b[3])


class Cube(_Prim):
    def __init__(self, bbox, rotate=(0, 0, 0), stretch=None, translate=None):
        self.dim = 3
        self.corners = corners(bbox)
        self.bbox0 = bbox
        self.bbox = bbox
        self._setup(rotate, stretch, translate)

    def eval(self, x):
        x = _manipulate(self, x)
        b = self.bbox0
        return dblock(x, b[0], b[1], b[2], b[3], b[4], b[5])


class Torus(_Prim):
    def __init__(self, r1, r2, rotate=(0, 0, 0), stretch=None, translate=None):
        assert r1 > 0.0 and r2 > 0.0
        z = 2 * max(r1, r2)
        self.dim = 3
        self.bbox = (-2 * z, 2 * z, -2 * z, 2 * z, -2 * z, 2 * z)
        self.t = (r1, r2)
        self._setup(rotate, stretch, translate)
        self.corners = None

    def eval(self, x):
        x = _manipulate(self, x)
        xz = np.column_stack((x[:, 0], x[:, 2]))
        q = np.column_stack((_length(xz) - self.t[0], x[:, 1]))
        return _length(q) - self.t[1]


class Prism(_Prim):
    def __init__(self, b, h, rotate=(0, 0, 0), stretch=None, translate=None):
        self.bbox = (-b, +b, -b, +b, -h, +h)
        self.h = (b, h)
        self.dim = 3
        self._setup(rotate, stretch, translate)
        self.corners = None

    def eval(self, x):
        x = _manipulate(self, x)
        q = np.abs(x)
        return np.maximum(
            q[:, 2] - self.h[1],
            np.maximum(q[:, 0] * 0.866025 + x[:, 1] * 0.5, -x[:, 1]) - self.h[0] * 0.5,
        )


class Cylinder(_Prim):
    def __init__(self, h=1.0, r=0.5, rotate=(0, 0, 0), stretch=None, translate=None):
        assert h > 0.0 and r > 0.0
        h /= 2.0
        sz = max(h, r)
        self.dim = 3
        self.bbox = (-2 * sz, 2 * sz, -2 * sz, 2 * sz, -2 * sz, 2 * sz)
        self.h = (r, h)
        self._setup(rotate, stretch, translate)
        self.corners = None

    def eval(self, x):
        x = _manipulate(self, x)
        xz = np.column_stack((x[:, 0], x[:, 2]))
        lxz = np.column_stack((_length(xz), x[:, 1]))
        d = np.abs(lxz) - self.h
        return np.minimum(np.maximum(d[:, 0], d[:, 1]), 0.0) + _length(np.maximum(d, 0.0))


# --------------------------------------------------------------------------
# CSG (signed_distance_functions.py:283-426)
# --------------------------------------------------------------------------
def _loop_call(func, d):
    # signed_distance_functions.py:300-304
    tmp = d[0]
    for i in range(0, len(d) - 1):
        tmp = func(tmp, d[i + 1])
    return tmp


class _CSG:
    def __init__(self, domains, smoothness=0.0):
        self.dim = domains[0].dim
        self.k = smoothness
        b = [d.bbox for d in domains]
        self.bbox = tuple(
            min(x[i] for x in b) if i % 2 == 0 else max(x[i] for x in b)
            for i in range(2 * self.dim)
        )
        self.corners = _gather_corners(domains)
        self.domains = domains


class Union(_CSG):
    def _smooth(self, d1, d2):
        h = np.maximum(self.k - np.abs(d1 - d2), 0.0)
        return np.minimum(d1, d2) - np.divide(h * h * 0.25, self.k)

    def eval(self, x):
        d = [dm.eval(x) for dm in self.domains]
        if self.k == 0.0:
            return np.minimum.reduce(d)
        return _loop_call(self._smooth, d)


class Intersection(_CSG):
    def _smooth(self, d1, d2):
        h = np.maximum(self.k - np.abs(d1 - d2), 0.0)
        return np.maximum(d1, d2) + h * h * 0.25 / self.k

    def eval(self, x):
        d = [dm.eval(x) for dm in self.domains]
        if self.k == 0.0:
            return np.maximum.reduce(d)
        return _loop_call(self._smooth, d)


class Difference(_CSG):
    def _smooth(self, d1, d2):
        h = np.maximum(self.k - np.abs(-d1 - d2), 0.0)
        return np.maximum(-d1, d2) + np.divide(h * h * 0.25, self.k)

    def eval(self, x):
        if self.k == 0.0:
            return np.maximum.reduce(
                [-dm.eval(x) if n > 0 else dm.eval(x) for n, dm in enumerate(self.domains)]
            )
        d = [dm.eval(x) for dm in self.domains]
        return _loop_call(self._smooth, d[::-1])


class Repeat:
    # signed_distance_functions.py:283-297
    def __init__(self, bbox, domain, period):
        self.bbox = bbox
        self.domain = domain
        self.corners = None
        self.period = np.array(period)
        self.parent = Cube(bbox)
        self.dim = 3

    def eval(self, x):
        q = np.mod(x + 0.5 * self.period, self.period) - 0.5 * self.period
        return np.maximum(self.domain.eval(q), self.parent.eval(x))


# --------------------------------------------------------------------------
# native helpers (fast_geometry.cpp)
# --------------------------------------------------------------------------
def unique_edges(edges):
    """fast_geometry.cpp:49-77: (min,max) per pair, lexicographic sort, unique."""
    e = np.asarray(edges, dtype=np.int32).reshape(-1, 2)
    if len(e) == 0:
        return np.zeros((0, 2), dtype=np.int32)
    e = np.sort(e, axis=1)
    return np.unique(e, axis=0).astype(np.int32)


_DH_EDGES = ((2, 3), (1, 3), (1, 2), (0, 3), (0, 2), (0, 1))


def calc_dihedral_angles(p, t):
    """fast_geometry.cpp:351-418: 6 dihedral angles per tet, (6C,1) radians."""
    p = np.asarray(p, dtype=float)
    t = np.asarray(t)
    out = np.empty((len(t), 6))
    for i in range(6):
        i0 = t[:, _DH_EDGES[i][0]]
        i1 = t[:, _DH_EDGES[i][1]]
        i2 = t[:, _DH_EDGES[5 - i][0]]
        i3 = t[:, _DH_EDGES[5 - i][1]]
        p0 = p[i0]
        v1 = p[i1] - p0
        v2 = p[i2] - p0
        v3 = p[i3] - p0
        v1 = v1 / np.sqrt(((v1 * v1)[:, 0] + (v1 * v1)[:, 1]) + (v1 * v1)[:, 2])[:, None]
        v2 = v2 / np.sqrt(((v2 * v2)[:, 0] + (v2 * v2)[:, 1]) + (v2 * v2)[:, 2])[:, None]
        v3 = v3 / np.sqrt(((v3 * v3)[:, 0] + (v3 * v3)[:, 1]) + (v3 * v3)[:, 2])[:, None]
        d23 = _dot3(v2, v3)
        d12 = _dot3(v1, v2)
        d13 = _dot3(v1, v3)
        c12 = _cross3(v1, v2)
        c13 = _cross3(v1, v3)
        n12 = np.sqrt(_dot3(c12, c12))
        n13 = np.sqrt(_dot3(c13, c13))
        with np.errstate(invalid="ignore", divide="ignore"):
            out[:, i] = np.arccos((d23 - d12 * d13) / (n12 * n13))
    return out.reshape(-1, 1)


def _dot3(a, b):
    return ((a[:, 0] * b[:, 0]) + (a[:, 1] * b[:, 1])) + (a[:, 2] * b[:, 2])


def _cross3(a, b):
    return np.column_stack((
        a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1],
        a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2],
        a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0],
    ))


def calc_circumsphere_grad(p0, p1, p2, p3):
    """fast_geometry.cpp:580-666: gradient of the circumradius w.r.t. p0."""
    p0, p1, p2, p3 = (np.asarray(a, dtype=float).reshape(-1, 3) for a in (p0, p1, p2, p3))
    x1, y1, z1 = (p0 - p3).T
    x2, y2, z2 = (p1 - p3).T
    x3, y3, z3 = (p2 - p3).T
    sq1 = x1 * x1 + y1 * y1 + z1 * z1
    sq2 = x2 * x2 + y2 * y2 + z2 * z2
    sq3 = x3 * x3 + y3 * y3 + z3 * z3
    da_dx = y2 * z3 - y3 * z2
    da_dy = z2 * x3 - x2 * z3
    da_dz = x2 * y3 - x3 * y2
    dDx_dx = -2.0 * x1 * da_dx
    dDx_dy = -2.0 * y1 * da_dx + sq2 * z3 - sq3 * z2
    dDx_dz = -2.0 * z1 * da_dx - sq2 * y3 + sq3 * y2
    dDy_dx = -2.0 * x1 * da_dy - sq2 * z3 + sq3 * z2
    dDy_dy = -2.0 * y1 * da_dy
    dDy_dz = -2.0 * z1 * da_dy + sq2 * x3 - sq3 * x2
    dDz_dx = -2.0 * x1 * da_dz + sq2 * y3 - sq3 * y2
    dDz_dy = -2.0 * y1 * da_dz - sq2 * x3 + sq3 * x2
    dDz_dz = -2.0 * z1 * da_dz
    a = x1 * da_dx + y1 * da_dy + z1 * da_dz
    Dx = -sq1 * da_dx + y1 * (sq2 * z3 - sq3 * z2) - z1 * (sq2 * y3 - sq3 * y2)
    Dy = -sq1 * da_dy - x1 * (sq2 * z3 - sq3 * z2) + z1 * (sq2 * x3 - sq3 * x2)
    Dz = -sq1 * da_dz + x1 * (sq2 * y3 - sq3 * y2) - y1 * (sq2 * x3 - sq3 * x2)
    s = Dx * Dx + Dy * Dy + Dz * Dz
    with np.errstate(invalid="ignore", divide="ignore"):
        gx = (Dx * dDx_dx + Dy * dDy_dx + Dz * dDz_dx) / (2.0 * a * a) - (da_dx * s) / (2.0 * a * a * a)
        gy = (Dx * dDx_dy + Dy * dDy_dy + Dz * dDz_dy) / (2.0 * a * a) - (da_dy * s) / (2.0 * a * a * a)
        gz = (Dx * dDx_dz + Dy * dDy_dz + Dz * dDz_dz) / (2.0 * a * a) - (da_dz * s) / (2.0 * a * a * a)
    return np.column_stack((gx, gy, gz))


# --------------------------------------------------------------------------
# mesh utilities (geometry/utils.py)
# --------------------------------------------------------------------------
def simp_vol(p, t):
    # geometry/utils.py:175-201
    dim = p.shape[1]
    if dim == 2:
        d01 = p[t[:, 1]] - p[t[:, 0]]
        d02 = p[t[:, 2]] - p[t[:, 0]]
        return (d01[:, 0] * d02[:, 1] - d01[:, 1] * d02[:, 0]) / 2
    d01 = p[t[:, 1]] - p[t[:, 0]]
    d02 = p[t[:, 2]] - p[t[:, 0]]
    d03 = p[t[:, 3]] - p[t[:, 0]]
    return np.einsum("ij,ij->i", np.cross(d01, d02), d03) / 6


def simp_qual(p, t):
    # geometry/utils.py:252-274 (2*inradius/circumradius of triangles)
    def length(a):
        return np.sqrt((a ** 2).sum(1))

    a = length(p[t[:, 1]] - p[t[:, 0]])
    b = length(p[t[:, 2]] - p[t[:, 0]])
    c = length(p[t[:, 2]] - p[t[:, 1]])
    r = 0.5 * np.sqrt((b + c - a) * (c + a - b) * (a + b - c) / (a + b + c))
    R = a * b * c / np.sqrt((a + b + c) * (b + c - a) * (c + a - b) * (a + b - c))
    return 2 * r / R


def unique_rows(A, return_index=False, return_inverse=False):
    # geometry/utils.py:130-172: rows compared as byte strings ('S' view)
    A = np.require(A, requirements="C")
    dt = np.dtype("S%d" % (A.dtype.itemsize * A.shape[1]))
    B, I, J = np.unique(A.view(dt), return_index=True, return_inverse=True)
    J = J.reshape(-1)
    B = B.view(A.dtype).reshape((-1, A.shape[1]))
    out = (B,)
    if return_index:
        out += (I,)
    if return_inverse:
        out += (J,)
    return out if len(out) > 1 else out[0]


def _unique_row_view(data):
    # geometry/utils.py:300-307
    b = np.ascontiguousarray(data).view(np.dtype((np.void, data.dtype.itemsize * data.shape[1])))
    u, cnts = np.unique(b, return_counts=True)
    return u.view(data.dtype).reshape(-1, data.shape[1]), cnts


def fix_mesh(p, t, ptol=2e-13, dim=2, delete_unused=False, fix_orientation=True):
    # geometry/utils.py:204-249
    snap = (p.max(0) - p.min(0)).max() * ptol
    _, ix, jx = unique_rows(np.round(p / snap) * snap, True, True)
    p = p[ix]
    t = jx[t]
    t = np.sort(t, axis=1)
    t, _ = _unique_row_view(t)
    if delete_unused:
        pix, _, jx = np.unique(t, return_index=True, return_inverse=True)
        t = np.reshape(jx, t.shape)
        p = p[pix, :]
    if fix_orientation:
        flip = simp_vol(p, t) < 0
        t[flip, :2] = t[flip, 1::-1]
    return p, t, jx


def get_edges(t, dim=2):
    # geometry/utils.py:277-297
    t = np.asarray(t)
    if dim == 2:
        return t[:, [[0, 1], [0, 2], [1, 2]]].reshape(-1, 2)
    return t[:, [[0, 1], [1, 2], [2, 0], [0, 3], [1, 3], [2, 3]]].reshape(-1, 2)


def get_facets(t):
    # geometry/utils.py:408-418
    ix = [[0, 1, 3], [1, 2, 3], [2, 0, 3], [1, 2, 0]]
    return np.array(t[:, ix]).reshape((len(t) * 4, 3))


def get_boundary_edges(t, dim=2):
    # geometry/utils.py:310-326
    e = np.sort(get_edges(t, dim=dim), axis=1)
    unq, cnt = _unique_row_view(e)
    return np.array([e for e, c in zip(unq, cnt) if c == (dim - 1)])


def get_boundary_facets(t):
    # geometry/utils.py:421-437
    f = np.sort(get_facets(t), axis=1)
    unq, cnt = _unique_row_view(f)
    return np.array([e for e, c in zip(unq, cnt) if c == 1])


def get_boundary_vertices(t, dim=2):
    # geometry/utils.py:364-382
    b = get_boundary_edges(t) if dim == 2 else get_boundary_facets(t)
    return np.unique(b.reshape(-1))


def vertex_to_entities(p, t, dim=2):
    # geometry/utils.py:97-128
    ne = len(t)
    ext = np.tile(np.arange(0, ne), (dim + 1, 1)).T.reshape((-1, 1), order="C")
    ve = np.hstack((np.reshape(t, (-1, 1)), ext))
    ve = ve[ve[:, 0].argsort(), :]
    idx = np.insert(np.diff(ve[:, 0]), 0, 0)
    ptr = np.argwhere(idx)
    ptr = np.insert(ptr, 0, 0)
    ptr = np.append(ptr, ne * (dim + 1))
    return ve[:, 1], ptr


def get_boundary_entities(p, t, dim=2):
    # geometry/utils.py:385-405
    bv = get_boundary_vertices(t, dim=dim)
    vtoe, ptr = vertex_to_entities(p, t, dim=dim)
    parts = [vtoe[ptr[v]:ptr[v + 1]] for v in bv]
    bele = np.concatenate(parts) if parts else np.array([], dtype=int)
    return np.unique(bele)


def delete_boundary_entities(p, t, dim=2, min_qual=0.10, verbose=1):
    # geometry/utils.py:440-469
    qual = simp_qual(p, t)
    bele = get_boundary_entities(p, t, dim=dim)
    delete = qual[bele] < min_qual
    if verbose:
        print("Deleting " + str(np.sum(delete)) + " poor quality boundary entities...", flush=True)
    delete = np.argwhere(delete == 1)
    t = np.delete(t, bele[delete], axis=0)
    p, t, _ = fix_mesh(p, t, delete_unused=True, dim=dim)
    return p, t


def laplacian2_fixed_point(p, t):
    """geometry/utils.py:494-547 with a direct sparse solve in place of pyamg.

    The reference solves the same linear system with pyamg's Ruge-Stuben
    solver (approximate, default tolerance); pyamg is absent here, so the
    oracle uses the exact solution.  Parity of this step is unpinned.
    """
    import scipy.sparse as sp
    from scipy.sparse.linalg import spsolve

    n = len(p)
    nds = t.T
    local_idx = np.array([[1, 2], [2, 0], [0, 1]]).T
    idx = nds[local_idx]
    row = np.array([idx[0], idx[1], idx[0], idx[1]]).ravel()
    col = np.array([idx[0], idx[1], idx[1], idx[0]]).ravel()
    a = np.ones(idx.shape[1:], dtype=int)
    val = np.array([+a, +a, -a, -a]).ravel()
    m = sp.coo_matrix((val.astype(float), (row, col)), shape=(n, n)).tocsr()
    bnd = get_boundary_vertices(t)
    for b in bnd:
        m.data[m.indptr[b]:m.indptr[b + 1]] = 0.0
    d = m.diagonal()
    d[bnd] = 1.0
    m.setdiag(d)
    rhs = np.zeros((n, 2))
    rhs[bnd] = p[bnd]
    m = m.tocsc()
    return np.column_stack([spsolve(m, rhs[:, 0]), spsolve(m, rhs[:, 1])]), t
