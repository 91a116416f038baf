"""Geometry primitives and CSG with the reference's API, evaluated on the GPU.

Mirrors SeismicMesh/geometry/signed_distance_functions.py (SeismicMesh
3.6.2): the same classes, constructor arguments, ``bbox`` / ``corners`` /
``dim`` attributes and transform semantics (translate, rotate, stretch;
:89-280).  Each object lowers itself to an SDF program (``tape()``) that the
device kernels interpret in FP64 (csrc/sdf_eval.h); ``eval(x)`` runs that
program on the GPU.  There is no CPU evaluation path.
"""
import ctypes
import itertools
from functools import reduce

import numpy as np

from . import _lib
from ._lib import Node

__all__ = [
    "Disk", "Ball", "Rectangle", "Cube", "Torus", "Prism", "Cylinder", "Union",
    "Intersection", "Difference", "Repeat", "corners", "eval_tape",
]


def corners(bbox):
    """Corners of an N-dim box (signed_distance_functions.py:74-78)."""
    mins = bbox[::2]
    maxs = bbox[1::2]
    return np.array(list(itertools.product(*zip(mins, maxs))))


def _gather_corners(domains):
    cs = [d.corners for d in domains if d.corners is not None]
    return None if len(cs) == 0 else np.concatenate(cs)


# --------------------------------------------------------------- transforms
class _Transformable:
    """Rotation / stretch / translation set-up of a primitive
    (signed_distance_functions.py:104-280).  The inverse maps are stored in
    the node's transform block and applied on the device in the reference's
    order: translate, Rx, Ry, Rz, stretch."""

    def _configure(self, rotate, stretch, translate):
        r = list(rotate) if rotate is not None else [0.0, 0.0, 0.0]
        while len(r) < 3:
            r.append(0.0)
        self.rotation = r
        self.v = None if stretch is None else np.array(stretch, dtype=float)
        self.translation_vec = None if translate is None else np.array(translate, dtype=float)
        self._build_stretch()
        self._build_rotation()
        self._build_translation()

    def _build_stretch(self):
        if self.v is None:
            return
        assert len(self.v) == self.dim, "Length of stretch vector does not equal dimension"
        self.alpha = np.sqrt(np.dot(self.v, self.v))
        self.v = self.v / self.alpha
        c = corners(self.bbox)
        vx = np.multiply.outer(np.dot(self.v, c.T), self.v)
        tmp = (vx * self.alpha + (c - vx)).T
        self.bbox = tuple(v for k in range(self.dim) for v in (np.min(tmp[k]), np.max(tmp[k])))
        self.corners = corners(self.bbox)

    def _build_rotation(self):
        r = self.rotation
        if self.dim == 2:
            self.R = np.array([[np.cos(r[0]), -np.sin(r[0])], [np.sin(r[0]), np.cos(r[0])]])
            self.Rx_inv = self.R.T
            if r[0] != 0.0:
                c = corners(self.bbox)
                tmp = np.dot(self.R, c.T).T
                self.bbox = (np.min(tmp[:, 0]), np.max(tmp[:, 0]),
                             np.min(tmp[:, 1]), np.max(tmp[:, 1]))
                self.corners = corners(self.bbox)
            return
        Rx = np.array([[1, 0, 0], [0, np.cos(r[0]), -np.sin(r[0])], [0, np.sin(r[0]), np.cos(r[0])]])
        Ry = np.array([[np.cos(r[1]), 0, np.sin(r[1])], [0, 1, 0], [-np.sin(r[1]), 0, np.cos(r[1])]])
        Rz = np.array([[np.cos(r[2]), -np.sin(r[2]), 0], [np.sin(r[2]), np.cos(r[2]), 0], [0, 0, 1]])
        self.Rx, self.Ry, self.Rz = Rx, Ry, Rz
        self.Rx_inv, self.Ry_inv, self.Rz_inv = Rx.T, Ry.T, Rz.T
        if reduce(lambda a, b: a or b, r):
            boxes = []
            for k, R in enumerate((Rx, Ry, Rz)):
                if r[k] != 0.0:
                    t = np.dot(R, corners(self.bbox).T).T
                    boxes.append(tuple(v for j in range(3) for v in (np.min(t[:, j]), np.max(t[:, j]))))
                else:
                    boxes.append(self.bbox)
            self.bbox = tuple(min(b[i] for b in boxes) if i % 2 == 0 else max(b[i] for b in boxes)
                              for i in range(6))
            self.corners = corners(self.bbox)

    def _build_translation(self):
        v = self.translation_vec
        if v is None:
            return
        assert len(v) == self.dim, "Length of translation vector does not equal dimension"
        self.bbox = tuple(b + v[i // 2] for i, b in enumerate(self.bbox))

    def _node(self, op, prm):
        nd = Node()
        nd.op = op
        nd.dim = self.dim
        for i, v in enumerate(prm):
            nd.prm[i] = float(v)
        flags = 0
        if self.translation_vec is not None:
            flags |= _lib.XF_TRANSLATE
            for i, v in enumerate(self.translation_vec):
                nd.xf[i] = float(v)
        if self.dim == 2:
            if self.rotation[0] != 0.0:
                flags |= _lib.XF_ROT_X
                for i, v in enumerate(np.asarray(self.Rx_inv).ravel()):
                    nd.xf[3 + i] = float(v)
        else:
            for k, (bit, R) in enumerate(((_lib.XF_ROT_X, self.Rx_inv), (_lib.XF_ROT_Y, self.Ry_inv),
                                          (_lib.XF_ROT_Z, self.Rz_inv))):
                if self.rotation[k] != 0.0:
                    flags |= bit
                    for i, v in enumerate(np.asarray(R).ravel()):
                        nd.xf[3 + 9 * k + i] = float(v)
        if self.v is not None:
            flags |= _lib.XF_STRETCH
            for i, v in enumerate(self.v):
                nd.xf[30 + i] = float(v)
            nd.xf[33] = float(self.alpha)
        nd.flags = flags
        return nd


class _Evaluable:
    def eval(self, x):
        """Signed distance at points x (n, dim), FP64, evaluated on the GPU."""
        return eval_tape(self.tape(), self.dim, x)

    def show(self, filename=None, samples=10000):
        raise NotImplementedError("plotting is not part of this package")


# --------------------------------------------------------------- primitives
class Disk(_Transformable, _Evaluable):
    def __init__(self, x0, r, rotate=(0, 0, 0), stretch=None, translate=None):
        self.dim = 2
        self.xc, self.yc, self.r = x0[0], x0[1], r
        self.bbox = (x0[0] - r, x0[0] + r, x0[1] - r, x0[1] + r)
        self._configure(rotate, stretch, translate)
        self.corners = None

    def tape(self):
        return [self._node(_lib.OP_DISK, (self.xc, self.yc, self.r))]


class Ball(_Transformable, _Evaluable):
    def __init__(self, x0, r, rotate=(0, 0, 0), stretch=None, translate=None):
        if stretch is not None:
            assert len(stretch) == 3
        self.dim = 3
        self.xc, self.yc, self.zc, self.r = x0[0], x0[1], x0[2], r
        self.bbox = (x0[0] - r, x0[0] + r, x0[1] - r, x0[1] + r, x0[2] - r, x0[2] + r)
        self._configure(rotate, stretch, translate)
        self.corners = None

    def tape(self):
        return [self._node(_lib.OP_BALL, (self.xc, self.yc, self.zc, self.r))]


class Rectangle(_Transformable, _Evaluable):
    def __init__(self, bbox, rotate=(0.0, 0.0, 0.0), stretch=None, translate=None):
        self.dim = 2
        self.corners = corners(bbox)
        self.bbox0 = bbox
        self.bbox = bbox
        self._configure(rotate, stretch, translate)

    def tape(self):
        return [self._node(_lib.OP_RECT, self.bbox0[:4])]


class Cube(_Transformable, _Evaluable):
    def __init__(self, bbox, rotate=(0, 0, 0), stretch=None, translate=None):
        self.dim = 3
        self.corners = corners(bbox)
        self.bbox0 = bbox
        self.bbox = bbox
        self._configure(rotate, stretch, translate)

    def tape(self):
        return [self._node(_lib.OP_CUBE, self.bbox0[:6])]


class Torus(_Transformable, _Evaluable):
    def __init__(self, r1, r2, rotate=(0, 0, 0), stretch=None, translate=None):
        assert r1 > 0.0 and r2 > 0.0
        z = 2 * max(r1, r2)
        self.dim = 3
        self.bbox = (-2 * z, 2 * z, -2 * z, 2 * z, -2 * z, 2 * z)
        self.t = (r1, r2)
        self._configure(rotate, stretch, translate)
        self.corners = None

    def tape(self):
        return [self._node(_lib.OP_TORUS, self.t)]


class Prism(_Transformable, _Evaluable):
    def __init__(self, b, h, rotate=(0, 0, 0), stretch=None, translate=None):
        self.bbox = (-b, +b, -b, +b, -h, +h)
        self.h = (b, h)
        self.dim = 3
        self._configure(rotate, stretch, translate)
        self.corners = None

    def tape(self):
        return [self._node(_lib.OP_PRISM, self.h)]


class Cylinder(_Transformable, _Evaluable):
    def __init__(self, h=1.0, r=0.5, rotate=(0, 0, 0), stretch=None, translate=None):
        assert h > 0.0 and r > 0.0
        h /= 2.0
        sz = max(h, r)
        self.dim = 3
        self.bbox = (-2 * sz, 2 * sz, -2 * sz, 2 * sz, -2 * sz, 2 * sz)
        self.h = (r, h)
        self._configure(rotate, stretch, translate)
        self.corners = None

    def tape(self):
        return [self._node(_lib.OP_CYLINDER, self.h)]


# --------------------------------------------------------------- CSG
class _CSG(_Evaluable):
    _op = None

    def __init__(self, domains, smoothness=0.0):
        dims = [d.dim for d in domains]
        assert all(d == dims[0] for d in dims), "mixed 2D/3D domains"
        self.dim = dims[0]
        self.k = smoothness
        b = [d.bbox for d in domains]
        self.bbox = tuple(min(x[i] for x in b) if i % 2 == 0 else max(x[i] for x in b)
                          for i in range(2 * self.dim))
        self.corners = _gather_corners(domains)
        self.domains = domains

    def _fold_node(self, op):
        nd = Node()
        nd.op = op
        nd.nargs = 2
        nd.dim = self.dim
        nd.prm[0] = float(self.k)
        return nd

    def tape(self):
        # left fold in the reference's order, two operands at a time
        out = list(self.domains[0].tape())
        for d in self.domains[1:]:
            out += d.tape()
            out.append(self._fold_node(self._op))
        return out


class Union(_CSG):
    """min of the children; smooth: signed_distance_functions.py:332-334"""
    _op = _lib.OP_UNION


class Intersection(_CSG):
    """max of the children; smooth: signed_distance_functions.py:372-374"""
    _op = _lib.OP_INTERSECT


class Difference(_CSG):
    """max(d0, -d1, ...); smooth: a left fold over the reversed children
    (signed_distance_functions.py:412-423)"""
    _op = _lib.OP_DIFFERENCE

    def tape(self):
        if self.k == 0.0:
            return super().tape()
        ds = self.domains[::-1]
        out = list(ds[0].tape())
        for d in ds[1:]:
            out += d.tape()
            out.append(self._fold_node(_lib.OP_SDIFF_STEP))
        return out


class Repeat(_Evaluable):
    """Periodic repetition clipped to a box (signed_distance_functions.py:283-297)."""

    def __init__(self, bbox, domain, period):
        self.bbox = bbox
        self.domain = domain
        self.corners = None
        self.period = np.array(period, dtype=float)
        self.parent = Cube(bbox)
        self.dim = 3

    def tape(self):
        push = Node()
        push.op = _lib.OP_PUSH_MOD
        push.dim = self.dim
        for i, v in enumerate(self.period):
            push.prm[i] = float(v)
        pop = Node()
        pop.op = _lib.OP_POP_POINT
        mx = Node()
        mx.op = _lib.OP_INTERSECT
        mx.nargs = 2
        mx.dim = self.dim
        return [push] + list(self.domain.tape()) + [pop] + self.parent.tape() + [mx]


# --------------------------------------------------------------- evaluation
def tape_array(tape):
    depth, points = validate_tape(tape)
    arr = (Node * len(tape))()
    for i, nd in enumerate(tape):
        arr[i] = nd
    # programs deeper than the register interpreter's 8 values (or with
    # nested Repeat points) are flagged for the scratch-stack interpreter
    arr[0].pad[0] = 1.0 if (depth > 8 or points > 1) else 0.0
    return arr


def eval_tape(tape, dim, x):
    """Evaluate an SDF program at points x on the GPU (FP64)."""
    import torch

    _lib.require_gpu()
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1, dim))
    n = len(x)
    if n == 0:
        return np.zeros(0)
    dev = torch.device("cuda", torch.cuda.current_device())
    xt = torch.from_numpy(x).to(dev)
    dt = torch.empty(n, dtype=torch.float64, device=dev)
    arr = tape_array(tape)
    tt = torch.frombuffer(bytearray(bytes(arr)), dtype=torch.uint8).to(dev)
    stream = torch.cuda.current_stream().cuda_stream
    _lib.check(_lib.lib.sm_sdf_eval(ctypes.c_void_p(tt.data_ptr()), len(tape), dim,
                                    ctypes.c_void_p(xt.data_ptr()), n,
                                    ctypes.c_void_p(dt.data_ptr()), ctypes.c_void_p(stream)),
               "sm_sdf_eval")
    return dt.cpu().numpy()


def validate_tape(tape):
    """Check an SDF program's stack discipline before it reaches the GPU:
    at most 24 live values and 3 pushed points, exactly one value left."""
    sp, pp = 0, 0
    max_sp, max_pp = 0, 0
    for nd in tape:
        op = nd.op
        if op < _lib.OP_UNION:
            sp += 1
        elif op <= _lib.OP_DIFFERENCE:
            if nd.nargs < 1 or nd.nargs > sp:
                raise ValueError("malformed SDF program (CSG arity)")
            sp -= nd.nargs - 1
        elif op == _lib.OP_SDIFF_STEP:
            if sp < 2:
                raise ValueError("malformed SDF program (smooth difference)")
            sp -= 1
        elif op == _lib.OP_PUSH_MOD:
            pp += 1
        elif op == _lib.OP_POP_POINT:
            pp -= 1
        else:
            raise ValueError("unknown SDF opcode %d" % op)
        if sp > 24 or pp > 3 or pp < 0:
            raise ValueError("SDF program too deep for the device evaluator")
        max_sp, max_pp = max(max_sp, sp), max(max_pp, pp)
    if sp != 1 or pp != 0:
        raise ValueError("malformed SDF program")
    return max_sp, max_pp
