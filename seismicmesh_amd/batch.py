"""Many independent 2D domains meshed together on one GPU.

The reference meshes one domain per generate_mesh call
(SeismicMesh/generation/mesh_generator.py:291-529).  generate_mesh_batch
takes a list of domains, each with a uniform edge length, and runs all of
their force loops in the same kernel launches (one captured graph per
iteration): the points of domain d are a contiguous range of one point
array, each domain keeps its own triangulation (no triangle or neighbour
link crosses domains), its own L0 scale, centroid filter and projection
(csrc: sm_mesher_create_batch).  Set-up (initial points) and termination
are the reference's, per domain, on the host.  Forces are FP32.

Each domain is triangulated over the convex hull of its points, as the
reference's CGAL triangulation is, by default.  far_square=True
triangulates it inside a fixed square of four far vertices at 100 times its
extent instead: fewer flip rounds along straight, nearly collinear hull
edges, at the price of leaving out the hull triangles whose circumcircle
reaches a far vertex (flat triangles of nearly collinear hull points, which
the reference keeps when their centroid passes the filter).
"""
import ctypes

import numpy as np

from . import _lib
from . import generation as gen
from . import geometry as geo

__all__ = ["BatchMesher", "prepare_batch", "generate_mesh_batch"]


class BatchMesher(gen.Mesher):
    """A libsmgpu batch handle: `domains` (geometry objects), their uniform
    sizes `h`, fixed-point counts `nfix` and point counts `counts`."""

    def __init__(self, domains, h, nfix, counts, delta_t=0.30, geps_mult=0.1, device=0,
                 far_square=False):
        _lib.require_gpu()
        self.dim = 2
        nd = len(domains)
        self._tapes = [geo.tape_array(d.tape()) for d in domains]
        ptrs = (ctypes.c_void_p * nd)(*[ctypes.addressof(t) for t in self._tapes])
        lens = (ctypes.c_int * nd)(*[len(t) for t in self._tapes])
        self.hsize = np.ascontiguousarray(h, dtype=np.float64)
        self.nfix = np.ascontiguousarray(nfix, dtype=np.int32)
        self.voff = np.zeros(nd + 1, dtype=np.int64)
        self.voff[1:] = np.cumsum(counts)
        prm = _lib.MesherParams()
        prm.dim = 2
        prm.precision = 32
        prm.nlevels = 1
        prm.delta_t = delta_t
        prm.geps = geps_mult
        prm.deps = float(np.sqrt(np.finfo(np.double).eps))
        prm.L0mult = 1 + 0.4 / 2 ** (2 - 1)
        prm.far_square = int(bool(far_square))
        hd = ctypes.c_void_p()
        _lib.check(_lib.lib.sm_mesher_create_batch(
            ctypes.byref(prm), nd, ptrs, lens, self.hsize.ctypes.data, self.nfix.ctypes.data,
            self.voff.ctypes.data, device, ctypes.byref(hd)), "sm_mesher_create_batch")
        self.h = hd  # the handle, as in Mesher

    def split(self, p, t):
        """Per-domain (points, cells) from the batch's point array and kept
        cells (cells are grouped by the domain of their first vertex)."""
        out = []
        dom = np.searchsorted(self.voff, t[:, 0], side="right") - 1
        for d in range(len(self.voff) - 1):
            b, e = self.voff[d], self.voff[d + 1]
            out.append((p[b:e].copy(), t[dom == d] - b))
        return out


def prepare_batch(domains, edge_lengths, seed=0, **kwargs):
    """Initial points of every domain exactly as generate_mesh would make
    them (mesh_generator.py:425-458): fixed corners first, then the
    staggered lattice with the seeded rejection."""
    if len(domains) != len(edge_lengths):
        raise ValueError("one edge length per domain")
    sts = []
    for dom, h in zip(domains, edge_lengths):
        if not np.isscalar(h):
            raise ValueError("batched domains take a uniform (scalar) edge length")
        sts.append(gen.prepare(dom, float(h), verbose=0, seed=seed, precision=32, **kwargs))
    return sts


def generate_mesh_batch(domains, edge_lengths, max_iter=50, seed=0, verbose=0, far_square=False,
                        **kwargs):
    """Mesh every domain of `domains` with its uniform edge length; returns
    a list of (points, cells), one per domain, each as generate_mesh would
    return it (FP32 forces)."""
    sts = prepare_batch(domains, edge_lengths, seed=seed, max_iter=max_iter, **kwargs)
    m = make_batch_mesher(sts, far_square=far_square)
    try:
        m.set_points(np.vstack([st["p"] for st in sts]))
        m.run(max(max_iter - 1, 0))
        m.triangulate()
        parts = m.split(m.points(), m.cells())
    finally:
        m.close()
    out = []
    for st, (p, t) in zip(sts, parts):
        if verbose:
            print("Termination reached...maximum number of iterations reached.", flush=True)
        p, t = gen._termination(p, t, st["opts"], verbose=verbose)
        p = gen._improve_level_set_newton(p, t, st["dom"], st["deps"])
        t = gen._remove_triangles_outside(p, t, st["dom"], st["h0"] * 0.001)
        out.append((p, t))
    return out


def make_batch_mesher(sts, far_square=False):
    """A batch handle for prepared domains (initial points not set)."""
    return BatchMesher([st["dom"] for st in sts], [st["h0"] for st in sts],
                       [st["nfix"] for st in sts], [len(st["p"]) for st in sts],
                       delta_t=sts[0]["delta_t"],
                       geps_mult=sts[0]["opts"]["geps_mult"], far_square=far_square)
