"""sliver_removal with the reference's signature (3D).

Mirrors SeismicMesh/generation/mesh_generator.py:59-288 (serial path):
options (:107-120), domain and sizing unpacking (:122-166), the bounded
perturbation loop (:197-285) and termination (:655-677).

Each iteration re-tetrahedralises the points exactly on the host (the
reference moves the vertices of its CGAL triangulation, which yields the
same Delaunay tetrahedralisation for points in general position), filters
cells by their centroid's signed distance on the GPU, and selects and
perturbs slivers on the GPU (sm_mesher_find_slivers /
sm_mesher_perturb_slivers).  The step-size control runs on the host.
"""
import ctypes
import math
import time

import numpy as np

from . import _lib
from . import generation as gen
from . import geometry as geo
from . import mesh_utils as mu

__all__ = ["sliver_removal"]

_DEFAULTS = {
    "verbose": 1, "max_iter": 50, "perform_checks": False, "pfix": None, "axis": 1,
    "min_dh_angle_bound": 10.0, "max_dh_angle_bound": 180.0, "points": None,
    "delta_t": 0.30, "geps_mult": 0.1, "subdomains": [], "gamma": 1.0, "preserve": False,
    "precision": 64,
}


def _say(flag, msg):
    if flag:
        print(msg, flush=True)


def sliver_removal(points, domain, edge_length, comm=None, **kwargs):
    """Improve the minimum dihedral angle of a 3D mesh by perturbing the
    vertices of slivers (mesh_generator.py:59-288)."""
    opts = dict(_DEFAULTS)
    opts.update(kwargs)
    gen._parse_kwargs(kwargs)
    if opts["verbose"] < 0:
        raise ValueError("Unknown verbosity level")
    v1 = opts["verbose"] >= 1
    v2 = opts["verbose"] > 1

    points = np.asarray(points, dtype=np.float64)
    dim = points.shape[1]
    if dim == 2:
        raise Exception("Mesh improvement currently on works in 3D")

    dom, bbox0, _ = gen._unpack_domain(domain, opts)
    fh, bbox1, hmin, _ = gen._unpack_sizing(edge_length)
    bbox = bbox0 if bbox1 is None else gen._minmax(bbox0, bbox1)
    if bbox1 is not None and bbox0 != bbox1:
        dom = geo.Cube(bbox)
    if not isinstance(bbox, tuple):
        raise ValueError("`bbox` must be a tuple")

    h0 = hmin if hmin is not None else opts["h0"]
    if h0 < 0:
        raise ValueError("`h0` must be > 0")
    if opts["max_iter"] < 0:
        raise ValueError("`max_iter` must be > 0")
    max_iter = opts["max_iter"]
    _say(v1, f"Will attempt {max_iter} iterations to bound the dihedral angles...")

    geps = opts["geps_mult"] * h0
    deps = np.sqrt(np.finfo(np.double).eps) * h0
    min_dh = opts["min_dh_angle_bound"] * math.pi / 180
    max_dh = opts["max_dh_angle_bound"] * math.pi / 180
    _say(v1, f"Enforcing a min. dihedral bound of: {min_dh * 180 / math.pi} degrees...")
    _say(v1, f"Enforcing a max. dihedral bound of: {max_dh * 180 / math.pi} degrees...")

    p = points.copy()
    _say(v1, "Commencing sliver removal with %d vertices on rank %d." % (len(p), 0))

    # the handle's force parameters are unused here: only its tetrahedra,
    # centroid filter and sliver kernels run
    m = gen.Mesher([dom], gen._Sizing("const", value=h0), h0, opts["delta_t"], geps, deps, 0,
                   precision=64, dim=3)
    ns = gen.default_slabs(len(p), 3, opts.get("slabs"))
    if ns > 1:
        m.set_slabs(ns, axis=int(opts.get("axis", 1)))
    try:
        m.set_points(p)
        count = 0
        step = 0.10
        gamma = opts["gamma"]
        num_old_bad = np.inf
        while True:
            start = time.time()
            m.triangulate()
            nbad = ctypes.c_int64(0)
            _lib.check(_lib.lib.sm_mesher_find_slivers(m.h, min_dh, max_dh, ctypes.byref(nbad)),
                       "sm_mesher_find_slivers")
            num_bad = nbad.value

            if count == max_iter - 1:
                _say(v1, "FAILURE: Termination...maximum number of iterations reached. "
                         "Try increasing max_iter when generating the mesh")
                p, t = m.points(), m.cells().astype(np.int64)
                p, t = _termination(p, t, opts)
                break

            _say(v1, f"On rank: 0. There are {num_bad} slivers...")

            if num_bad == 0:
                _say(v1, f"Termination reached in {count} iterations...no slivers detected!")
                p, t = m.points(), m.cells().astype(np.int64)
                p, t, _ = mu.fix_mesh(p, t, dim=dim, delete_unused=True)
                return p, t

            if num_bad < num_old_bad:
                step /= gamma
            elif num_bad > num_old_bad:
                step *= gamma
            else:
                step /= 0.8

            _lib.check(_lib.lib.sm_mesher_perturb_slivers(m.h, step * h0),
                       "sm_mesher_perturb_slivers")

            if opts["preserve"]:
                # boundary vertices of the cells the perturbation started from
                p, t = m.points(), m.cells()
                p = gen._improve_level_set_newton(p, t, dom, deps)
                m.set_points(p)

            count += 1
            num_old_bad = num_bad
            _say(v2, "     Elapsed wall-clock time %f : " % (time.time() - start))
    finally:
        m.close()
    return p, t


def _termination(p, t, opts):
    # mesh_generator.py:655-677 with sliver=True (3D: lint or fix)
    return gen._termination(p, t, opts)
