"""generate_mesh / sliver_removal with the reference's signatures.

Mirrors SeismicMesh/generation/mesh_generator.py (serial path): option
parsing and defaults (:344-357, :624-652), domain/sizing unpacking
(:562-621), fixed points and box corners (:425-430), the staggered initial
lattice with seeded rejection (:808-852, generation/utils.py:15-25), the
force loop (:460-527) and termination (:485-494, :655-677, :741-759).

The force loop runs on the GPU through libsmgpu (one handle per call,
csrc/mesher.hip): device Delaunay repair, centroid filter, sorted unique
edges, bar forces with a deterministic gather, update and Newton
projection.  Set-up (lattice, RNG) and termination (one-shot mesh clean-up)
run on the host, as in the reference.

Extension: ``precision`` (64 default, or 32) selects FP64 forces that
reproduce the reference's float64 arithmetic, or FP32 forces.
"""
import ctypes
import math
import time
import warnings

import numpy as np

from . import _lib
from . import geometry as geo
from . import mesh_utils as mu
from .sizing import GridSizing, SizeFunction

__all__ = ["generate_mesh", "sliver_removal", "Mesher"]

_KNOWN = {
    "verbose", "max_iter", "seed", "perform_checks", "pfix", "axis", "points", "domain",
    "edge_length", "bbox", "min_dh_angle_bound", "max_dh_angle_bound", "delta_t", "h0",
    "geps_mult", "subdomains", "gamma", "preserve", "mesh_improvement", "r0m_is_h0",
    "precision", "surrogate", "surrogate_band", "slabs",
}


def _parse_kwargs(kwargs):
    for key in kwargs:
        if key not in _KNOWN:
            raise ValueError("Option %s with parameter %s not recognized " % (key, kwargs[key]))


def _check_bbox(bbox):
    if bbox is not None:
        for b in bbox:
            if isinstance(b, int):
                raise ValueError("bbox must contain all floats")


_DOMAINS = (geo.Ball, geo.Rectangle, geo.Disk, geo.Cube, geo.Cylinder, geo.Torus, geo.Prism,
            geo.Union, geo.Intersection, geo.Difference, geo.Repeat)


class HostDomain:
    """A domain given as a plain function sdf(x: (n, dim)) -> (n,) with a
    bbox keyword (mesh_generator.py:614-618).  The device cannot run Python:
    the force loop's centroid filter and projection call it on the host
    (sm_mesher_set_host_sdf); everything else stays on the device."""

    def __init__(self, fn, bbox):
        self.fn = fn
        self.bbox = bbox
        self.corners = None
        self.dim = len(bbox) // 2 if bbox is not None else None

    def eval(self, x):
        return np.asarray(self.fn(np.asarray(x, dtype=float)), dtype=float).reshape(-1)


def _unpack_domain(domain, opts):
    if isinstance(domain, _DOMAINS):
        bbox, corners = domain.bbox, domain.corners
    elif isinstance(domain, HostDomain):
        bbox, corners = domain.bbox, None
    elif callable(domain):
        # mesh_generator.py:614-618: the bbox comes from the keywords
        bbox, corners = opts.get("bbox"), None
        _check_bbox(bbox)
        return HostDomain(domain, bbox), bbox, corners
    else:
        raise ValueError("`domain` must be a function or a :class:`geometry` object")
    _check_bbox(bbox)
    return domain, bbox, corners


class _Sizing:
    """How the loop obtains h at edge midpoints."""

    def __init__(self, kind, value=None, grid=None, func=None):
        self.kind = kind  # "const" | "grid" | "host"
        self.value = value
        self.grid = grid
        self.func = func

    def __call__(self, x):
        if self.kind == "const":
            return np.array([self.value] * len(x))
        if self.kind == "grid":
            return self.grid(x)
        return np.asarray(self.func(x), dtype=float)


def _unpack_sizing(edge_length):
    bbox = None
    hmin = None
    if isinstance(edge_length, SizeFunction):
        bbox = edge_length.bbox
        hmin = edge_length.hmin
        lsf = True
        if isinstance(edge_length.cell_size, GridSizing):
            fh = _Sizing("grid", grid=edge_length.cell_size)
        else:
            fh = _Sizing("host", func=edge_length.eval)
    elif callable(edge_length):
        fh = _Sizing("host", func=edge_length)
        lsf = False
    elif np.isscalar(edge_length):
        fh = _Sizing("const", value=edge_length)
        hmin = edge_length
        lsf = True
    else:
        raise ValueError(
            "`edge_length` must either be a function, a `edge_length` object, or a scalar")
    _check_bbox(bbox)
    return fh, bbox, hmin, lsf


def _minmax(b0, b1):
    return tuple(max(a, b) if i % 2 else min(a, b) for i, (a, b) in enumerate(zip(b0, b1)))


def create_staggered_grid(h0, dim, bbox):
    """generation/utils.py:15-25: lattice of spacing h0, odd rows shifted."""
    points = np.mgrid[tuple(slice(mn, mx + h0, h0) for mn, mx in bbox)].astype(float)
    odd_rows = [i for i in range(points[0].shape[0]) if i % 2 != 0]
    odd_cols = [i for i in range(points[1].shape[0]) if i % 2 != 0]
    points[1][odd_rows] += h0 / 2
    if dim == 3:
        points[2][odd_cols] += h0 / 2
    return points.reshape(dim, -1).T


class Mesher:
    """Owns one libsmgpu force-loop handle (2D triangles or 3D tetrahedra)."""

    def __init__(self, levels, sizing, h0, delta_t, geps, deps, nfix, precision=64, device=0,
                 dim=2, far_square=False):
        _lib.require_gpu()
        self.dim = dim
        host = any(isinstance(l, HostDomain) for l in levels)
        # a host-function level has no program: the handle gets a placeholder
        # (never run) and the callback below evaluates every level
        self._tapes = [geo.tape_array(_placeholder(dim) if host else l.tape()) for l in levels]
        ptrs = (ctypes.c_void_p * len(levels))(*[ctypes.addressof(t) for t in self._tapes])
        lens = (ctypes.c_int * len(levels))(*[len(t) for t in self._tapes])
        prm = _lib.MesherParams()
        prm.dim = dim
        prm.precision = precision
        prm.nfix = nfix
        prm.nlevels = len(levels)
        prm.h0 = h0
        prm.delta_t = delta_t
        prm.geps = geps
        prm.deps = deps
        prm.L0mult = 1 + 0.4 / 2 ** (dim - 1)
        prm.far_square = int(bool(far_square))
        self._sizing = sizing
        g = _lib.Grid()
        g.dim = dim
        cb = _lib.HOST_SIZING_FN()
        if sizing.kind == "const":
            g.hconst = float(sizing.value)
        elif sizing.kind == "grid":
            g = sizing.grid.grid_struct()
        else:
            def _cb(ctx, mid, n, h):
                try:
                    x = np.ctypeslib.as_array(mid, shape=(n, dim)).copy()
                    out = np.ctypeslib.as_array(h, shape=(n,))
                    out[:] = np.asarray(sizing.func(x), dtype=float).reshape(-1)
                    return 0
                except Exception:  # reported by the library as a failed callback
                    return 1
            cb = _lib.HOST_SIZING_FN(_cb)
        self._cb = cb
        self._grid = g
        h = ctypes.c_void_p()
        _lib.check(_lib.lib.sm_mesher_create(ctypes.byref(prm), ptrs, lens, ctypes.byref(g), cb,
                                             None, device, ctypes.byref(h)), "sm_mesher_create")
        self.h = h
        self._sdf_cb = None
        if host:
            def _sdf(ctx, level, x, n, d):
                try:
                    xs = np.ctypeslib.as_array(x, shape=(n, dim)).copy()
                    out = np.ctypeslib.as_array(d, shape=(n,))
                    out[:] = levels[level].eval(xs)
                    return 0
                except Exception:  # reported by the library as a failed callback
                    return 1
            self._sdf_cb = _lib.HOST_SDF_FN(_sdf)
            _lib.check(_lib.lib.sm_mesher_set_host_sdf(self.h, self._sdf_cb, None),
                       "sm_mesher_set_host_sdf")

    def close(self):
        h = getattr(self, "h", None)
        if h is not None and h.value and _lib is not None and _lib.lib is not None:
            _lib.lib.sm_mesher_destroy(h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_points(self, p):
        p = np.ascontiguousarray(p, dtype=np.float64)
        self.n = len(p)
        _lib.check(_lib.lib.sm_mesher_set_points(self.h, p.ctypes.data, len(p)),
                   "sm_mesher_set_points")

    def run(self, iters):
        maxdp = ctypes.c_double(0.0)
        _lib.check(_lib.lib.sm_mesher_run(self.h, int(iters), ctypes.byref(maxdp)), "sm_mesher_run")
        return maxdp.value

    def triangulate(self):
        _lib.check(_lib.lib.sm_mesher_triangulate(self.h), "sm_mesher_triangulate")

    def sizes(self):
        n, c, e = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _lib.check(_lib.lib.sm_mesher_sizes(self.h, ctypes.byref(n), ctypes.byref(c),
                                            ctypes.byref(e)), "sm_mesher_sizes")
        return n.value, c.value, e.value

    def points(self):
        n, _, _ = self.sizes()
        p = np.empty((n, self.dim))
        _lib.check(_lib.lib.sm_mesher_get_points(self.h, p.ctypes.data, n), "sm_mesher_get_points")
        return p

    def cells(self):
        _, c, _ = self.sizes()
        t = np.empty((c, self.dim + 1), dtype=np.int32)
        _lib.check(_lib.lib.sm_mesher_get_cells(self.h, t.ctypes.data, c), "sm_mesher_get_cells")
        return t

    def triangulation(self):
        """2D: (triangles, neighbours) of every slot, ghosts (-1) and dead
        slots (-2) included."""
        cap = 2 * (len(self.points()) + 4 * 65536) + 64
        tri = np.empty((cap, 3), dtype=np.int32)
        adj = np.empty((cap, 3), dtype=np.int32)
        n = _lib.lib.sm_mesher_get_triangulation(self.h, tri.ctypes.data, adj.ctypes.data, cap)
        _lib.check(int(n) if n < 0 else 0, "sm_mesher_get_triangulation")
        return tri[:n].copy(), adj[:n].copy()

    def edges(self):
        _, _, e = self.sizes()
        out = np.empty((e, 2), dtype=np.int32)
        _lib.check(_lib.lib.sm_mesher_get_edges(self.h, out.ctypes.data, e), "sm_mesher_get_edges")
        return out

    PHASES = ("flip_round", "centroid_filter", "unique_edges", "edge_geom", "l0_sums",
              "edge_force", "vertex_update", "flip_mark", "centroid_network")

    def set_graph(self, enable=True):
        """Run FP32 iterations as one captured graph launch each (default)."""
        _lib.check(_lib.lib.sm_mesher_set_graph(self.h, int(enable)), "sm_mesher_set_graph")

    def profile(self, enable=True):
        """Enable (and clear) per-phase HIP-event timing on the handle's stream."""
        _lib.check(_lib.lib.sm_mesher_profile(self.h, int(enable)), "sm_mesher_profile")

    def profile_read(self):
        k = len(self.PHASES)
        ms = (ctypes.c_double * k)()
        n = (ctypes.c_int64 * k)()
        _lib.check(_lib.lib.sm_mesher_profile_read(self.h, ms, n, k), "sm_mesher_profile_read")
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(self.PHASES)}

    def set_surrogate(self, net, band, bands=None):
        """Let `net` (a neural.NeuralSDF of this handle's dimension) answer
        the centroid filter and the projection test where |d_nn| >= band;
        net=None switches it off.  `bands` = (grid, lo, hi) from
        NeuralSDF.calibrate_bands gives each region its own band (at least
        `band`).  The handle keeps the weights the network has now (train
        before, not after)."""
        if net is None:
            self._net = None
            _lib.check(_lib.lib.sm_mesher_set_surrogate(self.h, None, 0.0),
                       "sm_mesher_set_surrogate")
            _lib.check(_lib.lib.sm_mesher_set_surrogate_bands(self.h, 0, None, None, None),
                       "sm_mesher_set_surrogate_bands")
            return
        st = net.struct()
        self._net = (net, st)
        _lib.check(_lib.lib.sm_mesher_set_surrogate(self.h, ctypes.byref(st), float(band)),
                   "sm_mesher_set_surrogate")
        if bands is None:
            _lib.check(_lib.lib.sm_mesher_set_surrogate_bands(self.h, 0, None, None, None),
                       "sm_mesher_set_surrogate_bands")
        else:
            grid, lo, hi = bands
            grid = np.ascontiguousarray(grid, dtype=np.float32)
            lo = np.ascontiguousarray(lo, dtype=np.float64)
            hi = np.ascontiguousarray(hi, dtype=np.float64)
            if grid.ndim != self.dim or len(set(grid.shape)) != 1:
                raise ValueError("bands must be a res^dim grid")
            self._bands = (grid, lo, hi)
            _lib.check(_lib.lib.sm_mesher_set_surrogate_bands(
                self.h, int(grid.shape[0]), lo.ctypes.data, hi.ctypes.data, grid.ctypes.data),
                "sm_mesher_set_surrogate_bands")

    def surrogate_stats(self):
        """(exact evaluations skipped, exact evaluations made, max |d_nn - d|
        over the evaluations made)."""
        a, b, e = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_double()
        _lib.check(_lib.lib.sm_mesher_surrogate_stats(self.h, ctypes.byref(a), ctypes.byref(b),
                                                      ctypes.byref(e)),
                   "sm_mesher_surrogate_stats")
        return {"skipped": a.value, "exact": b.value, "max_err": e.value}

    TUNABLE = ("centroid_filter", "edge_force", "vertex_update", "flip_orient", "flip_mark",
               "flip_apply")

    def autotune(self, reps=20):
        """Pick the workgroup size of each tunable kernel by timing 64..1024
        threads on the current state (run one iteration first).  Returns
        {kernel: (threads, ms per launch)}."""
        k = len(self.TUNABLE)
        chosen = (ctypes.c_int32 * k)()
        ms = (ctypes.c_double * k)()
        _lib.check(_lib.lib.sm_mesher_autotune(self.h, int(reps), chosen, ms, k),
                   "sm_mesher_autotune")
        return {name: (int(chosen[i]), float(ms[i])) for i, name in enumerate(self.TUNABLE)}

    def set_workgroups(self, sizes):
        """Set the tunable kernels' workgroup sizes ({kernel: threads} for
        the names in TUNABLE, or a sequence in that order)."""
        if isinstance(sizes, dict):
            sizes = [int(sizes[k]) for k in self.TUNABLE]
        arr = (ctypes.c_int32 * len(self.TUNABLE))(*[int(v) for v in sizes])
        _lib.check(_lib.lib.sm_mesher_set_workgroups(self.h, arr, len(self.TUNABLE)),
                   "sm_mesher_set_workgroups")

    def time_kernel(self, phase="flip_mark", reps=200):
        """Average ms of `reps` back-to-back launches of a phase's kernel on
        the current state (HIP events on the handle's stream)."""
        ms = ctypes.c_double()
        _lib.check(_lib.lib.sm_mesher_time_kernel(self.h, self.PHASES.index(phase), int(reps),
                                                  ctypes.byref(ms)), "sm_mesher_time_kernel")
        return ms.value

    STATS = ("iterations", "flip_rounds", "flips", "host_rebuilds", "exact_predicates",
             "rollbacks", "kinetic_repairs", "unresolved", "relocations", "parallel_builds",
             "retried_points", "tets_last_build", "host_tet_us")

    def stats(self):
        k = len(self.STATS)
        s = (ctypes.c_int64 * k)()
        _lib.check(_lib.lib.sm_mesher_stats(self.h, s, k), "sm_mesher_stats")
        return {name: int(s[i]) for i, name in enumerate(self.STATS)}

    def set_slabs(self, nslabs, axis=1, threads=None):
        """3D: build each iteration's tetrahedralisation on `threads` host
        threads (default: one per slab, at most the CPU count) inserting the
        points of `nslabs` equal-width slabs along coordinate `axis` (the
        reference's `axis` option) into one shared tetrahedralisation: the
        cells are those of the serial build."""
        import os
        if threads is None:
            threads = max(1, min(int(nslabs), os.cpu_count() or 1))
        _lib.check(_lib.lib.sm_mesher_set_slabs(self.h, int(axis), int(nslabs), int(threads)),
                   "sm_mesher_set_slabs")


def _placeholder(dim):
    return (geo.Disk([0.0, 0.0], 1.0) if dim == 2 else geo.Ball([0.0, 0.0, 0.0], 1.0)).tape()


def _fd(domain):
    return domain.eval


def _remove_triangles_outside(p, t, domain, geps):
    # mesh_generator.py:734-738
    dim = p.shape[1]
    pmid = p[t].sum(1) / (dim + 1)
    return t[domain.eval(pmid) < -geps]


def _improve_level_set_newton(p, t, domain, deps):
    # mesh_generator.py:741-759 (five damped Newton steps on boundary vertices)
    dim = p.shape[1]
    bid = mu.get_boundary_vertices(t, dim)
    alpha = 1
    for it in range(5):
        d = domain.eval(p[bid])
        grads = []
        for i in range(dim):
            q = p[bid].copy()
            q[:, i] = q[:, i] + deps
            grads.append((domain.eval(q) - d) / deps)
        g2 = sum(g ** 2 for g in grads)
        g2 = np.where(g2 < deps, deps, g2)
        p[bid] -= alpha * (d * np.vstack(grads) / g2).T
        alpha /= it + 1
    return p


def _termination(p, t, opts, verbose=1):
    # mesh_generator.py:655-677 (serial)
    dim = p.shape[1]
    if dim == 2:
        p, t, _ = mu.fix_mesh(p, t, dim=dim, delete_unused=True)
        p, t = mu.delete_boundary_entities(p, t, dim=2, min_qual=0.15, verbose=verbose)
        if opts["subdomains"] is None and opts["mesh_improvement"]:
            p, t = mu.laplacian2_fixed_point(p, t)
    if opts.get("perform_checks"):
        p, t = mu.linter(p, t, dim=dim)
    else:
        p, t, _ = mu.fix_mesh(p, t, dim=dim, delete_unused=True)
    return p, t


def prepare(domain, edge_length, **kwargs):
    """The set-up half of generate_mesh (mesh_generator.py:343-458): option
    checks, domain and sizing unpacking, fixed points, initial points."""
    opts = {
        "verbose": 1, "max_iter": 50, "seed": 0, "perform_checks": False, "pfix": None,
        "axis": 1, "points": None, "delta_t": 0.30, "geps_mult": 0.1, "subdomains": None,
        "mesh_improvement": True, "r0m_is_h0": False, "precision": 64,
    }
    opts.update(kwargs)
    _parse_kwargs(kwargs)
    if opts["verbose"] < 0:
        raise ValueError("Unknown verbosity level")
    v1 = opts["verbose"] >= 1
    v2 = opts["verbose"] > 1

    dom, bbox0, corners = _unpack_domain(domain, opts)
    fh, bbox1, hmin, lsf = _unpack_sizing(edge_length)
    bbox = bbox0 if bbox1 is None else _minmax(bbox0, bbox1)
    if not isinstance(bbox, tuple):
        raise ValueError("`bbox` must be a tuple")
    dim = int(len(bbox) / 2)
    if dim not in (2, 3):
        raise ValueError("`bbox` must describe a 2D or 3D box")
    # mesh_generator.py:388-390: corners are not constrained in 3D with graded sizing
    if fh.kind != "const" and dim == 3:
        corners = None
    if bbox1 is not None and bbox0 != bbox1:
        dom = geo.Rectangle(bbox) if dim == 2 else geo.Cube(bbox)
    bbox = np.array(bbox).reshape(-1, 2)
    h0 = hmin if hmin is not None else opts["h0"]
    if h0 < 0:
        raise ValueError("`h0` must be > 0")
    delta_t = opts["delta_t"]
    geps = opts["geps_mult"] * h0
    deps = np.sqrt(np.finfo(np.double).eps) * h0

    if opts["pfix"] is not None:
        pfix = np.array(opts["pfix"], dtype="d")
    else:
        pfix = np.empty((0, dim))
    if corners is not None:
        corners = corners[dom.eval(corners) > -1000 * deps]
        pfix = np.append(pfix, corners, axis=0)
    nfix = len(pfix)
    if v1:
        print(f"Constraining {nfix} fixed points...", flush=True)

    if opts["points"] is None:
        p = create_staggered_grid(h0, dim, bbox)
        p = p[dom.eval(p) < geps]
        r0 = fh(p)
        if opts["r0m_is_h0"]:
            r0m = 1.1 * h0 if 1.1 * h0 < r0.min() else h0
        else:
            r0m = r0.min()
        np.random.seed(opts["seed"])
        p = np.vstack((pfix, p[np.random.rand(p.shape[0]) < r0m ** dim / r0 ** dim]))
    else:
        p = np.array(opts["points"], dtype=float)

    if opts["max_iter"] < 0:
        raise ValueError("`max_iter` must be > 0")
    max_iter = opts["max_iter"]
    N = p.shape[0]
    assert N > 0, "No vertices to mesh with!"
    if v1:
        print("Commencing mesh generation with %d vertices on rank %d." % (N, 0), flush=True)

    levels = [dom]
    if opts["subdomains"] is not None:
        for sd in opts["subdomains"]:
            levels.append(_unpack_domain(sd, opts)[0])

    return dict(opts=opts, dom=dom, fh=fh, h0=h0, delta_t=delta_t, geps=geps, deps=deps,
                nfix=nfix, p=p, levels=levels, max_iter=max_iter, dim=dim)


# 3D: below this many points one thread builds the tetrahedralisation
SLAB_MIN_POINTS = 20000


def default_slabs(npoints, dim, requested=None):
    """Slab count for the 3D host tetrahedralisation: `requested` if given
    (1 = serial), else one slab per CPU (at most 16) for 3D problems of at
    least SLAB_MIN_POINTS points."""
    import os
    if dim != 3:
        return 1
    if requested is not None:
        return max(1, int(requested))
    if npoints < SLAB_MIN_POINTS:
        return 1
    return max(1, min(16, os.cpu_count() or 1))


def make_mesher(st, far_square=False, slabs=None):
    """A force-loop handle for a prepared problem (initial points not set).
    3D handles build their tetrahedralisation on one thread per slab along
    the `axis` option (default_slabs; `slabs` or the `slabs` option
    overrides)."""
    m = Mesher(st["levels"], st["fh"], st["h0"], st["delta_t"], st["geps"], st["deps"],
               st["nfix"], precision=st["opts"]["precision"], dim=st["dim"],
               far_square=far_square)
    if st["dim"] == 3:
        req = slabs if slabs is not None else st["opts"].get("slabs")
        ns = default_slabs(len(st["p"]), 3, req)
        if ns > 1:
            m.set_slabs(ns, axis=int(st["opts"].get("axis", 1)))
    return m


Mesher2D = Mesher


def generate_mesh(domain, edge_length, comm=None, **kwargs):
    """Generate a 2D or 3D mesh: see mesh_generator.py:291-529 for the semantics."""
    st = prepare(domain, edge_length, **kwargs)
    opts, dom, h0, deps = st["opts"], st["dom"], st["h0"], st["deps"]
    v1 = opts["verbose"] >= 1
    v2 = opts["verbose"] > 1
    m = make_mesher(st)
    if opts.get("surrogate") is not None:
        # band |d_nn| < surrogate_band * h0 falls back to the exact SDF
        m.set_surrogate(opts["surrogate"], opts.get("surrogate_band", 2.0) * h0)
    m.set_points(st["p"])
    updates = max(st["max_iter"] - 1, 0)
    if v2:
        for count in range(updates):
            start = time.time()
            maxdp = m.run(1)
            n, c, _ = m.sizes()
            print("Iteration #%d, max movement is %f, there are %d vertices and %d cells"
                  % (count + 1, maxdp, n, c), flush=True)
            print("     Elapsed wall-clock time %f : " % (time.time() - start), flush=True)
    else:
        m.run(updates)
    m.triangulate()
    p, t = m.points(), m.cells()
    if v1:
        print("Termination reached...maximum number of iterations reached.", flush=True)
    p, t = _termination(p, t, opts, verbose=opts["verbose"])
    p = _improve_level_set_newton(p, t, dom, deps)
    t = _remove_triangles_outside(p, t, dom, h0 * 0.001)
    return p, t


def sliver_removal(points, domain, edge_length, comm=None, **kwargs):
    """3D sliver removal (mesh_generator.py:59-288)."""
    from .sliver import sliver_removal as _impl
    return _impl(points, domain, edge_length, comm=comm, **kwargs)
