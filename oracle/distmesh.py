"""Oracle restatement of the DistMesh loops (FP64 numpy).

Test infrastructure only (tests/, the smoke check and the bench's CPU
baseline import it; the product never does).

Follows SeismicMesh/generation/mesh_generator.py (serial path) and
SeismicMesh/generation/utils.py.  Differences, all forced by what this image
lacks:

* Triangulation: scipy's Qhull stands in for CGAL's Delaunay
  (delaunay_class.cpp / delaunay_class3.cpp, absent here).  For inputs in
  general position both give the unique Delaunay triangulation of the convex
  hull; the oracle keeps the caller's vertex order, whereas CGAL renumbers
  vertices in its container order each iteration (mesh_generator.py:469).
  Vertex order only changes the order of the sorted edge list and hence of
  floating point summation, so the oracle defines the order that the GPU
  path is compared against.
* The L0 normalisation sums use numpy's own summation (np.sum), exactly as
  the reference does (mesh_generator.py:700).
* Termination smoothing solves its linear system directly (pyamg absent).
"""
import math
import time

import numpy as np
from scipy.spatial import Delaunay

from . import geometry as geo
from .sizing import SizeFunction

GEN_DEFAULTS = {
    # mesh_generator.py:344-357
    "verbose": 1, "max_iter": 50, "seed": 0, "perform_checks": False, "pfix": None,
    "axis": 1, "points": None, "delta_t": 0.30, "geps_mult": 0.1, "subdomains": None,
    "mesh_improvement": True, "r0m_is_h0": False,
}
SLIVER_DEFAULTS = {
    # mesh_generator.py:107-120
    "verbose": 1, "max_iter": 50, "perform_checks": False, "axis": 1,
    "min_dh_angle_bound": 10.0, "max_dh_angle_bound": 180.0, "points": None,
    "delta_t": 0.30, "geps_mult": 0.1, "subdomains": [], "gamma": 1.0, "preserve": False,
}
_KNOWN = {
    # mesh_generator.py:624-652
    "verbose", "max_iter", "seed", "perform_checks", "pfix", "axis", "points", "domain",
    "edge_length", "bbox", "min_dh_angle_bound", "max_dh_angle_bound", "delta_t", "h0",
    "geps_mult", "subdomains", "gamma", "preserve", "mesh_improvement", "r0m_is_h0",
}


def parse_kwargs(kwargs):
    for k in kwargs:
        if k not in _KNOWN:
            raise ValueError("Option %s with parameter %s not recognized " % (k, kwargs[k]))


# --------------------------------------------------------------------------
# initial points (generation/utils.py:15-45, mesh_generator.py:808-852)
# --------------------------------------------------------------------------
def create_staggered_grid(h0, dim, bbox):
    points = np.mgrid[tuple(slice(mn, mx + h0, h0) for mn, mx in bbox)].astype(float)
    odd_rows = [i for i in range(points[0].shape[0]) if i % 2 != 0]
    odd_cols = [i for i in range(points[1].shape[0]) if i % 2 != 0]
    points[1][odd_rows] += h0 / 2
    if dim == 3:
        points[2][odd_cols] += h0 / 2
    return points.reshape(dim, -1).T


def generate_initial_points(h0, geps, dim, bbox, fh, fd, pfix, opts):
    """Serial path of _generate_initial_points (mesh_generator.py:808-852)."""
    p = create_staggered_grid(h0, dim, bbox)
    p = p[fd(p) < geps]
    r0 = fh(p)
    if opts["r0m_is_h0"]:
        if 1.1 * h0 < r0.min():
            r0m = 1.1 * h0
        else:
            r0m = h0
    else:
        r0m = r0.min()
    np.random.seed(opts["seed"])
    return np.vstack((pfix, p[np.random.rand(p.shape[0]) < r0m ** dim / r0 ** dim]))


# --------------------------------------------------------------------------
# one force iteration (mesh_generator.py:460-527)
# --------------------------------------------------------------------------
def delaunay(p):
    """Delaunay cells of the convex hull of p, positively oriented."""
    dim = p.shape[1]
    if len(p) < dim + 1:
        return np.zeros((0, dim + 1), dtype=np.int32)
    t = Delaunay(p).simplices.astype(np.int32)
    if len(t) == 0:
        return t
    v = geo.simp_vol(p, t)
    flip = v < 0
    t[flip, :2] = t[flip, 1::-1]
    return t


def remove_triangles_outside(p, t, fd, geps):
    # mesh_generator.py:734-738
    dim = p.shape[1]
    pmid = p[t].sum(1) / (dim + 1)
    return t[fd(pmid) < -geps]


def get_edges(t):
    # mesh_generator.py:680-688
    dim = t.shape[1] - 1
    e = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]])
    if dim == 3:
        e = np.concatenate((e, t[:, [0, 3]], t[:, [1, 3]], t[:, [2, 3]]), axis=0)
    return geo.unique_edges(e)


def scatter_dense(edges, Fvec, N):
    """generation/utils.py:48-68 (coo_matrix(...).toarray()): every
    contribution is added in entry order -- for vertex v that is increasing
    edge order, +Fvec where v is the first endpoint, -Fvec where second."""
    import scipy.sparse as sp

    dim = Fvec.shape[1]
    Ix = edges[:, [0] * dim + [1] * dim]
    J = np.repeat([list(range(dim)) * 2], len(Fvec), axis=0)
    S = np.hstack((Fvec, -Fvec))
    return sp.coo_matrix((S.flat, (Ix.flat, J.flat)), (N, dim)).toarray()


def compute_forces(p, t, fh, h0, L0mult, return_parts=False):
    # mesh_generator.py:691-712
    dim = p.shape[1]
    N = p.shape[0]
    edges = get_edges(t)
    barvec = p[edges[:, 0]] - p[edges[:, 1]]
    L = np.sqrt((barvec ** 2).sum(1))
    L[L == 0] = np.finfo(float).eps
    hedges = fh(p[edges].sum(1) / 2)
    L0 = hedges * L0mult * ((L ** dim).sum() / (hedges ** dim).sum()) ** (1.0 / dim)
    F = L0 - L
    F[F < 0] = 0
    Fvec = F[:, None] / L[:, None].dot(np.ones((1, dim))) * barvec
    Ftot = scatter_dense(edges, Fvec, N)
    if return_parts:
        return Ftot, dict(edges=edges, L=L, hedges=hedges, L0=L0, F=F, Fvec=Fvec)
    return Ftot


def project_points_back_newton(p, fd, deps, hmin, idx):
    # mesh_generator.py:762-784
    dim = p.shape[1]
    d = fd(p)
    ix = d > 0.0 if idx == 0 else np.logical_and(d > 0.0, d < hmin / 1.5)
    if ix.any():
        def dv(i):
            a = [0] * dim
            a[i] = deps
            return a
        dgrads = [(fd(p[ix] + dv(i)) - d[ix]) / deps for i in range(dim)]
        dgrad2 = sum(g ** 2 for g in dgrads)
        dgrad2 = np.where(dgrad2 < deps, deps, dgrad2)
        p[ix] -= (d[ix] * np.vstack(dgrads) / dgrad2).T
    return p


def improve_level_set_newton(p, t, fd, deps, tol):
    # mesh_generator.py:741-759
    dim = p.shape[1]
    bid = geo.get_boundary_vertices(t, dim)
    alpha = 1
    for it in range(5):
        d = fd(p[bid])

        def dv(i):
            a = [0] * dim
            a[i] = deps
            return a
        dgrads = [(fd(p[bid] + dv(i)) - d) / deps for i in range(dim)]
        dgrad2 = sum(g ** 2 for g in dgrads)
        dgrad2 = np.where(dgrad2 < deps, deps, dgrad2)
        p[bid] -= alpha * (d * np.vstack(dgrads) / dgrad2).T
        alpha /= it + 1
    return p


def closest_node(node, nodes):
    # mesh_generator.py:911-915
    deltas = np.asarray(nodes) - node
    return np.argmin(np.einsum("ij,ij->i", deltas, deltas))


def termination(p, t, opts, sliver=False, verbose=1):
    # mesh_generator.py:655-677 (serial)
    dim = p.shape[1]
    if dim == 2:
        p, t, _ = geo.fix_mesh(p, t, dim=dim, delete_unused=True)
        p, t = geo.delete_boundary_entities(p, t, dim=2, min_qual=0.15, verbose=verbose)
        if opts["subdomains"] is None and opts["mesh_improvement"]:
            p, t = geo.laplacian2_fixed_point(p, t)
    p, t, _ = geo.fix_mesh(p, t, dim=dim, delete_unused=True)
    return p, t


# --------------------------------------------------------------------------
# drivers
# --------------------------------------------------------------------------
def unpack_domain(domain, opts):
    # mesh_generator.py:594-621
    if hasattr(domain, "eval") and hasattr(domain, "bbox"):
        return domain.eval, domain.bbox, getattr(domain, "corners", None)
    if callable(domain):
        return domain, opts["bbox"], None
    raise ValueError("`domain` must be a function or a :class:`geometry` object")


def unpack_sizing(edge_length):
    # mesh_generator.py:562-591
    if isinstance(edge_length, SizeFunction) or type(edge_length).__name__ == "SizeFunction":
        return edge_length.eval, edge_length.bbox, edge_length.hmin, True
    if callable(edge_length):
        return edge_length, None, None, False
    if np.isscalar(edge_length):
        def func(x):
            return np.array([edge_length] * len(x))
        return func, None, edge_length, True
    raise ValueError("`edge_length` must either be a function, a `edge_length` object, or a scalar")


def _minmax(b0, b1):
    return tuple(max(a, b) if i % 2 else min(a, b) for i, (a, b) in enumerate(zip(b0, b1)))


def setup(domain, edge_length, opts):
    """The set-up half of generate_mesh (mesh_generator.py:373-458)."""
    fd, bbox0, corners = unpack_domain(domain, opts)
    fh, bbox1, hmin, lsf = unpack_sizing(edge_length)
    bbox = bbox0 if bbox1 is None else _minmax(bbox0, bbox1)
    if not isinstance(bbox, tuple):
        raise ValueError("`bbox` must be a tuple")
    dim = len(bbox) // 2
    if not np.isscalar(edge_length) and dim == 3:
        corners = None
    if bbox1 is not None and bbox0 != bbox1:
        fd = (geo.Rectangle(bbox) if dim == 2 else geo.Cube(bbox)).eval
    bbox = np.array(bbox).reshape(-1, 2)
    h0 = hmin if hmin is not None else opts["h0"]
    if h0 < 0:
        raise ValueError("`h0` must be > 0")
    L0mult = 1 + 0.4 / 2 ** (dim - 1)
    geps = opts["geps_mult"] * h0
    deps = np.sqrt(np.finfo(np.double).eps) * h0
    if opts["pfix"] is not None:
        pfix = np.array(opts["pfix"], dtype="d")
    else:
        pfix = np.empty((0, dim))
    if corners is not None:
        corners = corners[fd(corners) > -1000 * deps]
        pfix = np.append(pfix, corners, axis=0)
    if opts["points"] is None:
        p = generate_initial_points(h0, geps, dim, bbox, fh, fd, pfix, opts)
    else:
        p = np.array(opts["points"], dtype=float)
    levels = [fd]
    if opts["subdomains"] is not None:
        for sd in opts["subdomains"]:
            levels.append(unpack_domain(sd, opts)[0])
    return dict(fd=fd, fh=fh, h0=h0, dim=dim, L0mult=L0mult, geps=geps, deps=deps,
                pfix=pfix, p=p, levels=levels, delta_t=opts["delta_t"])


def force_step(p, t_all, s, nfix_idx=None):
    """One position update of the loop (mesh_generator.py:481-506), on a
    triangulation t_all of p (all Delaunay cells of the hull).  Returns the
    new points, the kept cells and the iteration's parts."""
    t = remove_triangles_outside(p, t_all, s["fd"], s["geps"])
    Ftot, parts = compute_forces(p, t, s["fh"], s["h0"], s["L0mult"], return_parts=True)
    if nfix_idx is None:
        nfix_idx = [closest_node(f, p) for f in s["pfix"]]
    Ftot[list(nfix_idx)] = 0
    p = p + s["delta_t"] * Ftot
    for idx, level in enumerate(s["levels"]):
        p = project_points_back_newton(p, level, s["deps"], s["h0"], idx)
    parts["Ftot"] = Ftot
    parts["maxdp"] = s["delta_t"] * np.sqrt((Ftot ** 2).sum(1)).max()
    return p, t, parts


def generate_mesh(domain, edge_length, triangulate=delaunay, history=None, **kwargs):
    """Serial generate_mesh (mesh_generator.py:291-529).

    ``triangulate`` maps points to positively oriented hull cells.
    ``history``, if a list, receives (p, t_kept) of every iteration.
    """
    opts = dict(GEN_DEFAULTS)
    opts.update(kwargs)
    parse_kwargs(kwargs)
    v1 = opts["verbose"] >= 1
    v2 = opts["verbose"] > 1
    s = setup(domain, edge_length, opts)
    nfix = len(s["pfix"])
    if v1:
        print(f"Constraining {nfix} fixed points...", flush=True)
    if opts["max_iter"] < 0:
        raise ValueError("`max_iter` must be > 0")
    max_iter = opts["max_iter"]
    p = s["p"]
    assert p.shape[0] > 0, "No vertices to mesh with!"
    if v1:
        print("Commencing mesh generation with %d vertices on rank %d." % (p.shape[0], 0), flush=True)
    count = 0
    while True:
        start = time.time()
        t_all = triangulate(p)
        ifix = [closest_node(f, p) for f in s["pfix"]]
        if count == max_iter - 1:
            t = remove_triangles_outside(p, t_all, s["fd"], s["geps"])
            if v1:
                print("Termination reached...maximum number of iterations reached.", flush=True)
            p, t = termination(p, t, opts, verbose=opts["verbose"])
            p = improve_level_set_newton(p, t, s["fd"], s["deps"], s["deps"] * 1000)
            t = remove_triangles_outside(p, t, s["fd"], s["h0"] * 0.001)
            break
        p, t, parts = force_step(p, t_all, s, ifix)
        if history is not None:
            history.append((p.copy(), t.copy()))
        maxdp = parts["maxdp"]
        if v2:
            print("Iteration #%d, max movement is %f, there are %d vertices and %d cells"
                  % (count + 1, maxdp, len(p), len(t)), flush=True)
        assert maxdp < 1000 * s["h0"], "max movement indicates there's a convergence problem"
        count += 1
        if v2:
            print("     Elapsed wall-clock time %f : " % (time.time() - start), flush=True)
    return p, t


def calc_slivers(p, t, min_dh, max_dh):
    # mesh_generator.py:532-540
    dh = geo.calc_dihedral_angles(p, t)
    oob = np.argwhere((dh[:, 0] < min_dh) | (dh[:, 0] > max_dh))
    ele = np.floor(oob / 6).astype("int")
    ele, ix = np.unique(ele, return_index=True)
    return ele, ix


def sliver_removal(points, domain, edge_length, triangulate=delaunay, **kwargs):
    """Serial sliver_removal (mesh_generator.py:59-288).  CGAL's incremental
    move (mesh_generator.py:210-212) is a re-triangulation of the moved
    points; for points in general position the result is the same."""
    opts = dict(SLIVER_DEFAULTS)
    opts.update(kwargs)
    parse_kwargs(kwargs)
    v1 = opts["verbose"] >= 1
    dim = points.shape[1]
    if dim == 2:
        raise Exception("Mesh improvement currently on works in 3D")
    fd, bbox0, _ = unpack_domain(domain, opts)
    fh, bbox1, hmin, _ = unpack_sizing(edge_length)
    bbox = bbox0 if bbox1 is None else _minmax(bbox0, bbox1)
    if bbox1 is not None and bbox0 != bbox1:
        fd = geo.Cube(bbox).eval
    if not isinstance(bbox, tuple):
        raise ValueError("`bbox` must be a tuple")
    h0 = hmin if hmin is not None else opts["h0"]
    if h0 < 0:
        raise ValueError("`h0` must be > 0")
    if opts["max_iter"] < 0:
        raise ValueError("`max_iter` must be > 0")
    max_iter = opts["max_iter"]
    if v1:
        print(f"Will attempt {max_iter} iterations to bound the dihedral angles...", flush=True)
    geps = opts["geps_mult"] * h0
    deps = np.sqrt(np.finfo(np.double).eps) * h0
    min_dh = opts["min_dh_angle_bound"] * math.pi / 180
    max_dh = opts["max_dh_angle_bound"] * math.pi / 180
    if v1:
        print(f"Enforcing a min. dihedral bound of: {min_dh * 180 / math.pi} degrees...", flush=True)
        print(f"Enforcing a max. dihedral bound of: {max_dh * 180 / math.pi} degrees...", flush=True)
    p = np.array(points, dtype=float)
    if v1:
        print("Commencing sliver removal with %d vertices on rank %d." % (len(p), 0), flush=True)
    count = 0
    step = 0.10
    gamma = opts["gamma"]
    num_old_bad = np.inf
    while True:
        t = triangulate(p)
        t = remove_triangles_outside(p, t, fd, geps)
        ele, _ = calc_slivers(p, t, min_dh, max_dh)
        if count == max_iter - 1:
            if v1:
                print("FAILURE: Termination...maximum number of iterations reached. "
                      "Try increasing max_iter when generating the mesh", flush=True)
            p, t = termination(p, t, opts, sliver=True)
            break
        if v1:
            print(f"On rank: 0. There are {len(ele)} slivers...", flush=True)
        move = t[ele, 0]
        if move.size == 0:
            if v1:
                print(f"Termination reached in {count} iterations...no slivers detected!", flush=True)
            p, t, _ = geo.fix_mesh(p, t, dim=dim, delete_unused=True)
            return p, t
        perturb = geo.calc_circumsphere_grad(p[t[ele, 0]], p[t[ele, 1]], p[t[ele, 2]], p[t[ele, 3]])
        perturb[np.isinf(perturb)] = 1.0
        perturb /= (np.sum(np.abs(perturb) ** 2, axis=-1) ** 0.5)[:, None]
        num_bad = len(ele)
        if num_bad < num_old_bad:
            step /= gamma
        elif num_bad > num_old_bad:
            step *= gamma
        else:
            step /= 0.8
        p[move] += step * h0 * perturb
        if opts["preserve"]:
            p = improve_level_set_newton(p, t, fd, deps, deps * 1000)
        count += 1
        num_old_bad = num_bad
    return p, t
