"""Mesh utilities of the reference's geometry module (host, one-shot).

Mirrors SeismicMesh/geometry/utils.py: simp_vol (:175-201), fix_mesh
(:204-249), simp_qual (:252-274), edge/facet/boundary helpers (:277-437),
delete_boundary_entities (:440-469) and laplacian2_fixed_point (:494-547).
These run once, after the force loop (termination), on the host; the
Laplacian system is solved directly (the reference uses pyamg's
Ruge-Stuben solver on the same matrix).
"""
import numpy as np

__all__ = [
    "simp_vol", "simp_qual", "fix_mesh", "unique_rows", "get_edges", "get_facets",
    "get_boundary_edges", "get_boundary_facets", "get_boundary_vertices",
    "vertex_to_entities", "get_boundary_entities", "delete_boundary_entities",
    "laplacian2_fixed_point", "get_centroids", "linter", "do_any_overlap", "is_manifold",
    "dihedral_angles",
]


def simp_vol(p, t):
    dim = p.shape[1]
    if dim == 2:
        d01 = p[t[:, 1]] - p[t[:, 0]]
        d02 = p[t[:, 2]] - p[t[:, 0]]
        return (d01[:, 0] * d02[:, 1] - d01[:, 1] * d02[:, 0]) / 2
    if dim == 3:
        d01 = p[t[:, 1]] - p[t[:, 0]]
        d02 = p[t[:, 2]] - p[t[:, 0]]
        d03 = p[t[:, 3]] - p[t[:, 0]]
        return np.einsum("ij,ij->i", np.cross(d01, d02), d03) / 6
    raise NotImplementedError


def simp_qual(p, t):
    """2 * inradius / circumradius of triangles (1 = equilateral)."""
    assert p.ndim == 2 and t.ndim == 2 and p.shape[1] + 1 == t.shape[1]

    def length(a):
        return np.sqrt((a ** 2).sum(1))

    a = length(p[t[:, 1]] - p[t[:, 0]])
    b = length(p[t[:, 2]] - p[t[:, 0]])
    c = length(p[t[:, 2]] - p[t[:, 1]])
    r = 0.5 * np.sqrt((b + c - a) * (c + a - b) * (a + b - c) / (a + b + c))
    R = a * b * c / np.sqrt((a + b + c) * (b + c - a) * (c + a - b) * (a + b - c))
    return 2 * r / R


def unique_rows(A, return_index=False, return_inverse=False):
    """Unique rows compared as raw bytes (the reference's row view)."""
    A = np.require(A, requirements="C")
    assert A.ndim == 2, "array must be 2-dim'l"
    dt = np.dtype("S%d" % (A.dtype.itemsize * A.shape[1]))
    B, I, J = np.unique(A.view(dt), return_index=True, return_inverse=True)
    B = B.view(A.dtype).reshape((-1, A.shape[1]))
    J = J.reshape(-1)
    out = (B,)
    if return_index:
        out += (I,)
    if return_inverse:
        out += (J,)
    return out if len(out) > 1 else out[0]


def _unique_row_view(data):
    b = np.ascontiguousarray(data).view(np.dtype((np.void, data.dtype.itemsize * data.shape[1])))
    u, cnts = np.unique(b, return_counts=True)
    return u.view(data.dtype).reshape(-1, data.shape[1]), cnts


def _used_ids(t):
    # np.unique(t, return_inverse=True) for integer ids: the ids present,
    # sorted, and each entry's rank among them -- by counting when the ids
    # are non-negative (linear time), the same arrays
    flat = np.asarray(t).reshape(-1)
    if flat.size == 0 or not np.issubdtype(flat.dtype, np.integer) or flat.min() < 0:
        pix, _, jx = np.unique(t, return_index=True, return_inverse=True)
        return pix, jx
    present = np.bincount(flat) > 0
    rank = np.cumsum(present, dtype=np.intp) - 1
    return np.flatnonzero(present).astype(flat.dtype), rank[t]


def fix_mesh(p, t, ptol=2e-13, dim=2, delete_unused=False, fix_orientation=True):
    """Merge duplicate vertices/cells, drop unused vertices, orient cells."""
    snap = (p.max(0) - p.min(0)).max() * ptol
    _, ix, jx = unique_rows(np.round(p / snap) * snap, True, True)
    p = p[ix]
    t = jx[t]
    t = np.sort(t, axis=1)
    t, _ = _unique_row_view(t)
    if delete_unused:
        pix, jx = _used_ids(t)
        t = np.reshape(jx, t.shape)
        p = p[pix, :]
    if fix_orientation:
        flip = simp_vol(p, t) < 0
        t[flip, :2] = t[flip, 1::-1]
    return p, t, jx


def get_centroids(p, t):
    dim = p.shape[1]
    return p[t].sum(1) / (dim + 1)


def get_edges(t, dim=2):
    t = np.asarray(t)
    if dim == 2:
        return t[:, [[0, 1], [0, 2], [1, 2]]].reshape(-1, 2)
    return t[:, [[0, 1], [1, 2], [2, 0], [0, 3], [1, 3], [2, 3]]].reshape(-1, 2)


def get_facets(t):
    ix = [[0, 1, 3], [1, 2, 3], [2, 0, 3], [1, 2, 0]]
    return np.array(t[:, ix]).reshape((len(t) * 4, 3))


def get_boundary_edges(t, dim=2):
    e = np.sort(get_edges(t, dim=dim), axis=1)
    unq, cnt = _unique_row_view(e)
    return unq[cnt == (dim - 1)]


def get_boundary_facets(t):
    if t.shape[1] < 4:
        raise ValueError("Only works for triangles")
    f = np.sort(get_facets(t), axis=1)
    unq, cnt = _unique_row_view(f)
    return unq[cnt == 1]


def _boundary_facets_any_order(t):
    # the facets of one cell only, in no particular order: each sorted facet
    # as one int64 key, sorted, the keys seen once kept -- several times
    # faster than the unique-row view on millions of facets
    f = np.sort(get_facets(t), axis=1)
    nv = int(f.max()) + 1 if f.size else 0
    if nv == 0 or nv > (1 << 21):  # nv^3 must fit 64 bits
        return get_boundary_facets(t)
    w = np.uint64(nv)
    k = (f[:, 0].astype(np.uint64) * w + f[:, 1].astype(np.uint64)) * w + f[:, 2].astype(np.uint64)
    k.sort()
    new = np.ones(len(k), dtype=bool)
    new[1:] = k[1:] != k[:-1]
    start = np.flatnonzero(new)
    once = k[start[np.diff(np.append(start, len(k))) == 1]]
    return np.stack([once // (w * w), (once // w) % w, once % w], axis=1).astype(f.dtype)


def get_boundary_vertices(t, dim=2):
    if dim == 2:
        b = get_boundary_edges(t)
    elif dim == 3:
        b = _boundary_facets_any_order(t)
    else:
        raise ValueError("Dimension not supported.")
    return np.unique(b.reshape(-1))


def vertex_to_entities(p, t, dim=2):
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
    bv = get_boundary_vertices(t, dim=dim)
    vtoe, ptr = vertex_to_entities(p, t, dim=dim)
    parts = [vtoe[ptr[v]:ptr[v + 1]] for v in bv]
    bele = np.concatenate(parts) if parts else np.array([], dtype=int)
    return np.unique(bele)


def delete_boundary_entities(p, t, dim=2, min_qual=0.10, verbose=1):
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
    """Move interior vertices to the solution of the graph Laplace equation
    with boundary vertices fixed (one direct sparse solve per coordinate)."""
    import scipy.sparse as sp
    from scipy.sparse.linalg import splu

    if p.ndim != 2:
        raise NotImplementedError("Laplacian smoothing only works in 2D for now")
    n = len(p)
    nds = t.T
    local = np.array([[1, 2], [2, 0], [0, 1]]).T
    idx = nds[local]
    row = np.array([idx[0], idx[1], idx[0], idx[1]]).ravel()
    col = np.array([idx[0], idx[1], idx[1], idx[0]]).ravel()
    a = np.ones(idx.shape[1:], dtype=float)
    val = np.array([+a, +a, -a, -a]).ravel()
    m = sp.coo_matrix((val, (row, col)), shape=(n, n)).tocsr()
    bnd = get_boundary_vertices(t)
    for b in bnd:
        m.data[m.indptr[b]:m.indptr[b + 1]] = 0.0
    d = m.diagonal()
    d[bnd] = 1.0
    m.setdiag(d)
    rhs = np.zeros((n, 2))
    rhs[bnd] = p[bnd]
    lu = splu(m.tocsc())
    return np.column_stack([lu.solve(rhs[:, 0]), lu.solve(rhs[:, 1])]), t


def is_manifold(p, t, dim=2):
    """geometry/utils.py:634-654: a boundary edge belongs to one entity and
    boundary vertices and edges pair up."""
    bedges = get_boundary_edges(t, dim=dim)
    if bedges.size != p[np.unique(bedges), :].size:
        print("Mesh has a non-manifold boundary...", flush=True)
        return False
    return True


def _det4(m):
    """Determinants of a stack of 4x4 matrices by cofactor expansion along
    the last column (of ones for the point-in-cell test)."""
    def d3(a):
        return (a[:, 0, 0] * (a[:, 1, 1] * a[:, 2, 2] - a[:, 1, 2] * a[:, 2, 1])
                - a[:, 0, 1] * (a[:, 1, 0] * a[:, 2, 2] - a[:, 1, 2] * a[:, 2, 0])
                + a[:, 0, 2] * (a[:, 1, 0] * a[:, 2, 1] - a[:, 1, 1] * a[:, 2, 0]))
    out = 0.0
    for r in range(4):
        rows = [k for k in range(4) if k != r]
        minor = m[:, rows][:, :, :3]
        out = out + ((-1) ** (r + 3)) * m[:, r, 3] * d3(minor)
    return out


def _inside(cent, q, dim):
    """Point-in-entity tests of geometry/utils.py:657-726 for stacks of
    (point, entity) pairs: barycentric coordinates in [0, 1] (2D), equal
    orientation signs of the five 4x4 determinants (3D)."""
    if dim == 2:
        x, y = cent[:, 0], cent[:, 1]
        x1, y1, x2, y2, x3, y3 = (q[:, 0, 0], q[:, 0, 1], q[:, 1, 0], q[:, 1, 1],
                                  q[:, 2, 0], q[:, 2, 1])
        den = (y2 - y3) * (x1 - x3) + (x3 - x2) * (y1 - y3)
        with np.errstate(divide="ignore", invalid="ignore"):
            a = ((y2 - y3) * (x - x3) + (x3 - x2) * (y - y3)) / den
            b = ((y3 - y1) * (x - x3) + (x1 - x3) * (y - y3)) / den
        c = 1 - a - b
        return (0 <= a) & (a <= 1) & (0 <= b) & (b <= 1) & (0 <= c) & (c <= 1)
    ones = np.ones((len(q), 4, 1))
    base = np.concatenate((q, ones), axis=2)
    signs = [np.sign(_det4(base))]
    for k in range(4):
        m = base.copy()
        m[:, k, :3] = cent
        signs.append(np.sign(_det4(m)))
    s = np.stack(signs, 1)
    return (s == -1).all(1) | (s == 1).all(1)


def do_any_overlap(p, t, dim=2):
    """geometry/utils.py:745-817: pairs (ie, ele) where entity ie's centroid
    lies in an entity ele of its vertex 1-ring; progress and alerts printed
    as the reference does."""
    vtoe, ptr = vertex_to_entities(p, t, dim=dim)
    cents = get_centroids(p, t)
    C = len(t)
    k = dim + 1
    vflat = t.reshape(-1)
    cnt = ptr[vflat + 1] - ptr[vflat]
    start = ptr[vflat]
    ie = np.repeat(np.repeat(np.arange(C), k), cnt)
    off = np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt)
    ele = vtoe[np.repeat(start, cnt) + off]
    keep = ele != ie
    ie, ele = ie[keep], ele[keep]
    hit = _inside(cents[ie], p[t[ele]], dim) if len(ie) else np.zeros(0, bool)
    inter = []
    hi, he = ie[hit], ele[hit]
    j = 0
    for blk in range(0, C, 1000):
        print("INFO: " + str(100.0 * (blk / C)) + " % done...", flush=True)
        while j < len(hi) and hi[j] < blk + 1000:
            print("Alert: entity " + str(hi[j]) + " intersects with entity " + str(he[j])
                  + ". These will be adjusted.", flush=True)
            inter.append((int(hi[j]), int(he[j])))
            j += 1
    return inter


def linter(p, t, dim=2, min_qual=0.10):
    """geometry/utils.py:820-872: delete the lower-quality entity of every
    overlapping pair, clean up, drop poor boundary entities (2D) and report
    the final size and minimum quality."""
    print("Performing mesh linting...", flush=True)
    qual = simp_qual(p, t)
    inter = do_any_overlap(p, t, dim=dim)
    delete = np.array([], dtype=int)
    for pair in inter:
        ix = list(pair)
        delete = np.append(delete, pair[int(np.argmin(qual[ix]))])
    delete = np.unique(delete)
    print("Deleting " + str(len(delete)) + " overlapped entities", flush=True)
    t = np.delete(t, delete, axis=0)
    p, t, _ = fix_mesh(p, t, delete_unused=True)
    if dim == 2:
        p, t = delete_boundary_entities(p, t, min_qual=min_qual)
        _ = is_manifold(p, t)
    qual = simp_qual(p, t)
    print("There are " + str(len(p)) + " vertices and " + str(len(t))
          + " elements in the mesh", flush=True)
    print("The minimum element quality is " + str(np.amin(qual)), flush=True)
    return p, t


def dihedral_angles(p, t):
    """(ncells, 6) dihedral angles in radians of tetrahedra t of points p,
    computed on the GPU (sm_dihedral_angles, the reference's
    fast_geometry.cpp:351-452 edge table)."""
    import torch

    from . import _lib

    _lib.require_gpu()
    t = np.ascontiguousarray(t, dtype=np.int32)
    n = len(t)
    if n == 0:
        return np.zeros((0, 6))
    pt = torch.from_numpy(np.ascontiguousarray(p, dtype=np.float64)).cuda()
    tt = torch.from_numpy(t).cuda()
    ang = torch.empty(6 * n, dtype=torch.float64, device=pt.device)
    stream = torch.cuda.current_stream().cuda_stream
    _lib.check(_lib.lib.sm_dihedral_angles(pt.data_ptr(), tt.data_ptr(), n, ang.data_ptr(),
                                           stream), "sm_dihedral_angles")
    return ang.cpu().numpy().reshape(n, 6)
