"""Synthetic workloads (seeded) for benchmarks and tests.

* ``disk``: unit disk, uniform edge length 0.02 (the reference's plain
  DistMesh case; 7,879 initial vertices).
* ``bp2004``: a BP2004-shaped 2D velocity model on a 1911 x 5395 FP32 grid
  over z in [-12 km, 0], x in [0, 67 km] (6.25 m x 12.42 m cells): a
  1480 m/s water layer over the top 1 km, then 1500 + 0.55 (depth - 1000)
  m/s capped at 4000 m/s, plus 1-3 seeded elliptical salt bodies at
  4500 m/s.  Sizing v / (2 Hz * 10), clamped to hmin = 75 m, gradient
  limited at grade 0.15, on the Rectangle domain with its 4 corners fixed.
  ``scale`` < 1 shrinks the grid (same physical box) for quick runs.
"""
import numpy as np

from . import geometry as _geo
from .sizing import get_sizing_function_from_segy

BP_BBOX = (-12000.0, 0.0, 0.0, 67000.0)
BP_SHAPE = (1911, 5395)


def bp2004_velocity(shape=BP_SHAPE, bbox=BP_BBOX, seed=2004):
    nz, nx = shape
    z = np.linspace(bbox[0], bbox[1], nz)
    x = np.linspace(bbox[2], bbox[3], nx)
    depth = -z
    v = np.where(depth < 1000.0, 1480.0, np.minimum(1500.0 + 0.55 * (depth - 1000.0), 4000.0))
    vp = np.repeat(v[:, None], nx, axis=1).astype(np.float32)
    rng = np.random.default_rng(seed)
    nbody = int(rng.integers(1, 4))
    Z, X = np.meshgrid(depth, x, indexing="ij")
    for _ in range(nbody):
        cx = rng.uniform(8000.0, 59000.0)
        cz = rng.uniform(2500.0, 9000.0)
        ax = rng.uniform(3000.0, 9000.0)
        az = rng.uniform(800.0, 2500.0)
        inside = ((X - cx) / ax) ** 2 + ((Z - cz) / az) ** 2 <= 1.0
        vp[inside] = 4500.0
    return vp


def bp2004(scale=1.0, seed=2004, quiet=True):
    """Returns (domain, edge_length, info) for the BP2004-shaped workload."""
    shape = (max(16, int(round(BP_SHAPE[0] * scale))), max(16, int(round(BP_SHAPE[1] * scale))))
    vp = bp2004_velocity(shape, BP_BBOX, seed)
    kw = dict(velocity_data=vp, hmin=75.0, hmax=1e4, wl=10, freq=2.0, grade=0.15)
    if quiet:
        import contextlib
        import io
        with contextlib.redirect_stdout(io.StringIO()):
            ef = get_sizing_function_from_segy(None, BP_BBOX, **kw)
    else:
        ef = get_sizing_function_from_segy(None, BP_BBOX, **kw)
    domain = _geo.Rectangle(BP_BBOX)
    return domain, ef, {"grid": shape, "bbox": BP_BBOX, "hmin": 75.0, "vp": vp}


def disk(h=0.02):
    return _geo.Disk([0.0, 0.0], 1.0), h, {"bbox": (-1.0, 1.0, -1.0, 1.0)}


def batch_domain(i, seed0=1000, geom=None):
    """Domain i of the many-domains workload: a seeded CSG of 1-6 primitives
    in [-2, 2]^2 -- disks with r ~ U[0.2, 1] and axis-aligned rectangles with
    sides ~ U[0.3, 2] -- the first kept, each next one added (union) or cut
    out (difference).  ``geom`` is the module whose classes build it (this
    package's geometry by default; any module with the reference's
    geometry API)."""
    geo = geom or _geo
    rng = np.random.default_rng(seed0 + i)
    k = int(rng.integers(1, 7))
    prims = []
    for _ in range(k):
        if rng.random() < 0.5:
            r = rng.uniform(0.2, 1.0)
            c = rng.uniform(-2.0 + r, 2.0 - r, 2)
            prims.append(geo.Disk([float(c[0]), float(c[1])], float(r)))
        else:
            w, hgt = rng.uniform(0.3, 2.0, 2)
            c = rng.uniform([-2.0 + w / 2, -2.0 + hgt / 2], [2.0 - w / 2, 2.0 - hgt / 2])
            prims.append(geo.Rectangle((float(c[0] - w / 2), float(c[0] + w / 2),
                                        float(c[1] - hgt / 2), float(c[1] + hgt / 2))))
    ops = rng.random(k) < 0.65
    dom = prims[0]
    for prim, union in zip(prims[1:], ops[1:]):
        dom = geo.Union([dom, prim]) if union else geo.Difference([dom, prim])
    return dom, rng


def domain_area(dom, n=400):
    """Area of {sdf < 0} from an n x n sample of [-2, 2]^2."""
    ax = (np.arange(n) + 0.5) / n * 4.0 - 2.0
    X, Y = np.meshgrid(ax, ax, indexing="ij")
    d = dom.eval(np.column_stack((X.ravel(), Y.ravel())))
    return float((d < 0).mean() * 16.0)


def many_domains(n=512, seed0=1000, nmin=500, nmax=2.0e5, geom=None):
    """The many-domains workload: n seeded CSG domains with uniform sizes
    h = sqrt(area / N_i), N_i log-uniform in [nmin, nmax].  A domain whose
    cut-outs leave less than 0.1 of area keeps its first primitive only."""
    geo = geom or _geo
    domains, hs = [], []
    for i in range(n):
        dom, rng = batch_domain(i, seed0, geom)
        target = float(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        area = domain_area(dom)
        if area < 0.1:
            dom, rng2 = batch_domain(i, seed0, geom)
            while isinstance(dom, (geo.Union, geo.Difference)):
                dom = dom.domains[0]
            area = domain_area(dom)
        domains.append(dom)
        hs.append(float(np.sqrt(area / target)))
    return domains, hs


def csg3d(seed=3, nballs=48, nboxes=16, ncyl=8, ntori=4, geom=None):
    """The many-primitive 3D CSG: ((nballs balls U nboxes boxes) minus ncyl
    cylinders) U ntori tori, all inside [-1, 1]^3, seeded.  Balls have
    r ~ U[0.05, 0.3]; boxes sides ~ U[0.1, 0.5]; cylinders (axis z)
    r ~ U[0.05, 0.15], length ~ U[0.5, 1.5]; tori r1 ~ U[0.15, 0.35],
    r2 ~ U[0.03, 0.08].  Returns (domain, bbox); ``geom`` as in
    batch_domain."""
    geo = geom or _geo
    rng = np.random.default_rng(seed)
    solids = []
    for _ in range(nballs):
        r = rng.uniform(0.05, 0.3)
        c = rng.uniform(-1.0 + r, 1.0 - r, 3)
        solids.append(geo.Ball([float(v) for v in c], float(r)))
    for _ in range(nboxes):
        s = rng.uniform(0.1, 0.5, 3)
        c = rng.uniform(-1.0 + s / 2, 1.0 - s / 2)
        solids.append(geo.Cube(tuple(float(v) for k in range(3)
                                     for v in (c[k] - s[k] / 2, c[k] + s[k] / 2))))
    cuts = []
    for _ in range(ncyl):
        r = rng.uniform(0.05, 0.15)
        ln = rng.uniform(0.5, 1.5)
        c = rng.uniform(-1.0 + r, 1.0 - r, 3)
        c[2] = rng.uniform(-1.0 + ln / 2, 1.0 - ln / 2)
        cuts.append(geo.Cylinder(h=float(ln), r=float(r), translate=[float(v) for v in c]))
    tori = []
    for _ in range(ntori):
        r1 = rng.uniform(0.15, 0.35)
        r2 = rng.uniform(0.03, 0.08)
        c = rng.uniform(-1.0 + r1 + r2, 1.0 - r1 - r2, 3)
        tori.append(geo.Torus(float(r1), float(r2), translate=[float(v) for v in c]))
    body = geo.Union(solids) if len(solids) > 1 else solids[0]
    if cuts:
        body = geo.Difference([body] + cuts)
    if tori:
        body = geo.Union([body] + tori)
    bbox = (-1.0, 1.0, -1.0, 1.0, -1.0, 1.0)
    body.bbox = bbox
    return body, bbox


EAGE_BBOX = (-4200.0, 0.0, 0.0, 13520.0, 0.0, 13520.0)
EAGE_SHAPE = (210, 676, 676)


def eage3d_velocity(shape=EAGE_SHAPE, bbox=EAGE_BBOX, seed=4):
    """EAGE-shaped 3D velocity model on a (z, x, y) grid: layered
    1500 -> 4500 m/s with depth (linear over the box) plus one seeded
    ellipsoidal salt dome at 4480 m/s.  FP32, z-major."""
    nz, nx, ny = shape
    depth = -np.linspace(bbox[0], bbox[1], nz)
    x = np.linspace(bbox[2], bbox[3], nx)
    y = np.linspace(bbox[4], bbox[5], ny)
    zmax = bbox[1] - bbox[0]
    v = (1500.0 + 3000.0 * depth / zmax).astype(np.float32)
    vp = np.repeat(v[:, None], nx, axis=1)[:, :, None].repeat(ny, axis=2)
    rng = np.random.default_rng(seed)
    cx, cy = rng.uniform(0.3, 0.7, 2) * (bbox[3] - bbox[2])
    cz = rng.uniform(0.4, 0.7) * zmax
    ax, ay = rng.uniform(2000.0, 3500.0, 2)
    az = rng.uniform(800.0, 1500.0)
    dz = ((depth - cz) / az) ** 2
    dx = ((x - cx) / ax) ** 2
    dy = ((y - cy) / ay) ** 2
    inside = dz[:, None, None] + dx[None, :, None] + dy[None, None, :] <= 1.0
    vp[inside] = 4480.0
    return vp


def eage3d(scale=1.0, seed=4, quiet=True):
    """Returns (domain, edge_length, info) for the EAGE-shaped 3D workload:
    sizing v / (2 Hz * 8 cells per wavelength), hmin = 150 m, grade 0.15,
    on the Cube domain of the box.  ``scale`` < 1 coarsens the grid (same
    physical box and sizes)."""
    shape = tuple(max(8, int(round(s * scale))) for s in EAGE_SHAPE)
    vp = eage3d_velocity(shape, EAGE_BBOX, seed)
    kw = dict(velocity_data=vp, hmin=150.0, hmax=1e4, wl=8, freq=2.0, grade=0.15)
    if quiet:
        import contextlib
        import io
        with contextlib.redirect_stdout(io.StringIO()):
            ef = get_sizing_function_from_segy(None, EAGE_BBOX, **kw)
    else:
        ef = get_sizing_function_from_segy(None, EAGE_BBOX, **kw)
    return _geo.Cube(EAGE_BBOX), ef, {"grid": shape, "bbox": EAGE_BBOX, "hmin": 150.0}
