"""Oracle restatement of the reference's sizing layer (FP64 numpy).

Test infrastructure only (tests/, the smoke check and the bench's CPU
baseline import it; the product never does).

Cited reference files:
  SeismicMesh/sizing/size_function.py:1-12        (SizeFunction)
  SeismicMesh/sizing/mesh_size_function.py:391-523 (interpolant, wavelength
                                                    sizing, gradation call)
  SeismicMesh/sizing/cpp/FastHJ.cpp:27-157         (limgrad active set)
"""
import math

import numpy as np
from scipy.interpolate import RegularGridInterpolator


class SizeFunction:
    # size_function.py:1-12
    def __init__(self, bbox, cell_size, hmin):
        if not isinstance(bbox, tuple):
            raise ValueError("`bbox` must be a tuple")
        if not callable(cell_size):
            raise ValueError("`cell_size` must be callable")
        self.bbox = bbox
        self.cell_size = cell_size
        self.hmin = hmin

    def eval(self, x):
        return self.cell_size(x)


def grid_vectors(bbox, shape):
    """mesh_size_function.py:514-523: float32 linspace axes (z, x[, y])."""
    return tuple(
        np.linspace(bbox[2 * k], bbox[2 * k + 1], n, dtype=np.float32)
        for k, n in enumerate(shape)
    )


def build_interpolant(cell_size, bbox):
    """mesh_size_function.py:391-408: RegularGridInterpolator, linear,
    bounds_error=False, fill_value=None (linear extrapolation outside)."""
    vecs = grid_vectors(bbox, cell_size.shape)
    return RegularGridInterpolator(vecs, cell_size, bounds_error=False, fill_value=None)


def wavelength_sizing(vp, wl, freq):
    # mesh_size_function.py:411-425
    return vp / (freq * wl)


def limgrad(dims, elen, dfdx, imax, ffun):
    """FastHJ.cpp:63-157, restated with 0-based indices.

    ``ffun`` is the field flattened in Fortran order with shape ``dims``
    (nz, nx, ny); 2D fields pass ny=1.  Active nodes of pass ``it`` are those
    whose value changed in pass ``it-1`` (all nodes at pass 0), visited in
    increasing linear index; each visits its 6-neighbour stencil (clamped at
    the grid edges) in the order +x, -x, +z, -z, +y, -y and lowers the larger
    value of each pair to the smaller + elen*dfdx when it exceeds it by more
    than ftol = min(f)*sqrt(1e-9).  Pure Python: small grids only.
    """
    n0, n1, n2 = (int(d) for d in dims)
    f = np.array(ffun, dtype=float).ravel().copy()
    aset = np.full(f.size, -1, dtype=np.int64)
    ftol = f.min() * math.sqrt(1e-9)
    step = elen * dfdx
    for it in range(imax):
        active = np.nonzero(aset == it - 1)[0]
        if active.size == 0:
            break
        for node in active:
            k, rem = divmod(int(node), n0 * n1)
            j, i = divmod(rem, n0)
            nbrs = (
                i + n0 * min(j + 1, n1 - 1) + n0 * n1 * k,
                i + n0 * max(j - 1, 0) + n0 * n1 * k,
                min(i + 1, n0 - 1) + n0 * j + n0 * n1 * k,
                max(i - 1, 0) + n0 * j + n0 * n1 * k,
                i + n0 * j + n0 * n1 * min(k + 1, n2 - 1),
                i + n0 * j + n0 * n1 * max(k - 1, 0),
            )
            a = int(node)
            for b in nbrs:
                if f[a] > f[b]:
                    v = f[b] + step
                    if f[a] > v + ftol:
                        f[a] = v
                        aset[a] = it
                else:
                    v = f[a] + step
                    if f[b] > v + ftol:
                        f[b] = v
                        aset[b] = it
    return f


def gradation(cell_size, grade, elen, limiter=limgrad):
    """mesh_size_function.py:469-494 (calls limgrad on the F-ordered field)."""
    sz = cell_size.shape
    dims = (sz[0], sz[1], 1) if len(sz) == 2 else sz
    tmp = limiter(dims, elen, grade, 10000, cell_size.flatten("F"))
    return np.reshape(tmp, sz, "F")


def sizing_from_velocity(vp, bbox, hmin, hmax=10000.0, wl=0, freq=2.0, grade=0.0,
                         limiter=limgrad):
    """The parts of get_sizing_function_from_segy (mesh_size_function.py:34-239)
    that the synthetic workloads use: wavelength sizing, hmin/hmax clamp,
    gradation with elen=(bbox[1]-bbox[0])/nz, no pad, no Courant limit."""
    vp = np.asarray(vp)
    cell = np.full(vp.shape, hmin, dtype=float) if wl <= 0 else wavelength_sizing(vp, wl, freq)
    cell = np.array(cell, dtype=float)
    cell[cell < hmin] = hmin
    cell[cell > hmax] = hmax
    if grade > 0:
        cell = gradation(cell, grade, (bbox[1] - bbox[0]) / vp.shape[0], limiter)
    fh = build_interpolant(cell, bbox)
    return SizeFunction(tuple(float(b) for b in bbox), fh, hmin), cell


# ---------------------------------------------------------------------------
# velocity files, gradient sizing and domain pads
# (mesh_size_function.py:428-448, :526-587, :630-643)
# ---------------------------------------------------------------------------
def read_segy(filename):
    """SEG-Y (format 1, IBM float) as segyio reads it with
    ignore_geometry=True (mesh_size_function.py:630-643): traces are the
    columns, then the grid is flipped upside down.  Word-by-word decoding."""
    import struct

    with open(filename, "rb") as f:
        data = f.read()
    ns = struct.unpack(">H", data[3220:3222])[0]
    fmt = struct.unpack(">h", data[3224:3226])[0]
    assert fmt == 1, "the oracle reads IBM-float SEG-Y only"
    ntr = (len(data) - 3600) // (240 + 4 * ns)
    vp = np.zeros((ns, ntr))
    for j in range(ntr):
        base = 3600 + j * (240 + 4 * ns) + 240
        for i in range(ns):
            (w,) = struct.unpack(">I", data[base + 4 * i:base + 4 * i + 4])
            sign = -1.0 if w >> 31 else 1.0
            expo = ((w >> 24) & 0x7F) - 64
            frac = (w & 0xFFFFFF) / 16777216.0
            vp[i, j] = float(np.float32(sign * frac * 16.0 ** expo))
    return np.flipud(vp), ns, ntr, 0


def gradient_sizing(vp, grad, stencil_size):
    """mesh_size_function.py:428-448."""
    from scipy import ndimage

    if grad == 0.0:
        return 99999
    window = [stencil_size] * vp.ndim
    mean = ndimage.uniform_filter(vp, tuple(window))
    sqr = ndimage.uniform_filter(vp ** 2, tuple(window))
    var = sqr - mean ** 2
    var = np.divide(var, np.amax(var))
    var -= np.amin(var)
    return grad / (var + 0.10)


def domain_pad(cell, bbox, pad, style):
    """mesh_size_function.py:526-587 (2D): pad nz rows at z min and nx
    columns on both x sides; max value for the constant / ramp ends."""
    nz, nx = cell.shape
    dz = (bbox[1] - bbox[0]) / nz
    dx = (bbox[3] - bbox[2]) / nx
    nnz, nnx = int(pad / dz), int(pad / dx)
    bbox = (bbox[0] - pad, bbox[1], bbox[2] - pad, bbox[3] + pad)
    widths = ((nnz, 0), (nnx, nnx))
    top = np.amax(cell)
    if style == "edge":
        cell = np.pad(cell, widths, "edge")
    elif style == "constant":
        cell = np.pad(cell, widths, "constant", constant_values=(top, top))
    else:
        cell = np.pad(cell, widths, "linear_ramp", end_values=(top, top))
    return cell, bbox


def sizing_from_segy(filename, bbox, hmin, hmax, wl, freq, grad, grade, stencil_size=10.0,
                     domain_pad_m=0.0, pad_style="edge", limiter=limgrad):
    """get_sizing_function_from_segy for an IBM-float SEG-Y file
    (mesh_size_function.py:34-239, without Courant limiting)."""
    vp, nz, nx, _ = read_segy(filename)
    cell = np.minimum(vp / (freq * wl) if wl > 0 else 99999,
                      gradient_sizing(vp, grad, stencil_size))
    cell = np.array(cell, dtype=float)
    cell[cell < hmin] = hmin
    cell[cell > hmax] = hmax
    if grade > 0:
        cell = gradation(cell, grade, (bbox[1] - bbox[0]) / nz, limiter)
    if domain_pad_m > 0:
        cell, bbox = domain_pad(cell, bbox, domain_pad_m, pad_style)
    return SizeFunction(tuple(float(b) for b in bbox), build_interpolant(cell, bbox), hmin), cell
