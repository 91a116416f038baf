"""Sizing functions with the reference's API; grids are interpolated on the GPU.

Mirrors SeismicMesh/sizing/size_function.py:1-12 (SizeFunction) and the
in-memory part of SeismicMesh/sizing/mesh_size_function.py:34-239
(get_sizing_function_from_segy with ``velocity_data``).  The graded size
grid is held as FP32 values on float32-linspace axes (z, x[, y]) exactly as
the reference's RegularGridInterpolator is built (:391-408, :514-523);
queries are linear with linear extrapolation, evaluated on the device
(csrc/sizing_eval.h).  Gradient limiting (FastHJ.cpp) runs on the device
as a separable distance transform (csrc/api.hip, sm_limgrad).
"""
import ctypes

import numpy as np

from . import _lib

__all__ = ["SizeFunction", "GridSizing", "get_sizing_function_from_segy", "limgrad"]


class SizeFunction:
    """size_function.py:1-12"""

    def __init__(self, bbox, cell_size, hmin):
        if not isinstance(bbox, tuple):
            raise ValueError("`bbox` must be a tuple")
        self.bbox = bbox
        if not callable(cell_size):
            raise ValueError("`cell_size` must be callable")
        self.cell_size = cell_size
        self.hmin = hmin

    def eval(self, x):
        return self.cell_size(x)


class GridSizing:
    """Callable gridded sizing: linear interpolation of FP32 values on the
    float32 linspace axes of ``bbox`` (z, x[, y]); evaluated on the GPU."""

    def __init__(self, values, bbox):
        values = np.asarray(values)
        self.dim = values.ndim
        if self.dim not in (2, 3):
            raise ValueError("Dimension not supported")
        self.values = np.ascontiguousarray(values, dtype=np.float32)
        self.bbox = tuple(float(b) for b in bbox)
        self.axes = tuple(
            np.linspace(self.bbox[2 * k], self.bbox[2 * k + 1], n, dtype=np.float32).astype(np.float64)
            for k, n in enumerate(self.values.shape))

    def grid_struct(self, device_ptrs=None):
        """sm_grid with host (default) or given device pointers."""
        g = _lib.Grid()
        g.dim = self.dim
        for k in range(self.dim):
            g.n[k] = self.values.shape[k]
        if device_ptrs is None:
            for k in range(self.dim):
                g.axis[k] = self.axes[k].ctypes.data
            g.values = self.values.ctypes.data
        else:
            for k in range(self.dim):
                g.axis[k] = device_ptrs[k]
            g.values = device_ptrs[3]
        return g

    def __call__(self, x):
        import torch

        _lib.require_gpu()
        x = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1, self.dim))
        dev = torch.device("cuda", torch.cuda.current_device())
        axes = [torch.from_numpy(a).to(dev) for a in self.axes]
        vals = torch.from_numpy(self.values).to(dev)
        xt = torch.from_numpy(x).to(dev)
        h = torch.empty(len(x), dtype=torch.float32, device=dev)
        ptrs = [a.data_ptr() for a in axes] + [None] * (3 - self.dim) + [vals.data_ptr()]
        g = self.grid_struct(ptrs)
        stream = torch.cuda.current_stream().cuda_stream
        if len(x):
            _lib.check(_lib.lib.sm_sizing_eval(ctypes.byref(g), ctypes.c_void_p(xt.data_ptr()),
                                               len(x), ctypes.c_void_p(h.data_ptr()),
                                               ctypes.c_void_p(stream)), "sm_sizing_eval")
        return h.cpu().numpy().astype(np.float64)


def limgrad(cell_size, elen, grade):
    """Gradient-limit a 2D/3D size grid on the GPU (FastHJ.cpp:63-157 fixed
    point: neighbouring values differ by at most elen*grade)."""
    import torch

    _lib.require_gpu()
    sz = cell_size.shape
    dims = (sz[0], sz[1], 1) if len(sz) == 2 else tuple(sz)
    f = np.asarray(cell_size, dtype=np.float64).flatten("F")
    dev = torch.device("cuda", torch.cuda.current_device())
    ft = torch.from_numpy(f).to(dev)
    stream = torch.cuda.current_stream().cuda_stream
    _lib.check(_lib.lib.sm_limgrad(ctypes.c_void_p(ft.data_ptr()), dims[0], dims[1], dims[2],
                                   float(elen), float(grade), ctypes.c_void_p(stream)),
               "sm_limgrad")
    return np.reshape(ft.cpu().numpy(), sz, "F")


_SZ_KEYS = {
    "hmin", "hmax", "wl", "freq", "cr_max", "dt", "space_order", "grad", "grade",
    "stencil_size", "pad_style", "domain_pad", "units", "nz", "nx", "ny", "byte_order",
    "axes_order", "axes_order_sort", "dtype", "velocity_data", "vp_water",
}


def get_sizing_function_from_segy(filename, bbox, comm=None, **kwargs):
    """Velocity model -> graded size grid -> SizeFunction
    (mesh_size_function.py:34-239), for in-memory ``velocity_data``.

    Reading SEG-Y / binary files is not part of this package: pass the
    velocity array as ``velocity_data``.
    """
    opts = {
        "velocity_data": None, "vp_water": 1500.0, "hmin": 150.0, "hmax": 10000.0,
        "wl": 0, "freq": 2.0, "grad": 0.0, "grade": 0.0, "stencil_size": 10.0,
        "space_order": 1, "dt": 0.0, "cr_max": 1.0, "pad_style": "edge", "domain_pad": 0.0,
        "units": "m-s", "nz": None, "nx": None, "ny": None,
    }
    for key in kwargs:
        if key not in _SZ_KEYS:
            raise ValueError("Option %s with parameter %s not recognized " % (key, kwargs[key]))
    opts.update(kwargs)
    vp = opts["velocity_data"]
    if vp is None:
        raise NotImplementedError(
            "reading velocity files is not supported; pass velocity_data=<array>")
    vp = np.array(vp)
    if opts["units"] == "km-s":
        print("Converting from km-s to m-s...", flush=True)
        vp = vp * 1000.0
    elif opts["units"] == "ft-s":
        print("Converting from ft-s to m-s...", flush=True)
        vp = vp * 0.30
    pos = np.where(vp < 1e-3)
    if len(pos) > 0 and any(pos[0]):
        if opts["vp_water"] is None or opts["vp_water"] < 1300 or opts["vp_water"] > 1800:
            raise ValueError("vp_water is None or out of bounds. It should be >1300 and <1800 m/s")
        vp[pos] = opts["vp_water"]
    if len(bbox) == 4:
        dim = 2
    elif len(bbox) == 6:
        dim = 3
    else:
        raise ValueError("Dimension not supported")
    cell = np.full(vp.shape, opts["hmin"], dtype=float)
    if opts["wl"] > 0 or opts["grad"] > 0:
        wl_size = 99999 if opts["wl"] == 0.0 else _wavelength_sizing(vp, opts["wl"], opts["freq"])
        if opts["grad"] > 0:
            raise NotImplementedError("gradient-based sizing (grad > 0) is not supported")
        cell = np.minimum(wl_size, 99999)
    cell = np.array(cell)
    print(f"Enforcing minimum edge length of {opts['hmin']}")
    cell[cell < opts["hmin"]] = opts["hmin"]
    print(f"Enforcing maximum edge length of {opts['hmax']}")
    cell[cell > opts["hmax"]] = opts["hmax"]
    cell = _enforce_courant(vp, cell, opts["cr_max"], opts["dt"], opts["space_order"])
    if opts["grade"] < 0:
        raise ValueError("Parameter `grade` must be > 0.0")
    if opts["grade"] > 0:
        print(f"Enforcing mesh size gradation of {opts['grade']} decimal percent...")
        cell = limgrad(cell, (bbox[1] - bbox[0]) / vp.shape[0], opts["grade"])
    if opts["domain_pad"] > 0:
        raise NotImplementedError("domain padding is not supported")
    fh = GridSizing(cell, bbox)
    del dim
    return SizeFunction(tuple(bbox), fh, opts["hmin"])


def _wavelength_sizing(vp, wl, freq):
    # mesh_size_function.py:411-425
    if wl < 0:
        raise ValueError("Parameter `wl` must be set > 0")
    if freq < 0.0:
        raise ValueError("Parameter `freq` must be set > 0.0")
    print(f"Mesh sizes will be built to resolve an estimate of wavelength of a {freq}  hz "
          f"wavelet with {wl} vertices", flush=True)
    return vp / (freq * wl)


def _enforce_courant(vp, cell, cr_max, dt, space_order):
    # mesh_size_function.py:451-466
    if cr_max == 0.0 or dt == 0.0 or space_order == 0.0:
        return cell
    print(f"Enforcing timestep of {dt} seconds...", flush=True)
    if cr_max < 0:
        raise ValueError("Parameter `cr_max` must be > 0.0")
    if dt < 0:
        raise ValueError("Parameter `dt` must be > 0.0")
    if space_order < 1:
        raise ValueError("Parameter `space_order` must be >= 1 ")
    d = vp.ndim
    cr_old = (vp * dt) / (d * cell)
    crm = cr_max / (d * space_order)
    dxn = (vp * dt) / (d * crm)
    return np.where(cr_old > crm, dxn, cell)
