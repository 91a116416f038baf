"""ctypes binding of libsmgpu.so (the C ABI declared in include/smgpu.h).

The library is loaded from this package directory.  There is no fallback:
if it is missing, importing this module raises ImportError, and every entry
point that needs the GPU raises RuntimeError when no gfx950 device exists.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsmgpu.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        "libsmgpu.so is not built; run `python -m seismicmesh_amd._build` "
        "(hipcc, gfx950) before importing seismicmesh_amd"
    )

# One HIP runtime per process: torch ships its own libamdhip64 (same
# soname).  Loading torch first makes libsmgpu bind to that copy instead of
# loading a second runtime that would then keep torch from seeing the GPU.
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover - torch is only the buffer provider
    torch = None

lib = ctypes.CDLL(LIB_PATH)

SM_OK = 0
SM_ERR_NUMERIC = -5
# device scratch of sm_unique_edges_noalloc per GPU (SM_NOALLOC_SCRATCH_BYTES)
NOALLOC_SCRATCH_BYTES = 64 << 20

# SDF program opcodes and transform flags (smgpu.h)
OP_DISK, OP_RECT, OP_BALL, OP_CUBE, OP_TORUS, OP_PRISM, OP_CYLINDER = 1, 2, 3, 4, 5, 6, 7
OP_UNION, OP_INTERSECT, OP_DIFFERENCE, OP_SDIFF_STEP = 16, 17, 18, 19
OP_PUSH_MOD, OP_POP_POINT = 32, 33
XF_TRANSLATE, XF_ROT_X, XF_ROT_Y, XF_ROT_Z, XF_STRETCH = 1, 2, 4, 8, 16


class Node(ctypes.Structure):
    _fields_ = [
        ("op", ctypes.c_int32),
        ("nargs", ctypes.c_int32),
        ("flags", ctypes.c_int32),
        ("dim", ctypes.c_int32),
        ("prm", ctypes.c_double * 8),
        ("xf", ctypes.c_double * 34),
        ("pad", ctypes.c_double * 2),
    ]


assert ctypes.sizeof(Node) == 368


class Grid(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32),
        ("n", ctypes.c_int32 * 3),
        ("axis", ctypes.c_void_p * 3),
        ("values", ctypes.c_void_p),
        ("hconst", ctypes.c_double),
    ]


class MesherParams(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("nfix", ctypes.c_int32),
        ("nlevels", ctypes.c_int32),
        ("h0", ctypes.c_double),
        ("delta_t", ctypes.c_double),
        ("geps", ctypes.c_double),
        ("deps", ctypes.c_double),
        ("L0mult", ctypes.c_double),
        ("far_square", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


HOST_SIZING_FN = ctypes.CFUNCTYPE(
    ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int64,
    ctypes.POINTER(ctypes.c_double))

HOST_SDF_FN = ctypes.CFUNCTYPE(
    ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_int64,
    ctypes.POINTER(ctypes.c_double))


_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double
_flt = ctypes.c_float
_sz = ctypes.c_size_t

_PROTOS = {
    "sm_version": (_int, []),
    "sm_last_error": (_int, [ctypes.c_char_p, _sz]),
    "sm_device_count": (_int, [ctypes.POINTER(_int)]),
    "sm_sdf_eval": (_int, [_vp, _int, _int, _vp, _i64, _vp, _vp]),
    "sm_sizing_eval": (_int, [ctypes.POINTER(Grid), _vp, _i64, _vp, _vp]),
    "sm_unique_edges_workspace": (_sz, [_i64, _i64]),
    "sm_unique_edges": (_int, [_vp, _i64, _i64, _vp, _vp, _vp, _sz, _vp]),
    "sm_unique_edges_noalloc": (_int, [_vp, _i64, _i64, _vp, _vp, _vp]),
    "sm_cell_edges_workspace": (_sz, [_i64, _int, _i64]),
    "sm_cell_edges": (_int, [_vp, _i64, _int, _i64, _vp, _vp, _vp, _sz, _vp]),
    "sm_mesher_create": (_int, [ctypes.POINTER(MesherParams), _vp, _vp, ctypes.POINTER(Grid),
                                HOST_SIZING_FN, _vp, _int, ctypes.POINTER(_vp)]),
    "sm_mesher_create_batch": (_int, [ctypes.POINTER(MesherParams), _int, _vp, _vp, _vp, _vp,
                                      _vp, _int, ctypes.POINTER(_vp)]),
    "sm_mesher_set_host_sdf": (_int, [_vp, HOST_SDF_FN, _vp]),
    "sm_mesher_destroy": (_int, [_vp]),
    "sm_mesher_set_points": (_int, [_vp, _vp, _i64]),
    "sm_mesher_run": (_int, [_vp, _int, ctypes.POINTER(_dbl)]),
    "sm_mesher_triangulate": (_int, [_vp]),
    "sm_mesher_sizes": (_int, [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64),
                               ctypes.POINTER(_i64)]),
    "sm_mesher_get_points": (_int, [_vp, _vp, _i64]),
    "sm_mesher_get_cells": (_int, [_vp, _vp, _i64]),
    "sm_mesher_get_edges": (_int, [_vp, _vp, _i64]),
    "sm_mesher_get_triangulation": (_i64, [_vp, _vp, _vp, _i64]),
    "sm_mesher_stats": (_int, [_vp, _vp, _int]),
    "sm_mesher_set_slabs": (_int, [_vp, _int, _int, _int]),
    "sm_mesher_set_graph": (_int, [_vp, _int]),
    "sm_mesher_profile": (_int, [_vp, _int]),
    "sm_mesher_profile_read": (_int, [_vp, _vp, _vp, _int]),
    "sm_mesher_time_kernel": (_int, [_vp, _int, _int, ctypes.POINTER(_dbl)]),
    "sm_mesher_autotune": (_int, [_vp, _int, _vp, _vp, _int]),
    "sm_mesher_set_workgroups": (_int, [_vp, _vp, _int]),
    "sm_mesher_device_points": (_int, [_vp, ctypes.POINTER(_vp)]),
    "sm_mesher_stream": (_int, [_vp, ctypes.POINTER(_vp)]),
    "sm_mesher_find_slivers": (_int, [_vp, _dbl, _dbl, ctypes.POINTER(_i64)]),
    "sm_mesher_perturb_slivers": (_int, [_vp, _dbl]),
    "sm_mesher_set_surrogate": (_int, [_vp, _vp, _dbl]),
    "sm_mesher_set_surrogate_bands": (_int, [_vp, _int, _vp, _vp, _vp]),
    "sm_mesher_surrogate_stats": (_int, [_vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64),
                                         ctypes.POINTER(_dbl)]),
    "sm_mlp_eval": (_int, [_vp, _vp, _i64, _vp, _vp]),
    "sm_mlp_eval_centroids": (_int, [_vp, _vp, _i64, _vp, _vp, _vp]),
    "sm_gemm_f32": (_int, [_int, _int, _int, _flt, _vp, _i64, _i64, _vp, _i64, _i64, _flt,
                           _vp, _i64, _int, _vp, _vp]),
    "sm_mlp_bias_silu": (_int, [_i64, _int, _vp, _vp, _vp, _vp]),
    "sm_mlp_silu_backward": (_int, [_i64, _vp, _vp, _vp, _vp]),
    "sm_mlp_make_inputs": (_int, [_i64, _int, _vp, _vp, _vp, _flt, _vp, _vp]),
    "sm_mlp_sdf_loss": (_int, [_i64, _int, _vp, _vp, _flt, _flt, _flt, _vp, _vp, _vp]),
    "sm_mlp_adam": (_int, [_i64, _vp, _vp, _vp, _vp, _flt, _flt, _flt, _flt, _int, _vp]),
    "sm_f32_to_bf16": (_int, [_i64, _vp, _vp, _vp]),
    "sm_mlp_pack_hidden": (_int, [_int, _int, _vp, _vp, _vp]),
    "sm_host_delaunay2d": (_i64, [_vp, _i64, _vp, _vp, _i64, _vp]),
    "sm_host_delaunay3d": (_i64, [_vp, _i64, _vp, _i64, _vp]),
    "sm_host_delaunay3d_slabs": (_i64, [_vp, _i64, _int, _int, _int, _vp, _i64, _vp]),
    "sm_host_delaunay3d_check": (_int, [_vp, _i64, _vp, _i64, _int]),
    "sm_dihedral_angles": (_int, [_vp, _vp, _i64, _vp, _vp]),
    "sm_circumsphere_grad": (_int, [_vp, _vp, _vp, _vp, _i64, _vp, _vp]),
    "sm_limgrad": (_int, [_vp, _int, _int, _int, _dbl, _dbl, _vp]),
}

for _name, (_res, _args) in _PROTOS.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args

EXPORTED = tuple(_PROTOS)


def last_error():
    buf = ctypes.create_string_buffer(512)
    lib.sm_last_error(buf, 512)
    return buf.value.decode(errors="replace")


def check(rc, what):
    """Raise on a negative status from the library."""
    if rc < 0:
        msg = last_error()
        if rc == SM_ERR_NUMERIC:
            raise AssertionError(msg or "max movement indicates there's a convergence problem")
        raise RuntimeError("%s failed (%d): %s" % (what, rc, msg))
    return rc


def device_count():
    n = ctypes.c_int(0)
    check(lib.sm_device_count(ctypes.byref(n)), "sm_device_count")
    return n.value


def require_gpu():
    """The device path has no CPU fallback: fail loudly without a gfx950 GPU."""
    if device_count() < 1:
        raise RuntimeError("seismicmesh_amd needs an MI355X (gfx950) GPU; none was found")
