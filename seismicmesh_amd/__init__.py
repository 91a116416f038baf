"""seismicmesh_amd: the SeismicMesh DistMesh force loop on MI355X (gfx950).

Drop-in for the reference's generate_mesh / sliver_removal path: the same
geometry classes, SizeFunction and keyword arguments, with the force loop
running in hand-written HIP kernels behind a C ABI (include/smgpu.h,
libsmgpu.so loaded with ctypes).
"""
from . import _lib  # noqa: F401  (fails loudly when libsmgpu.so is missing)
from . import geometry, mesh_utils, sizing
from .generation import generate_mesh, sliver_removal
from .geometry import (Ball, Cube, Cylinder, Difference, Disk, Intersection, Prism,
                       Rectangle, Repeat, Torus, Union)
from .sizing import (SizeFunction, get_sizing_function_from_segy, read_velocity_model,
                     write_velocity_model)

__version__ = "0.1.0"

__all__ = [
    "__version__", "geometry", "mesh_utils", "sizing", "generate_mesh", "sliver_removal",
    "Ball", "Cube", "Cylinder", "Difference", "Disk", "Intersection", "Prism", "Rectangle",
    "Repeat", "Torus", "Union", "SizeFunction", "get_sizing_function_from_segy",
    "read_velocity_model", "write_velocity_model",
]
