"""MI355X DEM timestep engine for the LIGGGHTS-INL sphere hot path.

The compute path is the HIP library ``libliggghts_dem.so`` built from
``csrc/`` (C-ABI in ``include/liggghts_dem.h``).  This module is a thin
ctypes front end used by the tests and the benchmark; it has no fallback:
if the library is missing or no GPU is visible, it raises.

The directory name contains a hyphen, so import it by path, e.g.::

    import importlib.util, os
    spec = importlib.util.spec_from_file_location(
        "liggghts_inl_amd", os.path.join(root, "liggghts-inl_amd", "__init__.py"))

The gradient front end (``liggghts_inl_amd.adjoint``, a torch.autograd
Function) is imported on its own, so importing the package does not pull in
torch.
"""
from .engine import (  # noqa: F401
    DEM_FP64,
    DEM_MIXED,
    DemError,
    Simulation,
    build,
    library,
    library_path,
    material_tables,
    mesh_topology,
)
from . import datafile  # noqa: F401
from . import meshio  # noqa: F401
from . import workloads  # noqa: F401
