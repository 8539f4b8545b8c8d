"""relaxmap_amd: RelaxMap's hot path (steady-state flow, map-equation move
sweeps, module contraction) as hand-written HIP kernels for MI355X (gfx950).

The product is the C ABI in include/relaxmap.h (librelaxmap.so) and the
``relaxmap`` command line; this package is a thin ctypes view of that ABI
whose names follow the reference's ``Network`` class (Module.h:99-200).
"""
from .core import Network, NetworkFile, PartitionResult, device_count, library_path, load_library  # noqa: F401

__all__ = ["Network", "NetworkFile", "PartitionResult", "device_count", "library_path", "load_library"]
