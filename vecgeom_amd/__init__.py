"""vecgeom_amd -- MI355X-native (gfx950) VecGeom shape-query and navigation
hot path.

The compute path is libvecgeom_amd.so (hand-written HIP kernels behind the
C-ABI in include/vecgeom_amd.h).  This Python package is host-side plumbing:
ctypes bindings (capi), a geometry builder mirroring LogicalVolume /
PlaceDaughter / GeoManager (geometry), basket queries on torch device
tensors (queries), and synthetic workloads (workloads).  There is no CPU
fallback anywhere in the package.
"""
from . import capi, geometry, workloads  # noqa: F401

__all__ = ["capi", "geometry", "workloads"]
