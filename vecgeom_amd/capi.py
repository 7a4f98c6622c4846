"""ctypes binding of libvecgeom_amd.so (the C-ABI in include/vecgeom_amd.h).

The library is loaded from vecgeom_amd/lib/.  There is no CPU fallback: if
the library is missing, or a call fails, a RuntimeError is raised.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libvecgeom_amd.so")

VG_OK = 0
VG_ERR_INVALID = -1
VG_ERR_HIP = -2
VG_ERR_STATE = -3
VG_ERR_UNSUPPORTED = -4
VG_ERR_CAPACITY = -5

c_i32 = ctypes.c_int32
c_sz = ctypes.c_size_t
c_vp = ctypes.c_void_p


class FlatGeometry(ctypes.Structure):
    _fields_ = [
        ("n_shapes", c_i32),
        ("shape_type", c_vp),
        ("shape_param_begin", c_vp),
        ("params", c_vp),
        ("n_logical", c_i32),
        ("logical_shape", c_vp),
        ("logical_daughter_begin", c_vp),
        ("daughters", c_vp),
        ("n_placed", c_i32),
        ("placed_logical", c_vp),
        ("placed_transform", c_vp),
        ("world", c_i32),
    ]


# name -> argtypes (all return int status unless listed in _RESTYPE)
_SIGS = {
    "vg_last_error": [],
    "vg_version": [],
    "vg_init": [ctypes.c_int],
    "vg_shutdown": [],
    "vg_shape_supported": [c_i32],
    "vg_set_kernel_variant": [ctypes.c_int],
    "vg_geometry_upload": [ctypes.POINTER(FlatGeometry), ctypes.POINTER(c_vp)],
    "vg_geometry_free": [c_vp],
    "vg_geometry_max_depth": [c_vp, ctypes.POINTER(c_i32)],
    "vg_geometry_placed_extent": [c_vp, c_i32, c_vp, c_vp],
    "vg_geometry_placed_aabb": [c_vp, c_i32, c_vp, c_vp],
    "vg_contains": [c_vp, c_i32, c_vp, c_vp, c_vp, c_sz, c_vp, c_vp],
    "vg_inside": [c_vp, c_i32, c_vp, c_vp, c_vp, c_sz, c_vp, c_vp],
    "vg_distance_to_in": [c_vp, c_i32] + [c_vp] * 7 + [c_sz, c_vp, c_vp],
    "vg_distance_to_in_minimize": [c_vp, c_i32, c_i32] + [c_vp] * 6 + [c_sz, c_vp, c_vp, c_vp],
    "vg_distance_to_out": [c_vp, c_i32] + [c_vp] * 7 + [c_sz, c_vp, c_vp, c_vp],
    "vg_placed_distance_to_out": [c_vp, c_i32] + [c_vp] * 7 + [c_sz, c_vp, c_vp],
    "vg_safety_to_in": [c_vp, c_i32, c_vp, c_vp, c_vp, c_sz, c_vp, c_vp],
    "vg_safety_to_in_minimize": [c_vp, c_i32, c_vp, c_vp, c_vp, c_sz, c_vp, c_vp],
    "vg_safety_to_out": [c_vp, c_i32, c_vp, c_vp, c_vp, c_sz, c_vp, c_vp],
    "vg_safety_to_out_minimize": [c_vp, c_i32, c_vp, c_vp, c_vp, c_sz, c_vp, c_vp],
    "vg_contains_local": [c_vp, c_i32, c_vp, c_vp, c_vp, c_sz, c_vp, c_vp, c_vp, c_vp, c_vp],
    "vg_unplaced_contains": [c_vp, c_i32, c_vp, c_vp, c_vp, c_sz, c_vp, c_vp],
    "vg_normal": [c_vp, c_i32, c_vp, c_vp, c_vp, c_sz, c_vp, c_vp, c_vp, c_vp, c_vp],
    "vg_distance_to_out_normal": [c_vp, c_i32] + [c_vp] * 7 + [c_sz] + [c_vp] * 7,
    "vg_scalar_query": [c_vp, c_i32, c_i32, c_vp, c_vp, ctypes.c_double, c_vp, c_vp],
}
_RESTYPE = {"vg_last_error": ctypes.c_char_p, "vg_version": ctypes.c_char_p}
# Entry points added by later layers register here (navigator, batching,
# serving) via register().
_lib = None


def register(sigs, restypes=None):
    """Adds signatures for entry points declared elsewhere in the package."""
    _SIGS.update(sigs)
    if restypes:
        _RESTYPE.update(restypes)
    if _lib is not None:
        _bind(_lib, sigs)


def _bind(lib, sigs):
    for name, args in sigs.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPE.get(name, ctypes.c_int)


def lib():
    """Loads libvecgeom_amd.so once; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        _bind(L, _SIGS)
        _lib = L
    return _lib


def call(name, *args):
    """Calls an entry point and raises on a non-zero status."""
    rc = getattr(lib(), name)(*args)
    if rc != VG_OK:
        msg = lib().vg_last_error().decode(errors="replace")
        raise RuntimeError(f"{name} failed ({rc}): {msg}")
    return rc


def status(name, *args):
    """Calls an entry point and returns (status, message) without raising."""
    rc = getattr(lib(), name)(*args)
    msg = lib().vg_last_error().decode(errors="replace") if rc != VG_OK else ""
    return rc, msg


def exported_symbols():
    """Names declared in include/vecgeom_amd*.h that the library must export."""
    return sorted(_SIGS)
