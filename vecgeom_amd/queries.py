"""Basket shape queries on device tensors, through the C-ABI.

Each function mirrors a VPlacedVolume SOA3D virtual
(volumes/PlacedVolume.h:147-250).  Points and directions are (x, y, z)
triples of 1-D float64 CUDA tensors (SOA3D); outputs are new CUDA tensors
unless `out` is given.  Work is enqueued on torch's current stream.
"""
import torch

from . import capi


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _check_soa(p, name, n=None, device=None):
    """An (x, y, z) triple of contiguous float64 CUDA tensors of one length
    (n, when given) on one device (device, when given); returns the length."""
    if len(p) != 3:
        raise ValueError(f"{name} must be an (x, y, z) triple")
    n = p[0].numel() if n is None else n
    device = p[0].device if device is None else device
    for a in p:
        if (not isinstance(a, torch.Tensor) or a.dtype != torch.float64 or not a.is_cuda or a.device != device
                or a.numel() != n or not a.is_contiguous()):
            raise ValueError(f"{name} components must be contiguous float64 CUDA tensors of length {n} on {device}")
    return n


def _check_arr(a, name, dtype, n, device):
    """One caller array the kernel reads or writes: a contiguous CUDA tensor
    of the given dtype with n entries on the points' device.  A host tensor
    or a shorter one would hand the kernel a pointer it must not touch."""
    if (not isinstance(a, torch.Tensor) or a.dtype != dtype or not a.is_cuda or a.device != device
            or a.numel() != n or not a.is_contiguous()):
        raise ValueError(f"{name} must be a contiguous {dtype} CUDA tensor of length {n} on {device}")
    return a


def _out(out, dtype, n, device):
    if out is None:
        return torch.empty(n, dtype=dtype, device=device)
    return _check_arr(out, "out", dtype, n, device)


def _step_max(step_max, n, device):
    return None if step_max is None else _check_arr(step_max, "step_max", torch.float64, n, device).data_ptr()


def _ptrs(p):
    return [a.data_ptr() for a in p]


def contains(geom, placed, p, out=None):
    n = _check_soa(p, "points")
    out = _out(out, torch.uint8, n, p[0].device)
    capi.call("vg_contains", geom.handle, int(placed), *_ptrs(p), n, out.data_ptr(), _stream())
    return out


def contains_local(geom, placed, p, out=None):
    """Contains(point, localPoint&) (volumes/PlacedVolume.h:153-156):
    returns (uint8 contains, (lx, ly, lz) local points)."""
    n = _check_soa(p, "points")
    out = _out(out, torch.uint8, n, p[0].device)
    loc = tuple(torch.empty(n, dtype=torch.float64, device=p[0].device) for _ in range(3))
    capi.call("vg_contains_local", geom.handle, int(placed), *_ptrs(p), n, out.data_ptr(), *_ptrs(loc), _stream())
    return out, loc


def unplaced_contains(geom, placed, lp, out=None):
    """UnplacedContains(localPoint) (volumes/PlacedVolume.h:158-160)."""
    n = _check_soa(lp, "local points")
    out = _out(out, torch.uint8, n, lp[0].device)
    capi.call("vg_unplaced_contains", geom.handle, int(placed), *_ptrs(lp), n, out.data_ptr(), _stream())
    return out


def normal(geom, placed, lp):
    """VPlacedVolume::Normal(localPoint, normal&) (volumes/PlacedVolume.h:266-272
    and the shapes' overrides) on local points: returns ((nx, ny, nz) float64,
    uint8 valid).  Box, Tube, Orb, Sphere, Cone, Trd, Trapezoid and Polycone;
    the other shapes raise (their reference Normal asserts)."""
    n = _check_soa(lp, "local points")
    dev = lp[0].device
    nrm = tuple(torch.empty(n, dtype=torch.float64, device=dev) for _ in range(3))
    valid = torch.empty(n, dtype=torch.uint8, device=dev)
    capi.call("vg_normal", geom.handle, int(placed), *_ptrs(lp), n, *_ptrs(nrm), valid.data_ptr(), _stream())
    return nrm, valid


def distance_to_out_normal(geom, placed, lp, ld, step_max=None):
    """The USolids DistanceToOut(point, direction, normal&, convex&, stepMax)
    (volumes/ShapeImplementationHelper.h:277-295) on local rays: returns
    (distance, (nx, ny, nz), uint8 valid, uint8 convex)."""
    n = _check_soa(lp, "local points")
    _check_soa(ld, "local directions", n, lp[0].device)
    dev = lp[0].device
    out = torch.empty(n, dtype=torch.float64, device=dev)
    nrm = tuple(torch.empty(n, dtype=torch.float64, device=dev) for _ in range(3))
    valid = torch.empty(n, dtype=torch.uint8, device=dev)
    convex = torch.empty(n, dtype=torch.uint8, device=dev)
    capi.call("vg_distance_to_out_normal", geom.handle, int(placed), *_ptrs(lp), *_ptrs(ld),
              _step_max(step_max, n, dev), n, out.data_ptr(), *_ptrs(nrm), valid.data_ptr(), convex.data_ptr(),
              _stream())
    return out, nrm, valid, convex


CONTAINS_LOCAL, UNPLACED_CONTAINS = 7, 8


def scalar(geom, placed, query, point, direction=None, step_max=float("inf")):
    """One scalar VPlacedVolume virtual on the device, synchronously
    (vg_scalar_query): returns the result (and the local point for
    CONTAINS_LOCAL).  query: the batch.* query codes, CONTAINS_LOCAL or
    UNPLACED_CONTAINS."""
    import ctypes
    import numpy as np
    p = np.ascontiguousarray(point, dtype=np.float64)
    d = np.ascontiguousarray(direction if direction is not None else (0.0, 0.0, 0.0), dtype=np.float64)
    r = ctypes.c_double()
    loc = np.zeros(3)
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    capi.call("vg_scalar_query", geom.handle, int(placed), int(query), ptr(p), ptr(d), float(step_max),
              ctypes.byref(r), ptr(loc))
    return (r.value, tuple(loc)) if query == CONTAINS_LOCAL else r.value


def inside(geom, placed, p, out=None):
    n = _check_soa(p, "points")
    out = _out(out, torch.int32, n, p[0].device)
    capi.call("vg_inside", geom.handle, int(placed), *_ptrs(p), n, out.data_ptr(), _stream())
    return out


def distance_to_in(geom, placed, p, d, step_max=None, out=None):
    n = _check_soa(p, "points")
    _check_soa(d, "directions", n, p[0].device)
    out = _out(out, torch.float64, n, p[0].device)
    sm = _step_max(step_max, n, p[0].device)
    capi.call("vg_distance_to_in", geom.handle, int(placed), *_ptrs(p), *_ptrs(d), sm, n, out.data_ptr(), _stream())
    return out


def distance_to_in_minimize(geom, placed, daughter_index, p, d, inout_dist, inout_next):
    n = _check_soa(p, "points")
    _check_soa(d, "directions", n, p[0].device)
    _check_arr(inout_dist, "inout_dist", torch.float64, n, p[0].device)
    _check_arr(inout_next, "inout_next", torch.int32, n, p[0].device)
    capi.call("vg_distance_to_in_minimize", geom.handle, int(placed), int(daughter_index), *_ptrs(p), *_ptrs(d), n,
              inout_dist.data_ptr(), inout_next.data_ptr(), _stream())


def distance_to_out(geom, placed, p, d, step_max=None, out=None, next_node=None):
    n = _check_soa(p, "points")
    _check_soa(d, "directions", n, p[0].device)
    out = _out(out, torch.float64, n, p[0].device)
    sm = _step_max(step_max, n, p[0].device)
    nx = None if next_node is None else _check_arr(next_node, "next_node", torch.int32, n, p[0].device).data_ptr()
    capi.call("vg_distance_to_out", geom.handle, int(placed), *_ptrs(p), *_ptrs(d), sm, n, out.data_ptr(), nx,
              _stream())
    return out


def placed_distance_to_out(geom, placed, p, d, step_max=None, out=None):
    n = _check_soa(p, "points")
    _check_soa(d, "directions", n, p[0].device)
    out = _out(out, torch.float64, n, p[0].device)
    sm = _step_max(step_max, n, p[0].device)
    capi.call("vg_placed_distance_to_out", geom.handle, int(placed), *_ptrs(p), *_ptrs(d), sm, n, out.data_ptr(),
              _stream())
    return out


def safety_to_in(geom, placed, p, out=None):
    n = _check_soa(p, "points")
    out = _out(out, torch.float64, n, p[0].device)
    capi.call("vg_safety_to_in", geom.handle, int(placed), *_ptrs(p), n, out.data_ptr(), _stream())
    return out


def safety_to_in_minimize(geom, placed, p, inout):
    n = _check_soa(p, "points")
    _check_arr(inout, "inout", torch.float64, n, p[0].device)
    capi.call("vg_safety_to_in_minimize", geom.handle, int(placed), *_ptrs(p), n, inout.data_ptr(), _stream())


def safety_to_out(geom, placed, p, out=None):
    n = _check_soa(p, "points")
    out = _out(out, torch.float64, n, p[0].device)
    capi.call("vg_safety_to_out", geom.handle, int(placed), *_ptrs(p), n, out.data_ptr(), _stream())
    return out


def safety_to_out_minimize(geom, placed, p, inout):
    n = _check_soa(p, "points")
    _check_arr(inout, "inout", torch.float64, n, p[0].device)
    capi.call("vg_safety_to_out_minimize", geom.handle, int(placed), *_ptrs(p), n, inout.data_ptr(), _stream())


def init(device=0):
    """vg_init on the given device (torch must see the same device)."""
    capi.call("vg_init", int(device))
