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


def _check_soa(p, name):
    if len(p) != 3:
        raise ValueError(f"{name} must be an (x, y, z) triple")
    n = p[0].numel()
    for a in p:
        if a.dtype != torch.float64 or not a.is_cuda or a.numel() != n or not a.is_contiguous():
            raise ValueError(f"{name} components must be contiguous float64 CUDA tensors of equal length")
    return n


def _ptrs(p):
    return [a.data_ptr() for a in p]


def contains(geom, placed, p, out=None):
    n = _check_soa(p, "points")
    out = torch.empty(n, dtype=torch.uint8, device=p[0].device) if out is None else out
    capi.call("vg_contains", geom.handle, int(placed), *_ptrs(p), n, out.data_ptr(), _stream())
    return out


def inside(geom, placed, p, out=None):
    n = _check_soa(p, "points")
    out = torch.empty(n, dtype=torch.int32, device=p[0].device) if out is None else out
    capi.call("vg_inside", geom.handle, int(placed), *_ptrs(p), n, out.data_ptr(), _stream())
    return out


def distance_to_in(geom, placed, p, d, step_max=None, out=None):
    n = _check_soa(p, "points")
    _check_soa(d, "directions")
    out = torch.empty(n, dtype=torch.float64, device=p[0].device) if out is None else out
    sm = step_max.data_ptr() if step_max is not None else None
    capi.call("vg_distance_to_in", geom.handle, int(placed), *_ptrs(p), *_ptrs(d), sm, n, out.data_ptr(), _stream())
    return out


def distance_to_in_minimize(geom, placed, daughter_index, p, d, inout_dist, inout_next):
    n = _check_soa(p, "points")
    _check_soa(d, "directions")
    capi.call("vg_distance_to_in_minimize", geom.handle, int(placed), int(daughter_index), *_ptrs(p), *_ptrs(d), n,
              inout_dist.data_ptr(), inout_next.data_ptr(), _stream())


def distance_to_out(geom, placed, p, d, step_max=None, out=None, next_node=None):
    n = _check_soa(p, "points")
    _check_soa(d, "directions")
    out = torch.empty(n, dtype=torch.float64, device=p[0].device) if out is None else out
    sm = step_max.data_ptr() if step_max is not None else None
    nx = next_node.data_ptr() if next_node is not None else None
    capi.call("vg_distance_to_out", geom.handle, int(placed), *_ptrs(p), *_ptrs(d), sm, n, out.data_ptr(), nx,
              _stream())
    return out


def placed_distance_to_out(geom, placed, p, d, step_max=None, out=None):
    n = _check_soa(p, "points")
    _check_soa(d, "directions")
    out = torch.empty(n, dtype=torch.float64, device=p[0].device) if out is None else out
    sm = step_max.data_ptr() if step_max is not None else None
    capi.call("vg_placed_distance_to_out", geom.handle, int(placed), *_ptrs(p), *_ptrs(d), sm, n, out.data_ptr(),
              _stream())
    return out


def safety_to_in(geom, placed, p, out=None):
    n = _check_soa(p, "points")
    out = torch.empty(n, dtype=torch.float64, device=p[0].device) if out is None else out
    capi.call("vg_safety_to_in", geom.handle, int(placed), *_ptrs(p), n, out.data_ptr(), _stream())
    return out


def safety_to_in_minimize(geom, placed, p, inout):
    n = _check_soa(p, "points")
    capi.call("vg_safety_to_in_minimize", geom.handle, int(placed), *_ptrs(p), n, inout.data_ptr(), _stream())


def safety_to_out(geom, placed, p, out=None):
    n = _check_soa(p, "points")
    out = torch.empty(n, dtype=torch.float64, device=p[0].device) if out is None else out
    capi.call("vg_safety_to_out", geom.handle, int(placed), *_ptrs(p), n, out.data_ptr(), _stream())
    return out


def safety_to_out_minimize(geom, placed, p, inout):
    n = _check_soa(p, "points")
    capi.call("vg_safety_to_out_minimize", geom.handle, int(placed), *_ptrs(p), n, inout.data_ptr(), _stream())


def init(device=0):
    """vg_init on the given device (torch must see the same device)."""
    capi.call("vg_init", int(device))
