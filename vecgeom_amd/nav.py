"""Device navigation through the C-ABI.

Mirrors the reference's navigator API for baskets of tracks:
  NavStates                 NavStatePool (navigation/NavStatePool.h:34-110),
                            structure-of-arrays torch tensors
  locate_points             SimpleNavigator::LocatePoint(world, p, s, true)
                            (navigation/SimpleNavigator.h:205-241)
  find_next_boundary_and_step
                            ABBoxNavigator / SimpleNavigator
                            FindNextBoundaryAndStep
                            (navigation/ABBoxNavigator.h:201-424,
                            navigation/SimpleNavigator.h:341-482)
  get_safety                ABBoxNavigator::GetSafety (ABBoxNavigator.h:428-452)
  trace_tracks              the TraceTrack benchmark loop
                            (test/globalbenchmarks/TraceTrack.cpp:170-220)
Points and directions are (x, y, z) triples of contiguous float64 CUDA
tensors; work goes on torch's current stream.
"""
import ctypes

import torch

from . import capi

ABBOX = 0
SIMPLE = 1
ABBOX_MIXED = 2  # ABBox navigator with the FP32 daughter filter (same results)
SIMPLE_BASKET = 3  # SimpleNavigator's basket step: one current volume for the whole basket
MAX_DEPTH = 16
HASH_SEED = 1469598103934665603
HASH_MUL = 1099511628211

c_i32 = ctypes.c_int32
c_u64 = ctypes.c_uint64
c_sz = ctypes.c_size_t
c_vp = ctypes.c_void_p


class CNavStates(ctypes.Structure):
    _fields_ = [
        ("level", c_vp),
        ("path", c_vp),
        ("boundary", c_vp),
        ("capacity", c_u64),
        ("depth", c_i32),
        ("reserved", c_i32),
    ]


_P = ctypes.POINTER(CNavStates)
capi.register({
    "vg_nav_states_create": [c_vp, c_u64, c_i32, _P],
    "vg_nav_states_free": [_P],
    "vg_nav_states_to_host": [_P, c_u64, c_vp, c_vp, c_vp],
    "vg_nav_states_from_host": [_P, c_u64, c_vp, c_vp, c_vp],
    "vg_locate_points": [c_vp, c_vp, c_vp, c_vp, c_sz, _P, c_vp],
    "vg_find_next_boundary_and_step": [c_vp, c_i32] + [c_vp] * 7 + [c_sz, _P, _P, c_vp, c_vp],
    "vg_get_safety": [c_vp, c_vp, c_vp, c_vp, c_sz, _P, c_vp, c_vp],
    "vg_trace_tracks": [c_vp, c_i32] + [c_vp] * 6 + [c_sz, ctypes.c_double, c_i32] + [c_vp] * 8,
})


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _soa(p, name, n=None):
    if len(p) != 3:
        raise ValueError(f"{name} must be an (x, y, z) triple")
    m = p[0].numel() if n is None else n
    for a in p:
        if a.dtype != torch.float64 or not a.is_cuda or a.numel() != m or not a.is_contiguous():
            raise ValueError(f"{name} components must be contiguous float64 CUDA tensors of length {m}")
    return m


class NavStates:
    """Navigation states of `capacity` tracks held in torch tensors.

    level (int32 [capacity]), path (int32 [depth, capacity], -1 beyond the
    level), boundary (uint8 [capacity]).
    """

    def __init__(self, geom, capacity, depth=None, device="cuda"):
        depth = geom.max_depth() if depth is None else int(depth)
        if depth < geom.max_depth():
            raise ValueError("state depth below the geometry depth")
        self.capacity = int(capacity)
        self.depth = depth
        self.level = torch.zeros(self.capacity, dtype=torch.int32, device=device)
        self.path = torch.full((depth, self.capacity), -1, dtype=torch.int32, device=device)
        self.boundary = torch.zeros(self.capacity, dtype=torch.uint8, device=device)

    def c(self):
        return CNavStates(self.level.data_ptr(), self.path.data_ptr(), self.boundary.data_ptr(), self.capacity,
                          self.depth, 0)

    def top(self):
        """Placed id of the deepest volume per track (-1 outside)."""
        idx = (self.level.long() - 1).clamp(min=0)
        t = self.path.gather(0, idx.unsqueeze(0)).squeeze(0)
        return torch.where(self.level > 0, t, torch.full_like(t, -1))

    @classmethod
    def from_arrays(cls, geom, level, path, boundary=None, device="cuda"):
        level = torch.as_tensor(level, dtype=torch.int32)
        s = cls(geom, level.numel(), depth=path.shape[0], device=device)
        s.level.copy_(level)
        s.path.copy_(torch.as_tensor(path, dtype=torch.int32))
        if boundary is not None:
            s.boundary.copy_(torch.as_tensor(boundary, dtype=torch.uint8))
        return s


def locate_points(geom, p, states=None):
    n = _soa(p, "points")
    states = NavStates(geom, n, device=p[0].device) if states is None else states
    cs = states.c()
    capi.call("vg_locate_points", geom.handle, *[a.data_ptr() for a in p], n, ctypes.byref(cs), _stream())
    return states


def find_next_boundary_and_step(geom, p, d, cur, pstep=None, navigator=ABBOX, new=None, step=None):
    n = _soa(p, "points")
    _soa(d, "directions", n)
    new = NavStates(geom, cur.capacity, cur.depth, device=p[0].device) if new is None else new
    step = torch.empty(n, dtype=torch.float64, device=p[0].device) if step is None else step
    ps = None
    if pstep is not None:
        if pstep.dtype != torch.float64 or pstep.numel() != n or not pstep.is_cuda:
            raise ValueError("pstep must be a float64 CUDA tensor with one entry per track")
        ps = pstep.data_ptr()
    c0, c1 = cur.c(), new.c()
    capi.call("vg_find_next_boundary_and_step", geom.handle, int(navigator), *[a.data_ptr() for a in p],
              *[a.data_ptr() for a in d], ps, n, ctypes.byref(c0), ctypes.byref(c1), step.data_ptr(), _stream())
    return new, step


def get_safety(geom, p, cur, out=None):
    n = _soa(p, "points")
    out = torch.empty(n, dtype=torch.float64, device=p[0].device) if out is None else out
    c0 = cur.c()
    capi.call("vg_get_safety", geom.handle, *[a.data_ptr() for a in p], n, ctypes.byref(c0), out.data_ptr(),
              _stream())
    return out


class TraceResult:
    def __init__(self, n, device):
        self.nsteps = torch.empty(n, dtype=torch.int32, device=device)
        self.final = [torch.empty(n, dtype=torch.float64, device=device) for _ in range(3)]
        self.travelled = torch.empty(n, dtype=torch.float64, device=device)
        self.hash = torch.empty(n, dtype=torch.int64, device=device)  # bit pattern of the uint64 hash
        self.exact = torch.empty(n, dtype=torch.int32, device=device)

    def hash_u64(self):
        return self.hash.cpu().numpy().view("uint64")


def trace_tracks(geom, p, d, pstep=float("inf"), navigator=ABBOX, max_steps=100000, out=None):
    n = _soa(p, "points")
    _soa(d, "directions", n)
    out = TraceResult(n, p[0].device) if out is None else out
    capi.call("vg_trace_tracks", geom.handle, int(navigator), *[a.data_ptr() for a in p], *[a.data_ptr() for a in d],
              n, float(pstep), int(max_steps), out.nsteps.data_ptr(), *[a.data_ptr() for a in out.final],
              out.travelled.data_ptr(), out.hash.data_ptr(), out.exact.data_ptr(), _stream())
    return out
