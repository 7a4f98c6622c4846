"""Tasks, batches, FP32-storage and mixed-precision queries through the C-ABI.

A task is one basket query on one placed volume (VPlacedVolume SOA3D
virtuals, volumes/PlacedVolume.h:147-246) whose inputs may be float64 (the
FP64 form) or float32: FP32 storage (promoted to double before any
arithmetic) or, with mixed=True, the mixed-precision form (an FP32 filter
decides Contains / Inside and DistanceToIn misses with a rigorous rounding
margin; every other lane runs the FP64 code, so results equal FP32
storage's bit for bit).  A Batch runs a fixed table of tasks -- any mix of shapes,
queries and basket sizes -- in one kernel launch per shape weight class, and
can autotune its launch shape.  Work goes on torch's current stream.
"""
import ctypes

import torch

from . import capi

CONTAINS, INSIDE, DISTANCE_TO_IN, DISTANCE_TO_OUT, PLACED_DISTANCE_TO_OUT, SAFETY_TO_IN, SAFETY_TO_OUT = range(7)
F64, F32, MIXED = 0, 1, 2
# vg_set_mixed_policy: the FP32 filter where measured to pay, or everywhere.
POLICY_MEASURED, POLICY_ALL = 0, 1
OUT_DTYPE = {CONTAINS: torch.uint8, INSIDE: torch.int32}
DIRECTIONAL = {DISTANCE_TO_IN, DISTANCE_TO_OUT, PLACED_DISTANCE_TO_OUT}

c_vp = ctypes.c_void_p


class CQueryTask(ctypes.Structure):
    _fields_ = [
        ("placed", ctypes.c_int32),
        ("query", ctypes.c_int32),
        ("input_type", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("n", ctypes.c_uint64),
        ("x", c_vp), ("y", c_vp), ("z", c_vp),
        ("dx", c_vp), ("dy", c_vp), ("dz", c_vp),
        ("step_max", c_vp),
        ("out", c_vp),
        ("fp64_count", c_vp),
    ]


_T = ctypes.POINTER(CQueryTask)
capi.register({
    "vg_set_mixed_policy": [ctypes.c_int32],
    "vg_query": [c_vp, _T, c_vp],
    "vg_batch_create": [c_vp, _T, ctypes.c_int32, ctypes.POINTER(c_vp)],
    "vg_batch_run": [c_vp, c_vp],
    "vg_batch_set_launch": [c_vp, ctypes.c_int32, ctypes.c_uint32],
    "vg_batch_get_launch": [c_vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_uint32),
                            ctypes.POINTER(ctypes.c_uint64)],
    "vg_batch_autotune": [c_vp, ctypes.c_int32, c_vp, ctypes.POINTER(ctypes.c_float)],
    "vg_batch_free": [c_vp],
    "vg_batch_set_timing": [c_vp, ctypes.c_int32],
    "vg_batch_class_stats": [c_vp, ctypes.c_int32, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                             ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_float)],
})

# Weight classes of a batch (one kernel launch each); mixed-precision tasks
# run in classes of their own, 5 + the base class.
_BASE_CLASS_NAMES = ["light (Box, Tube, Orb, Trd, Trapezoid, Parallelepiped, Paraboloid)",
                     "Cone, Sphere; Torus Contains, Inside and safeties", "Torus distances",
                     "Polycone, Polyhedron (parameters staged in LDS)",
                     "Polycone, Polyhedron (parameters in global memory)"]
CLASS_NAMES = _BASE_CLASS_NAMES + [f"mixed: {n}" for n in _BASE_CLASS_NAMES]


def _stream():
    return torch.cuda.current_stream().cuda_stream


class Task:
    """One query: placed id, query code, point triple, optional direction
    triple and step_max, output tensor (allocated if None).  mixed=True
    (float32 inputs only) selects the mixed-precision form; fp64_count, a
    one-element int64 CUDA tensor, receives the number of queries each run
    evaluated in FP64."""

    def __init__(self, placed, query, p, d=None, step_max=None, out=None, mixed=False, fp64_count=None):
        if len(p) != 3 or (query in DIRECTIONAL and (d is None or len(d) != 3)):
            raise ValueError("points (and directions for distance queries) must be (x, y, z) triples")
        if query not in range(7):
            raise ValueError(f"unknown query code {query}")
        dt = p[0].dtype
        if dt not in (torch.float64, torch.float32):
            raise ValueError("inputs must be float64 or float32")
        n = p[0].numel()
        dev = p[0].device
        for a in list(p) + (list(d) if d is not None else []):
            if a.dtype != dt or a.numel() != n or not a.is_cuda or a.device != dev or not a.is_contiguous():
                raise ValueError("inputs must be contiguous CUDA tensors of one dtype, length and device")
        if step_max is not None and (step_max.dtype != torch.float64 or step_max.numel() != n or not step_max.is_cuda
                                     or step_max.device != dev or not step_max.is_contiguous()):
            raise ValueError("step_max must be a contiguous float64 CUDA tensor with one entry per point")
        odt = OUT_DTYPE.get(query, torch.float64)
        if out is None:
            out = torch.empty(n, dtype=odt, device=dev)
        elif (out.dtype != odt or out.numel() != n or not out.is_cuda or out.device != dev
              or not out.is_contiguous()):
            raise ValueError(f"out must be a contiguous {odt} CUDA tensor with one entry per point")
        if mixed and dt != torch.float32:
            raise ValueError("the mixed-precision form takes float32 inputs")
        if fp64_count is not None and (fp64_count.dtype != torch.int64 or fp64_count.numel() < 1
                                       or not fp64_count.is_cuda or fp64_count.device != dev):
            raise ValueError("fp64_count must be an int64 CUDA tensor on the inputs' device")
        self.placed, self.query, self.p, self.d, self.step_max, self.out, self.n = placed, query, p, d, step_max, out, n
        self.fp64_count = fp64_count
        self.input_type = MIXED if mixed else (F32 if dt == torch.float32 else F64)

    def c(self):
        d = self.d if self.d is not None else (None, None, None)
        ptr = lambda a: None if a is None else a.data_ptr()
        return CQueryTask(int(self.placed), int(self.query), self.input_type, 0, self.n,
                          *[ptr(a) for a in self.p], *[ptr(a) for a in d], ptr(self.step_max), self.out.data_ptr(),
                          ptr(self.fp64_count))


def set_mixed_policy(policy):
    """Selects which mixed-precision tasks run the FP32 filter (applies to
    batches and queries made afterwards; results do not depend on it)."""
    capi.call("vg_set_mixed_policy", int(policy))


def query(geom, task):
    """Runs one task (any query, float64 or float32 inputs); returns its out."""
    c = task.c()
    capi.call("vg_query", geom.handle, ctypes.byref(c), _stream())
    return task.out


class Batch:
    """A fixed table of tasks run together (vg_batch_*)."""

    def __init__(self, geom, tasks):
        self.geom = geom
        self.tasks = list(tasks)  # keeps the tensors alive
        arr = (CQueryTask * max(1, len(self.tasks)))(*[t.c() for t in self.tasks])
        h = c_vp()
        capi.call("vg_batch_create", geom.handle, arr, len(self.tasks), ctypes.byref(h))
        self.handle = h

    def run(self):
        capi.call("vg_batch_run", self.handle, _stream())

    def set_launch(self, items_per_lane, blocks=0):
        capi.call("vg_batch_set_launch", self.handle, int(items_per_lane), int(blocks))

    def launch(self):
        it, bl, items = ctypes.c_int32(), ctypes.c_uint32(), ctypes.c_uint64()
        capi.call("vg_batch_get_launch", self.handle, ctypes.byref(it), ctypes.byref(bl), ctypes.byref(items))
        return it.value, bl.value, items.value

    def autotune(self, repetitions=3):
        ms = ctypes.c_float()
        capi.call("vg_batch_autotune", self.handle, int(repetitions), _stream(), ctypes.byref(ms))
        return ms.value

    def set_timing(self, enable=True):
        capi.call("vg_batch_set_timing", self.handle, 1 if enable else 0)

    def class_stats(self):
        """Per weight class: queries and algorithmic bytes per run, timed runs
        and mean kernel duration (ms) since timing was enabled."""
        out = []
        for c in range(len(CLASS_NAMES)):
            it, by, runs, ms = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_int32(), ctypes.c_float()
            capi.call("vg_batch_class_stats", self.handle, c, ctypes.byref(it), ctypes.byref(by), ctypes.byref(runs),
                      ctypes.byref(ms))
            out.append({"class": c, "name": CLASS_NAMES[c], "items": it.value, "bytes": by.value, "runs": runs.value,
                        "mean_ms": ms.value})
        return out

    def free(self):
        if self.handle:
            capi.call("vg_batch_free", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
