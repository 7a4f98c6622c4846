"""GPU test of the no-allocation batch entry point (vg_batch_run_noalloc)
through the C-ABI: two calls in a row, the second with fewer and smaller
tasks, each compared with vg_batch_create + vg_batch_run on the same tasks
-- bit for bit, FP64, FP32-storage and mixed-precision tasks alike; and the
capacity error raised before any launch."""
import ctypes

import numpy as np
import pytest
import torch

from vecgeom_amd import batch as B
from vecgeom_amd import capi
from vecgeom_amd import geometry as G
from vecgeom_amd import workloads as W

pytestmark = pytest.mark.gpu

capi.register({"vg_batch_run_noalloc": [ctypes.c_void_p, ctypes.POINTER(B.CQueryTask), ctypes.c_int32,
                                        ctypes.c_void_p]})
QUERIES = [B.CONTAINS, B.INSIDE, B.DISTANCE_TO_IN, B.DISTANCE_TO_OUT, B.PLACED_DISTANCE_TO_OUT, B.SAFETY_TO_IN,
           B.SAFETY_TO_OUT]


def same_lanes(a, b):
    if a.dtype.kind == "f":
        return (a == b) | (np.isnan(a) & np.isnan(b))
    return a == b


@pytest.fixture(scope="module")
def zoo(gpu):
    shapes = [(n, u) for n, u in W.zoo_shapes() if capi.lib().vg_shape_supported(u.type)]
    rng = np.random.Generator(np.random.Philox(515))
    xfs = [W.random_placement(rng) for _ in shapes]
    geom, ids, _ = G.single_shape_geometry([u for _, u in shapes], xfs)
    return shapes, xfs, geom, ids


def tasks_for(zoo, n, every, device, seed):
    """Tasks over every `every`-th (shape, query) pair, cycling the input
    kinds, with n queries each."""
    shapes, xfs, geom, ids = zoo
    out = []
    k = 0
    for (name, u), xf, pid in zip(shapes, xfs, ids):
        inp = W.shape_inputs(geom.placed_extent(pid), xf, n, seed=seed + pid)
        for q in QUERIES:
            k += 1
            if k % every:
                continue
            kind = k % 3  # 0 FP64, 1 FP32 storage, 2 mixed
            dt = torch.float64 if kind == 0 else torch.float32
            local = q in (B.DISTANCE_TO_OUT, B.SAFETY_TO_OUT)
            p = tuple(torch.from_numpy(np.ascontiguousarray(inp["lp" if local else "p"][i])).to(device, dt)
                      for i in range(3))
            d = tuple(torch.from_numpy(np.ascontiguousarray(inp["ld" if local else "d"][i])).to(device, dt)
                      for i in range(3))
            directional = q in B.DIRECTIONAL
            out.append(B.Task(pid, q, p, d if directional else None, None, mixed=kind == 2))
    return out


def run_noalloc(geom, tasks):
    arr = (B.CQueryTask * max(1, len(tasks)))(*[t.c() for t in tasks])
    capi.call("vg_batch_run_noalloc", geom.handle, arr, len(tasks), torch.cuda.current_stream().cuda_stream)


def expected(geom, tasks):
    copies = [B.Task(t.placed, t.query, t.p, t.d, t.step_max, mixed=t.input_type == B.MIXED) for t in tasks]
    b = B.Batch(geom, copies)
    try:
        b.run()
        torch.cuda.synchronize()
        return [c.out.cpu().numpy() for c in copies]
    finally:
        b.free()


def test_noalloc_matches_batch_on_two_calls(gpu, zoo):
    geom = zoo[2]
    for n, every, seed in ((70001, 1, 11), (4099, 3, 12)):  # the second call: fewer, smaller tasks
        tasks = tasks_for(zoo, n, every, gpu, seed)
        for t in tasks:
            t.out.fill_(7)
        run_noalloc(geom, tasks)
        torch.cuda.synchronize()
        want = expected(geom, tasks)
        for t, w in zip(tasks, want):
            assert same_lanes(t.out.cpu().numpy(), w).all(), (n, t.placed, t.query, t.input_type)


def test_noalloc_capacity_error(gpu, zoo):
    geom, ids = zoo[2], zoo[3]
    x = torch.zeros(4, dtype=torch.float64, device=gpu)
    one = B.Task(ids[0], B.CONTAINS, (x, x, x))
    arr = (B.CQueryTask * 257)(*([one.c()] * 257))
    rc, msg = capi.status("vg_batch_run_noalloc", geom.handle, arr, 257, torch.cuda.current_stream().cuda_stream)
    assert rc == capi.VG_ERR_CAPACITY and "256" in msg
