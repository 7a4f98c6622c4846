"""GPU tests of tasks and batches (vg_query, vg_batch_*): a batch of mixed
shapes, queries and basket sizes gives the same results as the per-query
entry points for every launch shape, after autotuning, and when replayed
from a captured hipGraph; FP32-storage inputs give the oracle's results on
the same float values promoted to double."""
import ctypes

import numpy as np
import pytest
import torch

import pyoracle as po
from conftest import assert_agree, local_rays, same
from vecgeom_amd import batch as B
from vecgeom_amd import capi
from vecgeom_amd import geometry as G
from vecgeom_amd import queries as Q
from vecgeom_amd import workloads as W

pytestmark = pytest.mark.gpu

SIZES = [0, 1, 7, 513, 4099, 70001]
QUERIES = [B.CONTAINS, B.INSIDE, B.DISTANCE_TO_IN, B.DISTANCE_TO_OUT, B.PLACED_DISTANCE_TO_OUT, B.SAFETY_TO_IN,
           B.SAFETY_TO_OUT]
NAMES = {B.CONTAINS: "contains", B.INSIDE: "inside", B.DISTANCE_TO_IN: "dist_in", B.DISTANCE_TO_OUT: "dist_out",
         B.PLACED_DISTANCE_TO_OUT: "placed_dist_out", B.SAFETY_TO_IN: "safe_in", B.SAFETY_TO_OUT: "safe_out"}
ORACLE_Q = {B.CONTAINS: po.Q_CONTAINS, B.INSIDE: po.Q_INSIDE, B.DISTANCE_TO_IN: po.Q_DIST_IN,
            B.DISTANCE_TO_OUT: po.Q_DIST_OUT, B.PLACED_DISTANCE_TO_OUT: po.Q_PLACED_DIST_OUT,
            B.SAFETY_TO_IN: po.Q_SAFE_IN, B.SAFETY_TO_OUT: po.Q_SAFE_OUT}
LOCAL = {B.DISTANCE_TO_OUT, B.SAFETY_TO_OUT}


def dev(a, device, dtype=torch.float64):
    return tuple(torch.from_numpy(np.ascontiguousarray(a[i])).to(device=device, dtype=dtype) for i in range(3))


@pytest.fixture(scope="module")
def zoo(gpu):
    shapes = [(n, u) for n, u in W.zoo_shapes() if capi.lib().vg_shape_supported(u.type)]
    rng = np.random.Generator(np.random.Philox(31))
    xfs = [W.random_placement(rng) for _ in shapes]
    geom, ids, _ = G.single_shape_geometry([u for _, u in shapes], xfs)
    return shapes, xfs, geom, ids


def make_tasks(zoo, device, dtype=torch.float64, seed=0):
    """One task per (shape, query) with a basket size cycling through SIZES."""
    shapes, xfs, geom, ids = zoo
    tasks, inputs = [], []
    k = 0
    for (name, u), xf, pid in zip(shapes, xfs, ids):
        inp = W.shape_inputs(geom.placed_extent(pid), xf, max(SIZES), seed=100 + pid + seed)
        for q in QUERIES:
            n = SIZES[k % len(SIZES)]
            k += 1
            p = inp["lp"] if q in LOCAL else inp["p"]
            d = inp["ld"] if q in LOCAL else inp["d"]
            pt, dt = dev(p[:, :n], device, dtype), dev(d[:, :n], device, dtype)
            sm = torch.full((n,), 1e30, dtype=torch.float64, device=device)
            directional = q in B.DIRECTIONAL
            tasks.append(B.Task(pid, q, pt, dt if directional else None, sm if directional else None))
            inputs.append((name, u, xf, pid, q, p[:, :n], d[:, :n]))
    return tasks, inputs


def reference_outputs(geom, tasks):
    """The per-query entry points on the same float64 inputs."""
    ref = []
    for t in tasks:
        p, d = t.p, t.d
        fn = {B.CONTAINS: lambda: Q.contains(geom, t.placed, p), B.INSIDE: lambda: Q.inside(geom, t.placed, p),
              B.DISTANCE_TO_IN: lambda: Q.distance_to_in(geom, t.placed, p, d, t.step_max),
              B.DISTANCE_TO_OUT: lambda: Q.distance_to_out(geom, t.placed, p, d, t.step_max),
              B.PLACED_DISTANCE_TO_OUT: lambda: Q.placed_distance_to_out(geom, t.placed, p, d, t.step_max),
              B.SAFETY_TO_IN: lambda: Q.safety_to_in(geom, t.placed, p),
              B.SAFETY_TO_OUT: lambda: Q.safety_to_out(geom, t.placed, p)}[t.query]
        ref.append(fn().cpu().numpy())
    return ref


def test_batch_matches_per_query_kernels_every_launch_shape(gpu, zoo):
    geom = zoo[2]
    tasks, _ = make_tasks(zoo, gpu)
    ref = reference_outputs(geom, tasks)
    b = B.Batch(geom, tasks)
    for it in (1, 2, 4):
        for blocks in (0, 7, 512):
            for t in tasks:
                t.out.fill_(7)
            b.set_launch(it, blocks)
            b.run()
            torch.cuda.synchronize()
            for t, r in zip(tasks, ref):
                assert same(t.out.cpu().numpy(), r).all(), f"it={it} blocks={blocks} placed={t.placed} q={t.query}"
    ms = b.autotune(2)
    it, blocks, items = b.launch()
    assert ms > 0 and it in (1, 2, 4) and items == sum(t.n for t in tasks)
    for t in tasks:
        t.out.fill_(7)
    b.run()
    torch.cuda.synchronize()
    for t, r in zip(tasks, ref):
        assert same(t.out.cpu().numpy(), r).all()
    # vg_query on each task alone
    for t, r in zip(tasks, ref):
        t.out.fill_(7)
        B.query(geom, t)
        torch.cuda.synchronize()
        assert same(t.out.cpu().numpy(), r).all()
    b.free()


def test_fp32_storage_matches_oracle_on_promoted_inputs(gpu, zoo, oracle):
    geom = zoo[2]
    tasks, inputs = make_tasks(zoo, gpu, torch.float32, seed=5)
    b = B.Batch(geom, tasks)
    b.run()
    torch.cuda.synchronize()
    for t, (name, u, xf, pid, q, p, d) in zip(tasks, inputs):
        if t.n == 0:
            continue
        pf = np.stack([a.cpu().numpy().astype(np.float64) for a in t.p])
        df = np.stack([a.cpu().numpy().astype(np.float64) for a in t.d]) if t.d is not None else None
        o = oracle.placed(u.type, u.params, xf.array())
        args = (pf,) if df is None else (pf, df, np.full(t.n, 1e30))
        want = o.query(ORACLE_Q[q], *args)
        tor = (u.params, *local_rays(xf, pf, df, q in LOCAL)) if df is not None else None
        assert_agree(name, NAMES[q], t.out.cpu().numpy(), want, "fp32 storage", torus=tor)
    b.free()


def test_batch_graph_replay(gpu, zoo):
    """vg_batch_run inside a captured graph: every replay, with new inputs
    copied in eagerly between replays, equals the eager run."""
    geom = zoo[2]
    tasks, inputs = make_tasks(zoo, gpu, seed=9)
    b = B.Batch(geom, tasks)
    b.set_launch(2, 0)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        b.run()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        b.run()
    torch.cuda.synchronize()
    rng = np.random.Generator(np.random.Philox(77))
    for rep in range(3):
        for t in tasks:
            if t.n:
                for a in t.p:
                    a.add_(torch.from_numpy(rng.normal(0, 1e-3, t.n)).to(gpu))
        eager = reference_outputs(geom, tasks)
        for t in tasks:
            t.out.fill_(3)
        graph.replay()
        torch.cuda.synchronize()
        for t, r in zip(tasks, eager):
            assert same(t.out.cpu().numpy(), r).all(), f"replay {rep}"
    del graph
    b.free()


def test_task_argument_errors(gpu, zoo):
    geom = zoo[2]
    x = torch.zeros(4, dtype=torch.float64, device=gpu)
    bad = B.CQueryTask(10 ** 6, B.CONTAINS, B.F64, 0, 4, x.data_ptr(), x.data_ptr(), x.data_ptr(), None, None, None,
                       None, x.data_ptr())
    rc, msg = capi.status("vg_query", geom.handle, ctypes.byref(bad), None)
    assert rc == -1 and "placed" in msg
    bad.placed, bad.query = zoo[3][0], 99
    rc, _ = capi.status("vg_query", geom.handle, ctypes.byref(bad), None)
    assert rc == -1
    bad.query, bad.dx = B.DISTANCE_TO_IN, None
    rc, msg = capi.status("vg_query", geom.handle, ctypes.byref(bad), None)
    assert rc == -1 and "null" in msg
    h = ctypes.c_void_p()
    arr = (B.CQueryTask * 1)(bad)
    rc, _ = capi.status("vg_batch_create", geom.handle, arr, 1, ctypes.byref(h))
    assert rc == -1 and not h.value
