"""GPU tests of the Polyhedron's DistanceToIn and DistanceToOut walks with
lane refill (WalkKernel for single launches, the staged batch class for
batches):
bit-exact against the oracle for basket sizes around the 64-lane and
512-query chunk boundaries, for polyhedra whose parameter block is staged in
LDS and one that is read from global memory, in FP64 and with FP32-storage
inputs (against the oracle on the promoted inputs)."""
import numpy as np
import pytest
import torch

import pyoracle as po
from conftest import same
from test_shapes_hostcheck import EXTRA
from vecgeom_amd import batch as B
from vecgeom_amd import geometry as G
from vecgeom_amd import queries as Q
from vecgeom_amd import workloads as W

pytestmark = pytest.mark.gpu

SIZES = [1, 63, 64, 65, 511, 512, 513, 2049, 70001]
NAMES = ["Polyhedron", "PolyhedronPhi", "PolyhedronDupZPhi", "PolyhedronMany"]


def shape(name):
    return dict(W.zoo_shapes() + EXTRA)[name]


@pytest.fixture(scope="module", params=NAMES)
def case(request, gpu):
    u = shape(request.param)
    rng = np.random.Generator(np.random.Philox(4242))
    xf = W.random_placement(rng)
    geom, ids, _ = G.single_shape_geometry([u], [xf])
    inp = W.shape_inputs(geom.placed_extent(ids[0]), xf, max(SIZES), seed=4243)
    # every other ray aimed at a point of the extent, so that many hit
    lo, hi = geom.placed_extent(ids[0])
    tgt = rng.uniform(lo, hi, size=(max(SIZES), 3)).T
    v = tgt - inp["p"]
    v /= np.linalg.norm(v, axis=0)
    inp["d"][:, ::2] = v[:, ::2]
    return request.param, u, xf, geom, ids[0], inp


def dev(a, device, dtype=torch.float64):
    return tuple(torch.from_numpy(np.ascontiguousarray(a[i])).to(device=device, dtype=dtype) for i in range(3))


def test_walk_single_launch_matches_oracle(gpu, case, oracle):
    name, u, xf, geom, pid, inp = case
    o = oracle.placed(u.type, u.params, xf.array())
    for n in SIZES:
        p, d = inp["p"][:, :n], inp["d"][:, :n]
        sm = np.full(n, 1e30)
        sm[::3] = 5.0  # some step limits below the distance
        got = Q.distance_to_in(geom, pid, dev(p, gpu), dev(d, gpu),
                               torch.from_numpy(sm).to(gpu)).cpu().numpy()
        want = o.query(po.Q_DIST_IN, p, d, sm)
        ok = same(got, want)
        assert ok.all(), f"{name} n={n}: {int((~ok).sum())} lanes differ, first {np.flatnonzero(~ok)[:5]}"


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_walk_batched_matches_oracle(gpu, case, oracle, dtype):
    name, u, xf, geom, pid, inp = case
    o = oracle.placed(u.type, u.params, xf.array())
    tasks = []
    for n in SIZES:
        sm = torch.full((n,), 1e30, dtype=torch.float64, device=gpu)
        tasks.append(B.Task(pid, B.DISTANCE_TO_IN, dev(inp["p"][:, :n], gpu, dtype), dev(inp["d"][:, :n], gpu, dtype),
                            sm))
    b = B.Batch(geom, tasks)
    for it in (1, 2, 4):
        b.set_launch(it, 0)
        for t in tasks:
            t.out.fill_(7)
        b.run()
        torch.cuda.synchronize()
        for t, n in zip(tasks, SIZES):
            pf = np.stack([a.cpu().numpy().astype(np.float64) for a in t.p])
            df = np.stack([a.cpu().numpy().astype(np.float64) for a in t.d])
            want = o.query(po.Q_DIST_IN, pf, df, np.full(n, 1e30))
            ok = same(t.out.cpu().numpy(), want)
            assert ok.all(), f"{name} {dtype} it={it} n={n}: {int((~ok).sum())} lanes differ"
    b.free()


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_walk_out_matches_oracle(gpu, case, oracle, dtype):
    """DistanceToOut (local frame) and PlacedDistanceToOut from points inside
    the shape, single launches and batched."""
    name, u, xf, geom, pid, inp = case
    o = oracle.placed(u.type, u.params, xf.array())
    inside = np.flatnonzero(o.query(po.Q_CONTAINS, inp["p"]).astype(bool))
    lp, ld = inp["lp"][:, inside], inp["ld"][:, inside]
    p, d = inp["p"][:, inside], inp["d"][:, inside]
    sizes = [n for n in SIZES if n <= inside.size] + [inside.size]
    tasks, want = [], []
    for n in sizes:
        sm = np.full(n, 1e30)
        for q, pp, dd, oq in ((B.DISTANCE_TO_OUT, lp, ld, po.Q_DIST_OUT), (B.PLACED_DISTANCE_TO_OUT, p, d,
                                                                           po.Q_PLACED_DIST_OUT)):
            pt, dt = dev(pp[:, :n], gpu, dtype), dev(dd[:, :n], gpu, dtype)
            tasks.append(B.Task(pid, q, pt, dt, torch.from_numpy(sm).to(gpu)))
            pf = np.stack([a.cpu().numpy().astype(np.float64) for a in pt])
            df = np.stack([a.cpu().numpy().astype(np.float64) for a in dt])
            want.append(o.query(oq, pf, df, sm))
    if dtype == torch.float64:
        for t, w in zip(tasks, want):
            fn = Q.distance_to_out if t.query == B.DISTANCE_TO_OUT else Q.placed_distance_to_out
            got = fn(geom, pid, t.p, t.d, t.step_max).cpu().numpy()
            ok = same(got, w)
            assert ok.all(), f"{name} single q={t.query} n={t.n}: {int((~ok).sum())} lanes differ"
    b = B.Batch(geom, tasks)
    for it in (1, 2):
        b.set_launch(it, 0)
        for t in tasks:
            t.out.fill_(7)
        b.run()
        torch.cuda.synchronize()
        for t, w in zip(tasks, want):
            ok = same(t.out.cpu().numpy(), w)
            assert ok.all(), f"{name} {dtype} it={it} q={t.query} n={t.n}: {int((~ok).sum())} lanes differ"
    b.free()
