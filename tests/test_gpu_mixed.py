"""GPU tests of the mixed-precision shape form (VG_MIXED) on the mixed
workload: float32 inputs, a quarter of the points on the shape's surface
jittered by N(0, (1e-7 mm)^2), the rest uniform in 1.5x the extent, every
zoo shape under a random placement, all seven queries.

Bars:
  * against the oracle on the same float values promoted to double:
    Contains / Inside bit-exact, distances and safeties bit-exact or, for the
    shapes whose distances pass through library transcendentals, within the
    FP64 bar (conftest.assert_agree);
  * against the FP32-storage form (FP64 arithmetic on the promoted inputs):
    identical, lane for lane (NaN where it gives NaN) -- the FP32 filter
    only ever hands back the FP64 result;
  * the filter works: most classifications are decided in FP32 (the FP64
    counter of the task), in the single-task and the batched launch alike.
"""
import numpy as np
import pytest
import torch

import pyoracle as po
from conftest import assert_agree, local_rays
from vecgeom_amd import batch as B
from vecgeom_amd import capi
from vecgeom_amd import geometry as G
from vecgeom_amd import workloads as W

pytestmark = pytest.mark.gpu

N = 65536
QUERIES = [B.CONTAINS, B.INSIDE, B.DISTANCE_TO_IN, B.DISTANCE_TO_OUT, B.PLACED_DISTANCE_TO_OUT, B.SAFETY_TO_IN,
           B.SAFETY_TO_OUT]
NAMES = {B.CONTAINS: "contains", B.INSIDE: "inside", B.DISTANCE_TO_IN: "dist_in", B.DISTANCE_TO_OUT: "dist_out",
         B.PLACED_DISTANCE_TO_OUT: "placed_dist_out", B.SAFETY_TO_IN: "safe_in", B.SAFETY_TO_OUT: "safe_out"}
ORACLE_Q = {B.CONTAINS: po.Q_CONTAINS, B.INSIDE: po.Q_INSIDE, B.DISTANCE_TO_IN: po.Q_DIST_IN,
            B.DISTANCE_TO_OUT: po.Q_DIST_OUT, B.PLACED_DISTANCE_TO_OUT: po.Q_PLACED_DIST_OUT,
            B.SAFETY_TO_IN: po.Q_SAFE_IN, B.SAFETY_TO_OUT: po.Q_SAFE_OUT}
LOCAL = {B.DISTANCE_TO_OUT, B.SAFETY_TO_OUT}
IDENTITY = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1], np.float64)


def same_lanes(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype.kind == "f":
        return (a == b) | (np.isnan(a) & np.isnan(b))
    return a == b


@pytest.fixture(scope="module", autouse=True)
def filter_everywhere(gpu):
    """These tests exercise the FP32 filter on every shape: the policy that
    runs it for every Contains, Inside and DistanceToIn task."""
    B.set_mixed_policy(B.POLICY_ALL)
    yield
    B.set_mixed_policy(B.POLICY_MEASURED)


@pytest.fixture(scope="module")
def mixed_zoo(gpu, oracle):
    shapes = [(n, u) for n, u in W.zoo_shapes() if capi.lib().vg_shape_supported(u.type)]
    rng = np.random.Generator(np.random.Philox(3003))
    xfs = [W.random_placement(rng) for _ in shapes]
    geom, ids, _ = G.single_shape_geometry([u for _, u in shapes], xfs)
    data = []
    for k, ((name, u), xf) in enumerate(zip(shapes, xfs)):
        local = oracle.placed(u.type, u.params, IDENTITY)
        sampler = W.surface_sampler(local.extent(), lambda lp, o=local: o.query(po.Q_CONTAINS, lp).astype(bool),
                                    lambda lp, ld, o=local: o.query(po.Q_DIST_OUT, lp, ld))
        inp = W.mixed_inputs(local.extent(), xf, N, 7000 + k, sampler)
        # Local-frame inputs of DistanceToOut / SafetyToOut, stored as float.
        inp["lp32"] = inp["lp"].astype(np.float32)
        inp["ld32"] = inp["ld"].astype(np.float32)
        data.append(inp)
    return shapes, xfs, geom, ids, data


def tensors(a, device):
    return tuple(torch.from_numpy(np.ascontiguousarray(a[i])).to(device) for i in range(3))


def make_task(pid, q, inp, device, mixed, count=None):
    p = inp["lp32"] if q in LOCAL else inp["p32"]
    d = inp["ld32"] if q in LOCAL else inp["d32"]
    directional = q in B.DIRECTIONAL
    sm = torch.full((p.shape[1],), 1e30, dtype=torch.float64, device=device) if directional else None
    return B.Task(pid, q, tensors(p, device), tensors(d, device) if directional else None, sm, mixed=mixed,
                  fp64_count=count)


def promoted(q, inp):
    """The float inputs of query q promoted to double."""
    p = inp["lp32"].astype(np.float64) if q in LOCAL else inp["p"]
    d = inp["ld32"].astype(np.float64) if q in LOCAL else inp["d"]
    return p, d


def oracle_out(oracle, u, xf, q, inp):
    o = oracle.placed(u.type, u.params, xf.array())
    p, d = promoted(q, inp)
    sm = np.full(p.shape[1], 1e30) if q in B.DIRECTIONAL else None
    return o.query(ORACLE_Q[q], p, d if q in B.DIRECTIONAL else None, sm, nthreads=8)


@pytest.mark.parametrize("k", range(12))
def test_mixed_form_vs_oracle_and_fp32_storage(gpu, oracle, mixed_zoo, k):
    shapes, xfs, geom, ids, data = mixed_zoo
    if k >= len(shapes):
        pytest.skip("shape without device kernels")
    name, u = shapes[k]
    inp = data[k]
    for q in QUERIES:
        count = torch.zeros(1, dtype=torch.int64, device=gpu)
        got = B.query(geom, make_task(ids[k], q, inp, gpu, True, count)).cpu().numpy()
        ref = B.query(geom, make_task(ids[k], q, inp, gpu, False)).cpu().numpy()
        assert same_lanes(got, ref).all(), f"{name} {NAMES[q]}: mixed differs from FP32 storage"
        tor = (u.params, *local_rays(xfs[k], *promoted(q, inp), q in LOCAL)) if q in B.DIRECTIONAL else None
        assert_agree(name, NAMES[q], got, oracle_out(oracle, u, xfs[k], q, inp), "mixed vs oracle", torus=tor)
        n64 = int(count.item())
        assert 0 <= n64 <= N
        if q in (B.CONTAINS, B.INSIDE):
            # the surface quarter goes to FP64; nearly all the rest is decided in FP32
            assert n64 < 0.3 * N, f"{name} {NAMES[q]}: {n64} of {N} lanes evaluated in FP64"
        elif q == B.DISTANCE_TO_IN:
            assert n64 < 0.8 * N, f"{name} dist_in: {n64} of {N} lanes evaluated in FP64"
        else:
            assert n64 == N, f"{name} {NAMES[q]}: FP64 count {n64}, expected every lane"


def test_mixed_batch_equals_single_tasks(gpu, mixed_zoo):
    """The batched launch (every class kernel, its Torus compaction and
    Polyhedron lane-refill paths, every launch shape) gives the single-task
    results, and its FP64 counts equal theirs."""
    shapes, xfs, geom, ids, data = mixed_zoo
    tasks, singles, counts = [], [], []
    for k, inp in enumerate(data):
        for q in QUERIES:
            c = torch.zeros(1, dtype=torch.int64, device=gpu)
            tasks.append(make_task(ids[k], q, inp, gpu, True, c))
            counts.append(c)
            one = torch.zeros(1, dtype=torch.int64, device=gpu)
            singles.append((B.query(geom, make_task(ids[k], q, inp, gpu, True, one)).cpu().numpy(), one))
    batch = B.Batch(geom, tasks)
    try:
        for it, blocks in ((1, 0), (2, 0), (4, 0), (4, 512)):
            for c in counts:
                c.zero_()
            batch.set_launch(it, blocks)
            batch.run()
            torch.cuda.synchronize()
            for t, c, (ref, one) in zip(tasks, counts, singles):
                assert same_lanes(t.out.cpu().numpy(), ref).all(), (it, blocks, t.placed, NAMES[t.query])
                assert int(c.item()) == int(one.item()), (it, blocks, t.placed, NAMES[t.query])
    finally:
        batch.free()


def test_mixed_rejects_float64_inputs(gpu, mixed_zoo):
    shapes, xfs, geom, ids, data = mixed_zoo
    p = tensors(data[0]["p"], gpu)
    with pytest.raises(ValueError, match="float32"):
        B.Task(ids[0], B.CONTAINS, p, mixed=True)


def test_measured_policy_same_results(gpu, mixed_zoo):
    """Under the measured policy (the default) only the filters that pay run
    (Polycone DistanceToIn); every result is the same as under the
    filter-everywhere policy, and the other tasks count every lane as FP64."""
    shapes, xfs, geom, ids, data = mixed_zoo
    try:
        for k, inp in enumerate(data):
            for q in QUERIES:
                B.set_mixed_policy(B.POLICY_ALL)
                a = B.query(geom, make_task(ids[k], q, inp, gpu, True)).cpu().numpy()
                B.set_mixed_policy(B.POLICY_MEASURED)
                c = torch.zeros(1, dtype=torch.int64, device=gpu)
                b = B.query(geom, make_task(ids[k], q, inp, gpu, True, c)).cpu().numpy()
                assert same_lanes(a, b).all(), (shapes[k][0], NAMES[q])
                filtered = shapes[k][0] == "Polycone" and q == B.DISTANCE_TO_IN
                assert (int(c.item()) < N) == filtered, (shapes[k][0], NAMES[q], int(c.item()))
    finally:
        B.set_mixed_policy(B.POLICY_ALL)
