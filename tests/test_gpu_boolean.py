"""GPU parity of Boolean volumes (vecgeom_amd/csrc/shapes_boolean.hpp,
vg_boolean.hip) through the C-ABI: every basket query (with the Minimize
forms and the DistanceToOut next-node output) against the oracle on seeded
random baskets, bit for bit, under identity and random placements; the
scalar entry point on a sample; and the refusals (batches, navigation).  The
reference's golden vectors for these volumes run in test_gpu_shapes.py."""
import numpy as np
import pytest
import torch

import pyoracle as po
from conftest import same
from vecgeom_amd import batch as B
from vecgeom_amd import capi
from vecgeom_amd import geometry as G
from vecgeom_amd import nav as N
from vecgeom_amd import queries as Q
from vecgeom_amd import workloads as W

pytestmark = pytest.mark.gpu

NAMES = [n for n, _ in W.boolean_shapes()]


def dev(a, device):
    return tuple(torch.from_numpy(np.ascontiguousarray(a[i])).to(device) for i in range(3))


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.fixture(scope="module")
def booleans(gpu):
    shapes = W.boolean_shapes()
    rng = np.random.Generator(np.random.Philox(4242))
    xfs = [G.Transformation3D() if k % 2 == 0 else W.random_placement(rng) for k in range(len(shapes))]
    geom, ids, _ = G.single_shape_geometry([u for _, u in shapes], xfs)
    return shapes, xfs, geom, ids


@pytest.mark.parametrize("k", range(len(NAMES)))
def test_gpu_boolean_queries_vs_oracle(gpu, oracle, booleans, k):
    shapes, xfs, geom, ids = booleans
    name, u = shapes[k]
    xf, pid = xfs[k], ids[k]
    n = 65536 + 3
    inp = W.shape_inputs(((-60, -60, -60), (60, 60, 60)), xf, n, seed=900 + k)
    rng = np.random.Generator(np.random.Philox(77 + k))
    sm = np.where(rng.random(n) < 0.2, rng.uniform(0, 40, n), 1e30)
    p, d, lp, ld = dev(inp["p"], gpu), dev(inp["d"], gpu), dev(inp["lp"], gpu), dev(inp["ld"], gpu)
    smt = torch.from_numpy(sm).to(gpu)
    o = oracle.placed(u.type, u.params, xf.array())
    checks = [
        ("contains", host(Q.contains(geom, pid, p)), o.query(po.Q_CONTAINS, inp["p"])),
        ("inside", host(Q.inside(geom, pid, p)), o.query(po.Q_INSIDE, inp["p"])),
        ("dist_in", host(Q.distance_to_in(geom, pid, p, d, smt)), o.query(po.Q_DIST_IN, inp["p"], inp["d"], sm)),
        ("safe_in", host(Q.safety_to_in(geom, pid, p)), o.query(po.Q_SAFE_IN, inp["p"])),
        ("safe_out", host(Q.safety_to_out(geom, pid, lp)), o.query(po.Q_SAFE_OUT, inp["lp"])),
        ("placed_dist_out", host(Q.placed_distance_to_out(geom, pid, p, d, smt)),
         o.query(po.Q_PLACED_DIST_OUT, inp["p"], inp["d"], sm)),
    ]
    node = torch.empty(n, dtype=torch.int32, device=gpu)
    dout = host(Q.distance_to_out(geom, pid, lp, ld, smt, next_node=node))
    want = o.query(po.Q_DIST_OUT, inp["lp"], inp["ld"], sm)
    checks.append(("dist_out", dout, want))
    for qn, got, ref in checks:
        ok = same(got, ref)
        assert ok.all(), f"{name} {qn}: {int((~ok).sum())} of {n} lanes differ, first {np.flatnonzero(~ok)[:5]}"
    assert (host(node) == np.where(want < sm, -1, -2)).all()
    # Minimize forms (ShapeImplementationHelper.h:637-746) on a running minimum.
    cur0 = rng.uniform(0, 80, n)
    cur = torch.from_numpy(cur0.copy()).to(gpu)
    nxt = torch.full((n,), -7, dtype=torch.int32, device=gpu)
    Q.distance_to_in_minimize(geom, pid, 5, p, d, cur, nxt)
    r = o.query(po.Q_DIST_IN, inp["p"], inp["d"], cur0)
    upd = r < cur0
    assert same(host(cur), np.where(upd, r, cur0)).all()
    assert (host(nxt) == np.where(upd, 5, -7)).all()
    saf = torch.from_numpy(cur0.copy()).to(gpu)
    Q.safety_to_in_minimize(geom, pid, p, saf)
    s = o.query(po.Q_SAFE_IN, inp["p"])
    assert same(host(saf), np.where(s < cur0, s, cur0)).all()


def test_gpu_boolean_scalar_entry(gpu, oracle, booleans):
    shapes, xfs, geom, ids = booleans
    rng = np.random.Generator(np.random.Philox(31))
    for (name, u), xf, pid in zip(shapes, xfs, ids):
        o = oracle.placed(u.type, u.params, xf.array())
        for _ in range(4):
            p = rng.uniform(-40, 40, 3)
            d = W.isotropic(1, rng)[:, 0]
            got = Q.scalar(geom, pid, B.DISTANCE_TO_IN, p, d, step_max=1e30)
            want = o.query(po.Q_DIST_IN, p.reshape(3, 1), d.reshape(3, 1), np.full(1, 1e30))[0]
            assert got == want or (np.isnan(got) and np.isnan(want)), name
            assert Q.scalar(geom, pid, B.CONTAINS, p) == float(o.query(po.Q_CONTAINS, p.reshape(3, 1))[0])


def test_gpu_boolean_refused_by_batches_and_navigation(gpu, booleans):
    shapes, xfs, geom, ids = booleans
    p = dev(np.zeros((3, 4)), gpu)
    with pytest.raises(RuntimeError, match="Boolean"):
        B.Batch(geom, [B.Task(ids[0], B.CONTAINS, p)])
    with pytest.raises(RuntimeError, match="shape type 13"):
        N.locate_points(geom, p)


def test_gpu_boolean_nested_operand_refused_at_upload(gpu):
    """Operands must be primitive: a parameter block naming a Boolean operand
    is refused by the uploader (the Python constructor refuses it too)."""
    box = [G.VG_BOX, 3, 1.0, 1.0, 1.0] + list(G.Transformation3D().array())
    inner = [G.kUnion] + box + box
    bad = G.UnplacedVolume(G.VG_BOOLEAN, [G.kUnion, G.VG_BOOLEAN, len(inner)] + inner + list(
        G.Transformation3D().array()) + box)
    with pytest.raises(RuntimeError, match="primitive"):
        G.single_shape_geometry([bad], [G.Transformation3D()])
