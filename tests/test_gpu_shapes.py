"""GPU parity of the basket shape queries (HIP kernels through the C-ABI).

Bar: bit-exact agreement with the reference's outputs (golden vectors from
the reference build) and with the oracle on seeded random baskets, for every
query, including Inside classification in the 1e-9 mm tolerance band.
"""
import json
import os

import numpy as np
import pytest
import torch

import pyoracle as po
from test_shapes_hostcheck import EXTRA
from conftest import GOLDEN, agree, assert_agree, close, golden_cases, load_golden, same
from vecgeom_amd import capi
from vecgeom_amd import geometry as G
from vecgeom_amd import queries as Q
from vecgeom_amd import workloads as W

pytestmark = pytest.mark.gpu

KNOWN_DEVIATIONS = {("Torus", "rot", "dist_in")}


def dev_soa(a, device):
    return tuple(torch.from_numpy(np.ascontiguousarray(a[i])).to(device) for i in range(3))


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def run_all(geom, pid, g, device):
    p, d = dev_soa(g["p"], device), dev_soa(g["d"], device)
    lp, ld = dev_soa(g["lp"], device), dev_soa(g["ld"], device)
    sm = torch.from_numpy(g["sm"]).to(device)
    return {
        "contains": host(Q.contains(geom, pid, p)),
        "inside": host(Q.inside(geom, pid, p)),
        "dist_in": host(Q.distance_to_in(geom, pid, p, d, sm)),
        "dist_out": host(Q.distance_to_out(geom, pid, lp, ld, sm)),
        "safe_in": host(Q.safety_to_in(geom, pid, p)),
        "safe_out": host(Q.safety_to_out(geom, pid, lp)),
        "placed_dist_out": host(Q.placed_distance_to_out(geom, pid, p, d, sm)),
    }


@pytest.mark.parametrize("case", golden_cases())
def test_gpu_matches_reference_golden(gpu, case):
    g = load_golden(case)
    if not capi.lib().vg_shape_supported(int(g["type"])):
        pytest.skip("no device kernels for this shape yet")
    shape = case.split("_")[1].replace("C2", "")
    kind = case.split("_")[2][:-4]
    u = G.UnplacedVolume(int(g["type"]), g["params"])
    geom, ids, _ = G.single_shape_geometry([u], [G.Transformation3D.from_array(g["xf"])])
    got = run_all(geom, ids[0], g, gpu)
    for qn, v in got.items():
        if (shape, kind, qn) in KNOWN_DEVIATIONS:
            continue
        assert_agree(shape, qn, v, g[qn], case, torus=(g["params"], g["lp"], g["ld"]))


def test_gpu_known_answers(gpu):
    with open(os.path.join(GOLDEN, "known_answers.json")) as f:
        rows = json.load(f)["rows"]
    from test_known_answers import check
    for row in rows:
        if not capi.lib().vg_shape_supported(row["type"]):
            continue
        geom, ids, _ = G.single_shape_geometry([G.UnplacedVolume(row["type"], row["params"])], [G.Transformation3D()])
        p = tuple(torch.tensor([v], dtype=torch.float64, device=gpu) for v in row["p"])
        d = None if row["d"] is None else tuple(torch.tensor([v], dtype=torch.float64, device=gpu) for v in row["d"])
        fn = {"contains": lambda: Q.contains(geom, ids[0], p), "inside": lambda: Q.inside(geom, ids[0], p),
              "dist_in": lambda: Q.distance_to_in(geom, ids[0], p, d),
              "dist_out": lambda: Q.distance_to_out(geom, ids[0], p, d),
              "safe_in": lambda: Q.safety_to_in(geom, ids[0], p),
              "safe_out": lambda: Q.safety_to_out(geom, ids[0], p)}[row["query"]]
        got = host(fn())[0]
        assert check(row, got), f"{row['src']} {row['query']} p={row['p']} got {got} want {row['expect']}"


def shapes_under_test():
    out = []
    for name, u in W.config1_shapes() + W.zoo_shapes():
        if capi.lib().vg_shape_supported(u.type):
            out.append((name, u))
    return out


@pytest.mark.parametrize("placement", ["identity", "random"])
@pytest.mark.parametrize("n", [1, 3, 1000, 65536 + 5])
def test_gpu_matches_oracle_random(gpu, oracle, placement, n):
    """Seeded baskets of several sizes (odd sizes exercise the tail path)."""
    rng = np.random.Generator(np.random.Philox(99 + n))
    shapes = shapes_under_test()
    xfs = [G.Transformation3D() if placement == "identity" else W.random_placement(rng) for _ in shapes]
    geom, ids, _ = G.single_shape_geometry([u for _, u in shapes], xfs)
    for (name, u), xf, pid in zip(shapes, xfs, ids):
        inp = W.shape_inputs(geom.placed_extent(pid), xf, n, seed=1000 + pid + n)
        g = {"p": inp["p"], "d": inp["d"], "lp": inp["lp"], "ld": inp["ld"], "sm": np.full(n, 1e30)}
        got = run_all(geom, pid, g, gpu)
        o = oracle.placed(u.type, u.params, xf.array())
        want = {
            "contains": o.query(po.Q_CONTAINS, g["p"]), "inside": o.query(po.Q_INSIDE, g["p"]),
            "dist_in": o.query(po.Q_DIST_IN, g["p"], g["d"], g["sm"]),
            "dist_out": o.query(po.Q_DIST_OUT, g["lp"], g["ld"], g["sm"]),
            "safe_in": o.query(po.Q_SAFE_IN, g["p"]), "safe_out": o.query(po.Q_SAFE_OUT, g["lp"]),
            "placed_dist_out": o.query(po.Q_PLACED_DIST_OUT, g["p"], g["d"], g["sm"]),
        }
        for qn in want:
            # The oracle and the product share the TransformDirection form of
            # the Torus placement, so no deviation applies here.
            assert_agree(name, qn, got[qn], want[qn], f"{placement} n={n}", torus=(u.params, g["lp"], g["ld"]))


def test_gpu_unaligned_and_empty(gpu, oracle):
    """Pointers offset by one element take the 1-item-per-lane path; n=0 is a no-op."""
    name, u = shapes_under_test()[-1]
    xf = W.random_placement(np.random.Generator(np.random.Philox(5)))
    geom, ids, _ = G.single_shape_geometry([u], [xf])
    n = 5001
    inp = W.shape_inputs(geom.placed_extent(ids[0]), xf, n + 1, seed=77)
    p = tuple(torch.from_numpy(inp["p"][i]).to(gpu)[1:] for i in range(3))
    d = tuple(torch.from_numpy(inp["d"][i]).to(gpu)[1:] for i in range(3))
    out = torch.empty(n + 1, dtype=torch.float64, device=gpu)[1:]
    Q.distance_to_in(geom, ids[0], p, d, None, out=out)
    o = oracle.placed(u.type, u.params, xf.array())
    want = o.query(po.Q_DIST_IN, inp["p"][:, 1:], inp["d"][:, 1:])
    assert same(host(out), want).all()
    e = tuple(torch.empty(0, dtype=torch.float64, device=gpu) for _ in range(3))
    assert Q.contains(geom, ids[0], e).numel() == 0


def test_gpu_minimize_and_next_node(gpu, oracle):
    """DistanceToInMinimize / SafetyTo*Minimize / DistanceToOut next-node
    semantics (volumes/ShapeImplementationHelper.h:637-746)."""
    shapes = shapes_under_test()
    rng = np.random.Generator(np.random.Philox(3))
    xfs = [W.random_placement(rng) for _ in shapes]
    geom, ids, _ = G.single_shape_geometry([u for _, u in shapes], xfs)
    n = 20000
    p_np = rng.uniform(-150, 150, size=(3, n))
    d_np = W.isotropic(n, rng)
    p, d = dev_soa(p_np, gpu), dev_soa(d_np, gpu)
    cur = torch.full((n,), 1e30, dtype=torch.float64, device=gpu)
    nxt = torch.full((n,), -1, dtype=torch.int32, device=gpu)
    saf = torch.full((n,), 1e30, dtype=torch.float64, device=gpu)
    cur_np = np.full(n, 1e30)
    nxt_np = np.full(n, -1, np.int32)
    saf_np = np.full(n, 1e30)
    saf_src = np.full(n, -1)
    for k, ((name, u), xf, pid) in enumerate(zip(shapes, xfs, ids)):
        Q.distance_to_in_minimize(geom, pid, k, p, d, cur, nxt)
        Q.safety_to_in_minimize(geom, pid, p, saf)
        o = oracle.placed(u.type, u.params, xf.array())
        r = o.query(po.Q_DIST_IN, p_np, d_np, cur_np)
        if name == "Torus":
            # The chain takes the device's Torus distances, each within the
            # bar of the oracle's or of the exact first crossing.
            rd = host(Q.distance_to_in(geom, pid, p, d, torch.from_numpy(cur_np).to(gpu)))
            assert_agree(name, "dist_in", rd, r, "minimize chain",
                         torus=(u.params, W.master_to_local(xf, p_np), W.master_to_local(xf, d_np, direction=True)))
            r = rd
        upd = r < cur_np
        cur_np = np.where(upd, r, cur_np)
        nxt_np = np.where(upd, k, nxt_np)
        s = o.query(po.Q_SAFE_IN, p_np)
        saf_src = np.where(s < saf_np, k, saf_src)
        saf_np = np.where(s < saf_np, s, saf_np)
    ok = close(host(cur), cur_np) & (host(nxt) == nxt_np)
    assert ok.all(), f"{int((~ok).sum())} minimize lanes differ, first {np.flatnonzero(~ok)[:5]}"
    # Safeties: bit-exact, except lanes whose minimum comes from a shape whose
    # safety goes through libm transcendentals (held to the FP64 bar).
    got = host(saf)
    for k, (name, _) in enumerate(shapes):
        lanes = saf_src == k
        ok = agree(name, "safe_in", got[lanes], saf_np[lanes])
        assert ok.all(), f"{name} safe_in minimum: {int((~ok).sum())} lanes differ"
    lanes = saf_src < 0
    assert same(got[lanes], saf_np[lanes]).all()
    # DistanceToOut with next-node output
    name, u = shapes[0]
    sm_np = rng.uniform(0, 40, n)
    sm = torch.from_numpy(sm_np).to(gpu)
    node = torch.empty(n, dtype=torch.int32, device=gpu)
    dist = Q.distance_to_out(geom, ids[0], p, d, sm, next_node=node)
    want = oracle.placed(u.type, u.params, xfs[0].array()).query(po.Q_DIST_OUT, p_np, d_np, sm_np)
    assert same(host(dist), want).all()
    assert (host(node) == np.where(want < sm_np, -1, -2)).all()
    so = torch.full((n,), 1e30, dtype=torch.float64, device=gpu)
    Q.safety_to_out_minimize(geom, ids[0], p, so)
    s = oracle.placed(u.type, u.params, xfs[0].array()).query(po.Q_SAFE_OUT, p_np)
    assert same(host(so), np.where(s < 1e30, s, 1e30)).all()


def test_gpu_full_size_properties(gpu, oracle):
    """2^22 points per shape (the zoo basket size): Contains/Inside
    consistency everywhere and bit-exact agreement with the oracle on a
    strided sample of 65536 lanes."""
    n = 1 << 22
    shapes = shapes_under_test()
    rng = np.random.Generator(np.random.Philox(2024))
    xfs = [W.random_placement(rng) for _ in shapes]
    geom, ids, _ = G.single_shape_geometry([u for _, u in shapes], xfs)
    for (name, u), xf, pid in zip(shapes, xfs, ids):
        inp = W.shape_inputs(geom.placed_extent(pid), xf, n, seed=4242 + pid)
        p, d = dev_soa(inp["p"], gpu), dev_soa(inp["d"], gpu)
        c = Q.contains(geom, pid, p)
        ins = Q.inside(geom, pid, p)
        din = Q.distance_to_in(geom, pid, p, d)
        c_np, i_np, d_np = host(c), host(ins), host(din)
        # Contains without tolerance, Inside with +-kHalfTolerance:
        assert not np.any((c_np == 1) & (i_np == 2)), f"{name}: contained point classified outside"
        assert not np.any((c_np == 0) & (i_np == 0)), f"{name}: inside point not contained"
        sel = slice(0, n, 64)
        o = oracle.placed(u.type, u.params, xf.array())
        assert same(c_np[sel], o.query(po.Q_CONTAINS, inp["p"][:, sel])).all()
        assert same(i_np[sel], o.query(po.Q_INSIDE, inp["p"][:, sel])).all()
        assert_agree(name, "dist_in", d_np[sel], o.query(po.Q_DIST_IN, inp["p"][:, sel], inp["d"][:, sel]), "2^22",
                     torus=(u.params, inp["lp"][:, sel], inp["ld"][:, sel]))


def test_gpu_plain_variant_matches_tuned(gpu):
    """The plain FP64 form (benchmark denominator) gives identical results."""
    shapes = shapes_under_test()
    rng = np.random.Generator(np.random.Philox(8))
    xfs = [W.random_placement(rng) for _ in shapes]
    geom, ids, _ = G.single_shape_geometry([u for _, u in shapes], xfs)
    n = 30001
    try:
        for (name, u), xf, pid in zip(shapes, xfs, ids):
            inp = W.shape_inputs(geom.placed_extent(pid), xf, n, seed=55 + pid)
            g = {"p": inp["p"], "d": inp["d"], "lp": inp["lp"], "ld": inp["ld"], "sm": np.full(n, 1e30)}
            capi.call("vg_set_kernel_variant", 0)
            tuned = run_all(geom, pid, g, gpu)
            capi.call("vg_set_kernel_variant", 1)
            p, d = dev_soa(g["p"], gpu), dev_soa(g["d"], gpu)
            lp, ld = dev_soa(g["lp"], gpu), dev_soa(g["ld"], gpu)
            sm = torch.from_numpy(g["sm"]).to(gpu)
            plain = {
                "contains": host(Q.contains(geom, pid, p)), "inside": host(Q.inside(geom, pid, p)),
                "dist_in": host(Q.distance_to_in(geom, pid, p, d, sm)),
                "dist_out": host(Q.distance_to_out(geom, pid, lp, ld, sm)),
                "safe_in": host(Q.safety_to_in(geom, pid, p)), "safe_out": host(Q.safety_to_out(geom, pid, lp)),
            }
            for qn, v in plain.items():
                assert same(v, tuned[qn]).all(), f"{name} {qn}: plain form differs from tuned kernels"
    finally:
        capi.call("vg_set_kernel_variant", 0)


@pytest.mark.parametrize("name", [n for n, _ in EXTRA])
def test_gpu_extra_variants_vs_oracle(gpu, oracle, name):
    """Phi-sectioned and degenerate variants (Cone, Tube, Polycone, Torus,
    Polyhedron -- one with a parameter block beyond the LDS stage, read from
    global memory), random placement, all six queries against the oracle."""
    u = dict(EXTRA)[name]
    xf = W.random_placement(np.random.Generator(np.random.Philox(17)))
    geom, ids, _ = G.single_shape_geometry([u], [xf])
    n = 20000
    inp = W.shape_inputs(geom.placed_extent(ids[0]), xf, n, seed=606)
    g = {"p": inp["p"], "d": inp["d"], "lp": inp["lp"], "ld": inp["ld"], "sm": np.full(n, 1e30)}
    got = run_all(geom, ids[0], g, gpu)
    o = oracle.placed(u.type, u.params, xf.array())
    family = "Torus" if u.type == G.VG_TORUS else ("Sphere" if u.type == G.VG_SPHERE else name)
    tor = (u.params, g["lp"], g["ld"])
    for qn, q in (("contains", po.Q_CONTAINS), ("inside", po.Q_INSIDE), ("safe_in", po.Q_SAFE_IN)):
        assert_agree(family, qn, got[qn], o.query(q, g["p"]), name)
    assert_agree(family, "dist_in", got["dist_in"], o.query(po.Q_DIST_IN, g["p"], g["d"], g["sm"]), name, torus=tor)
    assert_agree(family, "dist_out", got["dist_out"], o.query(po.Q_DIST_OUT, g["lp"], g["ld"], g["sm"]), name,
                 torus=tor)
    assert_agree(family, "safe_out", got["safe_out"], o.query(po.Q_SAFE_OUT, g["lp"]), name)


def test_gpu_contains_local_unplaced_contains_and_scalar(gpu, oracle):
    """Contains(point, localPoint&) and UnplacedContains in basket form on
    every zoo shape under a random placement, against the oracle (Contains
    of the placed shape; the local point is the placement's Transform), and
    the scalar entry point on a sample of the same points against the
    basket results of every query."""
    from vecgeom_amd import batch as B
    import pyoracle as po
    shapes = [(n, u) for n, u in W.zoo_shapes() if Q.capi.lib().vg_shape_supported(u.type)]
    rng = np.random.Generator(np.random.Philox(4711))
    xfs = [W.random_placement(rng) for _ in shapes]
    geom, ids, _ = G.single_shape_geometry([u for _, u in shapes], xfs)
    for (name, u), xf, pid in zip(shapes, xfs, ids):
        inp = W.shape_inputs(geom.placed_extent(pid), xf, 5000, seed=pid + 4712)
        p = tuple(torch.from_numpy(np.ascontiguousarray(inp["p"][i])).to(gpu) for i in range(3))
        c, loc = Q.contains_local(geom, pid, p)
        want = oracle.placed(u.type, u.params, xf.array()).query(po.Q_CONTAINS, inp["p"])
        assert (c.cpu().numpy() == want).all(), name
        got_loc = np.stack([a.cpu().numpy() for a in loc])
        assert (got_loc == inp["lp"]).all(), name  # the reference's operation order, bit for bit
        uc = Q.unplaced_contains(geom, pid, loc)
        assert (uc.cpu().numpy() == want).all(), name
        dirs = tuple(torch.from_numpy(np.ascontiguousarray(inp["d"][i])).to(gpu) for i in range(3))
        basket = {B.CONTAINS: Q.contains(geom, pid, p).cpu().numpy(),
                  B.INSIDE: Q.inside(geom, pid, p).cpu().numpy(),
                  B.DISTANCE_TO_IN: Q.distance_to_in(geom, pid, p, dirs).cpu().numpy(),
                  B.SAFETY_TO_IN: Q.safety_to_in(geom, pid, p).cpu().numpy()}
        for i in range(0, 5000, 997):
            for q, arr in basket.items():
                r = Q.scalar(geom, pid, q, inp["p"][:, i], inp["d"][:, i])
                assert (r == float(arr[i])) or (np.isnan(r) and np.isnan(arr[i])), (name, q, i)
            r, lp = Q.scalar(geom, pid, Q.CONTAINS_LOCAL, inp["p"][:, i])
            assert r == float(want[i]) and lp == tuple(inp["lp"][:, i]), (name, i)
    geom.free()


@pytest.mark.parametrize("name", ["Torus", "TorusPhi"])
def test_gpu_torus_exact_crossings(gpu, oracle, name):
    """2^20 rays: on every ray the query is defined for (DistanceToIn from
    outside, DistanceToOut from inside) the device's Torus distance lies
    within the surface tolerance of the exact first crossing (binary128
    window, oracle/vgo_truth.cpp), under a random placement; every lane is
    within the FP64 bar of the oracle or passes that check."""
    u = dict(W.zoo_shapes() + EXTRA)[name]
    xf = W.random_placement(np.random.Generator(np.random.Philox(71)))
    geom, ids, _ = G.single_shape_geometry([u], [xf])
    n = 1 << 20
    inp = W.shape_inputs(geom.placed_extent(ids[0]), xf, n, seed=7171)
    p, d = dev_soa(inp["p"], gpu), dev_soa(inp["d"], gpu)
    lp, ld = dev_soa(inp["lp"], gpu), dev_soa(inp["ld"], gpu)
    o = oracle.placed(u.type, u.params, xf.array())
    for qn, out, got, want in (
            ("dist_in", False, host(Q.distance_to_in(geom, ids[0], p, d)), o.query(po.Q_DIST_IN, inp["p"], inp["d"],
                                                                                  nthreads=8)),
            ("dist_out", True, host(Q.distance_to_out(geom, ids[0], lp, ld)),
             o.query(po.Q_DIST_OUT, inp["lp"], inp["ld"], nthreads=8))):
        assert_agree("Torus", qn, got, want, name, torus=(u.params, inp["lp"], inp["ld"]))
        lo, _ = po.torus_window(u.params, inp["lp"], inp["ld"], out)
        sel = ~np.isnan(lo)
        ok = po.torus_distance_ok(u.params, inp["lp"][:, sel], inp["ld"][:, sel], got[sel], out)
        assert ok.all(), f"{name} {qn}: {int((~ok).sum())} of {int(sel.sum())} defined lanes outside the exact window"
