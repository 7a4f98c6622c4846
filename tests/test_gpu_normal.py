"""GPU parity of the surface normals (vecgeom_amd/csrc/shapes_normal.hpp,
vg_normal.hip) through the C-ABI (vg_normal, vg_distance_to_out_normal):
against the reference's golden vectors and against the oracle on seeded
random and surface points, bit for bit except where a Sphere output comes
from point-dependent sin/cos (its approximate normal) or from its
DistanceToOut (within the FP64 distance bar); and the refusals."""
import math

import numpy as np
import pytest
import torch

import pyoracle as po
from conftest import GOLDEN, close, same
from vecgeom_amd import geometry as G
from vecgeom_amd import queries as Q
from vecgeom_amd import workloads as W

pytestmark = pytest.mark.gpu

CASES = sorted(f for f in __import__("os").listdir(GOLDEN) if f.startswith("normal_") and f.endswith(".npz"))
I = G.Transformation3D()
# The Sphere's approximate normal (a point on no surface) takes cos and sin
# of the point's phi: device math against glibc, an ulp or two.
TRIG = 1e-15
# A normal at a DistanceToOut exit point moves with the distance: the
# Sphere's distances are held to the FP64 bar (1e-9 mm), its exit normals to
# this.
EXIT = 1e-12


def dev(a, device):
    return tuple(torch.from_numpy(np.ascontiguousarray(a[i])).to(device) for i in range(3))


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def load(name):
    with np.load(f"{GOLDEN}/{name}", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def unplaced(g):
    return G.UnplacedVolume(int(g["type"]), list(g["params"]))


def check_normals(t, got, want, gv, wv, what, tol=TRIG):
    assert (gv == wv).all(), f"{what}: valid differs on {int((gv != wv).sum())} of {len(wv)} points"
    ok = same(got, want).all(axis=0)
    if t == G.VG_SPHERE:
        ok |= (np.abs(got - want) <= tol).all(axis=0) & ((wv == 0) | (tol == EXIT))
    assert ok.all(), f"{what}: normal differs on {int((~ok).sum())} points, first {np.flatnonzero(~ok)[:5]}"


@pytest.mark.parametrize("case", CASES)
def test_gpu_normal_vs_reference_goldens(gpu, case):
    g = load(case)
    t = int(g["type"])
    geom, ids, _ = G.single_shape_geometry([unplaced(g)], [I])
    nrm, valid = Q.normal(geom, ids[0], dev(g["lp"], gpu))
    got = np.stack([host(a) for a in nrm])
    check_normals(t, got, g["normal"], host(valid), g["valid"], case)
    sm = torch.from_numpy(g["step_max"]).to(gpu)
    dist, dn, dv, dc = Q.distance_to_out_normal(geom, ids[0], dev(g["rp"], gpu), dev(g["rd"], gpu), sm)
    dist = host(dist)
    if t == G.VG_SPHERE:
        assert close(dist, g["dist"]).all()
        check_normals(t, np.stack([host(a) for a in dn]), g["dist_normal"], host(dv), g["dist_valid"],
                      case + " exit", tol=EXIT)
    else:
        assert same(dist, g["dist"]).all(), f"{case}: {int((~same(dist, g['dist'])).sum())} distances differ"
        check_normals(t, np.stack([host(a) for a in dn]), g["dist_normal"], host(dv), g["dist_valid"],
                      case + " exit")
    assert (host(dc) == 1).all()


@pytest.mark.parametrize("case", CASES)
def test_gpu_normal_vs_oracle_random(gpu, oracle, case):
    """Seeded points the goldens do not hold: uniform in 1.25x the extent,
    exact exit points of interior rays, and those jittered by 3e-10 mm."""
    g = load(case)
    t, prm = int(g["type"]), g["params"]
    o = oracle.placed(t, prm, I.array())
    lo, hi = o.extent()
    rng = np.random.Generator(np.random.Philox(8800 + CASES.index(case)))
    n = 1 << 15
    uni = W.sample_points(lo, hi, n, rng, scale=1.25)
    c = W.sample_points(lo, hi, 8 * n, rng, scale=1.0)
    c = c[:, o.query(po.Q_CONTAINS, c).astype(bool)][:, :n]
    d = W.isotropic(c.shape[1], rng)
    dist_o, dn_o, dv_o, dc_o = o.distance_to_out_normal(c, d)
    hit = c + d * np.where(np.isfinite(dist_o), dist_o, 0.0)
    pts = np.concatenate([uni, hit, hit + rng.normal(0.0, 3e-10, hit.shape)], axis=1)
    geom, ids, _ = G.single_shape_geometry([unplaced(g)], [I])
    nrm, valid = Q.normal(geom, ids[0], dev(pts, gpu))
    want, wv = o.normal(pts)
    check_normals(t, np.stack([host(a) for a in nrm]), want, host(valid), wv, case + " random")
    dist, dn, dv, dc = Q.distance_to_out_normal(geom, ids[0], dev(c, gpu), dev(d, gpu))
    dist = host(dist)
    if t == G.VG_SPHERE:
        assert close(dist, dist_o).all()
        check_normals(t, np.stack([host(a) for a in dn]), dn_o, host(dv), dv_o, case + " exit", tol=EXIT)
    else:
        assert same(dist, dist_o).all()
        check_normals(t, np.stack([host(a) for a in dn]), dn_o, host(dv), dv_o, case + " exit")


def test_gpu_normal_refused_and_empty(gpu):
    shapes = [G.UnplacedTorus(5, 10, 50, 0, 2 * math.pi), G.UnplacedParaboloid(10, 30, 40), G.UnplacedBox(1, 2, 3)]
    geom, ids, _ = G.single_shape_geometry(shapes, [I, I, I])
    p = dev(np.zeros((3, 4)), gpu)
    for pid in ids[:2]:
        with pytest.raises(RuntimeError, match="no Normal"):
            Q.normal(geom, pid, p)
        with pytest.raises(RuntimeError, match="no Normal"):
            Q.distance_to_out_normal(geom, pid, p, p)
    e = dev(np.zeros((3, 0)), gpu)
    nrm, valid = Q.normal(geom, ids[2], e)
    assert valid.numel() == 0
    # A step_max array is read per lane: DistanceToOut of a box is not
    # limited by it, as the reference's is not.
    sm = torch.full((4,), 0.5, dtype=torch.float64, device=gpu)
    dirs = dev(np.array([[1.0] * 4, [0.0] * 4, [0.0] * 4]), gpu)
    dist, dn, dv, dc = Q.distance_to_out_normal(geom, ids[2], p, dirs, sm)
    assert (host(dist) == 1.0).all() and (host(dn[0]) == 1.0).all() and (host(dv) == 1).all()
