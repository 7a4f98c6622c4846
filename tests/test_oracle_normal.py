"""Surface normals in the oracle (Shape::Normal and DistanceToOutNormal in
oracle/vgo*.cpp): against the reference's golden vectors
(tests/golden/normal_*.npz, written by make_normal_golden.py from the real
reference build), against the live reference build where it is present, and
against closed forms that fix what the reference's normals are."""
import math

import numpy as np
import pytest

import pyoracle as po
from conftest import GOLDEN, same
from vecgeom_amd import geometry as G

I = G.Transformation3D().array()
CASES = sorted(f for f in __import__("os").listdir(GOLDEN) if f.startswith("normal_") and f.endswith(".npz"))
# The reference's compiler fuses the cos and sin of one angle into glibc's
# sincos, which can differ from them by an ulp: the Sphere's approximate
# normal (a point on no surface) is compared within 1e-15 on each component.
SINCOS_ULP = 1e-15


def load(name):
    with np.load(f"{GOLDEN}/{name}", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def normals_agree(shape_type, got, want, valid):
    ok = same(got, want).all(axis=0)
    if shape_type == G.VG_SPHERE:
        near = (np.abs(got - want) <= SINCOS_ULP).all(axis=0) & (valid == 0)
        ok |= near
    return ok


def test_normal_goldens_present():
    assert len(CASES) >= 13


@pytest.mark.parametrize("case", CASES)
def test_oracle_normal_matches_reference_goldens(oracle, case):
    g = load(case)
    h = oracle.placed(int(g["type"]), g["params"], I)
    nrm, valid = h.normal(g["lp"])
    assert (valid == g["valid"]).all(), f"{case}: valid differs on {int((valid != g['valid']).sum())} points"
    ok = normals_agree(int(g["type"]), nrm, g["normal"], g["valid"])
    assert ok.all(), f"{case}: normal differs on {int((~ok).sum())} points, first {np.flatnonzero(~ok)[:5]}"
    dist, dn, dv, dc = h.distance_to_out_normal(g["rp"], g["rd"], g["step_max"])
    assert same(dist, g["dist"]).all()
    assert (dv == g["dist_valid"]).all() and (dc == g["convex"]).all() and (dc == 1).all()
    assert normals_agree(int(g["type"]), dn, g["dist_normal"], g["dist_valid"]).all()


@pytest.mark.parametrize("case", CASES)
def test_oracle_normal_against_live_reference(oracle, reference, case):
    """On fresh random points (the reference build's answers), where the Cone's
    kernel would not assert: it asserts for points on no surface in the strict
    build, so those go to the Release build."""
    g = load(case)
    t, prm = int(g["type"]), g["params"]
    o = oracle.placed(t, prm, I)
    lo, hi = o.extent()
    rng = np.random.default_rng(17)
    pts = rng.uniform(lo - 0.25 * (hi - lo), hi + 0.25 * (hi - lo), size=(4096, 3)).T.copy()
    pts = np.concatenate([pts, g["lp"] + rng.normal(0, 3e-10, g["lp"].shape)], axis=1)
    nrm, valid = o.normal(pts)
    sel = valid.astype(bool) if t == G.VG_CONE else np.ones(pts.shape[1], bool)
    rn, rv = reference.placed(t, prm, I).normal(pts[:, sel])
    assert (rv == valid[sel]).all()
    assert normals_agree(t, nrm[:, sel], rn, rv).all()
    if t == G.VG_CONE:
        fn, fv = po.reference(fast=True).placed(t, prm, I).normal(pts[:, ~sel])
        assert (fv == 0).all() and same(fn, nrm[:, ~sel]).all()


def q(u, p):
    n, v = po.oracle().placed(u.type, u.params, I).normal(np.asarray(p, np.float64).reshape(3, 1))
    return n[:, 0], int(v[0])


def test_normal_closed_forms():
    box = G.UnplacedBox(10, 15, 20)
    n, v = q(box, [10, 3, 4])
    assert v == 1 and (n == [1, 0, 0]).all()
    n, v = q(box, [-10, 15, 4])  # an edge: (-1, 1, 0)/sqrt(2)
    assert v == 1 and np.allclose(n, np.array([-1, 1, 0]) / math.sqrt(2), atol=1e-16)
    n, v = q(box, [10, -15, 20])  # a corner
    assert v == 1 and np.allclose(n, np.array([1, -1, 1]) / math.sqrt(3), atol=1e-16)
    n, v = q(box, [0, 0, 0])
    assert v == 0 and (n == 0).all()
    orb = G.UnplacedOrb(50)
    n, v = q(orb, [30, 40, 0])
    assert v == 1 and (n == [0.6, 0.8, 0]).all()
    assert q(orb, [3, 4, 0])[1] == 0
    tube = G.UnplacedTube(10, 20, 30, 0.0, 2 * math.pi)
    n, v = q(tube, [0, 20, 5])  # outer wall
    assert v == 1 and (n == [0, 1, 0]).all()
    n, v = q(tube, [-10, 0, 5])  # inner wall points to the axis
    assert v == 1 and (n == [1, 0, 0]).all()
    n, v = q(tube, [15, 0, -30])  # bottom lid
    assert v == 1 and (n == [0, 0, -1]).all()
    n, v = q(tube, [20, 0, 30])  # outer rim: (1, 0, 1)/sqrt(2)
    assert v == 1 and np.allclose(n, np.array([1, 0, 1]) / math.sqrt(2), atol=1e-16)
    wedge = G.UnplacedTube(0, 20, 30, 0.0, math.pi / 2)
    n, v = q(wedge, [10, 0, 0])  # the phi = 0 plane: outward -y
    assert v == 1 and np.allclose(n, [0, -1, 0], atol=1e-16)
    cone = G.UnplacedCone(5, 20, 10, 40, 30, 0, 2 * math.pi)
    n, v = q(cone, [0, 0, 30])  # on the axis: only the z = dz plane is within tolerance
    assert v == 1 and (n == [0, 0, 1]).all()
    n, v = q(cone, [15, 0, 30])
    assert v == 1 and (n == [0, 0, 1]).all()
    # Outer cone at z = 0: radius 30, slope (40 - 20) / 60 -> normal (1, 0, -1/3)/|.|.
    n, v = q(cone, [30, 0, 0])
    assert v == 1 and np.allclose(n, np.array([3, 0, -1]) / math.sqrt(10), atol=1e-15)
    trap = G.UnplacedTrapezoid(40.0, 0.0, 0.0, 20.0, 15.0, 15.0, 0.0, 20.0, 15.0, 15.0, 0.0)  # a box 15 x 20 x 40
    n, v = q(trap, [15, 0, 0])
    assert v == 1 and np.allclose(n, [1, 0, 0], atol=1e-15)
    n, v = q(trap, [0, 0, -40])
    assert v == 1 and (n == [0, 0, -1]).all()
    sph = G.UnplacedSphere(10, 50, 0, 2 * math.pi, 0, math.pi)
    n, v = q(sph, [0, 0, -10])  # inner shell points to the centre
    assert v == 1 and (n == [0, 0, 1]).all()
    pc = G.UnplacedPolycone(0.0, 2 * math.pi, rmin=[5, 5], rmax=[20, 20], z=[-10, 10])
    n, v = q(pc, [0, 0, -11])  # below the sections
    assert v == 1 and (n == [0, 0, -1]).all()
    n, v = q(pc, [0, 0, 11])
    assert v == 1 and (n == [0, 0, 1]).all()


def test_trd_normal_is_the_reference_code_as_it_runs():
    """UnplacedTrd::Normal's un-braced ifs (source/UnplacedTrd.cpp:191-196)
    set y and z from the last (-z) line, x from the last face test that holds,
    and count six surfaces whatever the point.  One of the +y and -y tests,
    (y - dy2) and (y + dy2) times a nonzero factor, always holds for a finite
    point, and their x-component is 0: the normal is (0, 0, 4 dx2 dy1) and
    valid everywhere."""
    trd = G.UnplacedTrd(15.0, 20.0, 25.0, 30.0, 35.0)
    for p in ([0, 0, 0], [17, 3, 0], [100, 100, 100], [20, 30, 35], [-20, -30, -35], [0, 30, 35]):
        n, v = q(trd, p)
        assert v == 1 and (n == [0.0, 0.0, 4 * 20.0 * 25.0]).all(), p


def test_dist_out_normal_closed_form():
    box = G.UnplacedBox(30, 40, 50)
    h = po.oracle().placed(box.type, box.params, I)
    p = np.array([[0.0], [0.0], [0.0]])
    d = np.array([[1.0], [0.0], [0.0]])
    dist, n, v, c = h.distance_to_out_normal(p, d)
    assert dist[0] == 30.0 and (n[:, 0] == [1, 0, 0]).all() and v[0] == 1 and c[0] == 1


@pytest.mark.parametrize("shape", [G.UnplacedParaboloid(10, 30, 40), G.UnplacedTorus(5, 10, 50, 0, 2 * math.pi),
                                   G.UnplacedParallelepiped(20, 30, 40, 30, 15, 30)])
def test_normal_refused_where_the_reference_has_none(shape):
    with pytest.raises(ValueError):
        po.oracle().placed(shape.type, shape.params, I).normal(np.zeros((3, 1)))
