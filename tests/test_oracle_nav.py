"""The navigation oracle (oracle/vgo_nav.cpp) against the reference's outputs.

Golden vectors come from the real reference navigators
(tests/golden/make_golden_nav.py).  Where the reference build is present the
oracle is also compared with it on fresh random tracks.  Closed-form cases
on the nested geometry pin the oracle without the reference: rays from the
origin along the axes, located points at known volume centres, and the
safety at the origin.
"""
import numpy as np
import pytest

import pyoracle as po
from conftest import load_golden, same
from vecgeom_amd import workloads as W

KINDS = ((0, "abbox"), (1, "simple"))
ZOO_REF_NAMES = [n for n, _ in W.NAV_ZOO if n != "Parallelepiped"]


@pytest.fixture(scope="module")
def nested():
    return W.nested_geometry()


@pytest.fixture(scope="module")
def nested_nav(oracle, nested):
    return po.Nav(oracle, nested[1])


def compare(nav, g):
    level, path = nav.locate(g["x"])
    assert (level == g["level"]).all()
    assert (path == g["path"]).all()
    for kind, tag in KINDS:
        lo, pa, bo, st = nav.step(g["x"], g["d"], g["level"], g["path"], kind=kind)
        assert (lo == g[f"step_{tag}_level"]).all(), tag
        assert (pa == g[f"step_{tag}_path"]).all(), tag
        assert (bo == g[f"step_{tag}_boundary"]).all(), tag
        assert same(st, g[f"step_{tag}_step"]).all(), tag
        ns, fp, tr, hs = nav.trace(g["x"], g["d"], kind=kind)
        assert (ns == g[f"trace_{tag}_nsteps"]).all(), tag
        assert (hs == g[f"trace_{tag}_hash"]).all(), tag
        assert same(fp, g[f"trace_{tag}_final"]).all(), tag
        assert same(tr, g[f"trace_{tag}_travelled"]).all(), tag
    assert same(nav.safety(g["x"], g["level"], g["path"]), g["safety"]).all()


def test_nested_golden(oracle, nested_nav):
    compare(nested_nav, load_golden("nav_nested.npz"))


def test_zoo_golden(oracle):
    _, flat, _ = W.nav_zoo_geometry(names=ZOO_REF_NAMES)
    compare(po.Nav(oracle, flat), load_golden("nav_zoo.npz"))


@pytest.mark.parametrize("geometry", ["nested", "zoo"])
def test_against_live_reference(oracle, reference, geometry):
    if geometry == "nested":
        _, flat = W.nested_geometry()
        rng = np.random.Generator(np.random.Philox(991))
        n = 3000
        r = rng.uniform(0.0, 1000.0, n)
        ph = rng.uniform(0, 2 * np.pi, n)
        p = np.stack([r * np.cos(ph), r * np.sin(ph), rng.uniform(-1900, 1900, n)])
        d = W.isotropic(n, rng)
    else:
        _, flat, placed = W.nav_zoo_geometry(names=ZOO_REF_NAMES)
        centres = [flat.placed_transform[12 * i:12 * i + 3] for _, i in placed]
        p, d = W.nav_zoo_tracks(3000, centres, seed=992)
    a, b = po.Nav(oracle, flat), po.Nav(reference, flat)
    la, pa = a.locate(p)
    lb, pb = b.locate(p)
    assert (la == lb).all() and (pa == pb).all()
    for kind, _ in KINDS:
        for x, y in zip(a.step(p, d, la, pa, kind=kind), b.step(p, d, lb, pb, kind=kind)):
            assert same(x, y).all()
        for x, y in zip(a.trace(p, d, kind=kind), b.trace(p, d, kind=kind)):
            assert same(x, y).all()
    assert same(a.safety(p, la, pa), b.safety(p, lb, pb)).all()


def fold(ids_boundary):
    """The trace hash of a sequence of (top placed id, boundary flag)."""
    h = 1469598103934665603
    for top, bnd in ids_boundary:
        v = (top + 1) * 2 + bnd
        h = ((h ^ v) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def shells(gm):
    return [pv.id for pv in gm.world.logical.daughters]


@pytest.mark.parametrize("kind", [0, 1])
def test_ray_along_x_crosses_every_shell(nested, nested_nav, kind):
    gm, _ = nested
    p = np.zeros((3, 1))
    d = np.array([[1.0], [0.0], [0.0]])
    ns, fp, tr, hs = nested_nav.trace(p, d, kind=kind)
    # z = 0 lies between the box slices, so the ray meets the 64 shells only:
    # world -> shell 0 -> world -> ... -> shell 63 -> world -> outside.
    seq = []
    for s in shells(gm):
        seq += [(s, 1), (gm.world.id, 1)]
    seq.append((-1, 1))
    assert ns[0] == 129
    assert hs[0] == fold(seq)
    # The last step ends on the world face x = 1000 before the 1e-6 push.
    assert abs(fp[0, 0] - (1000.0 + 1e-6)) < 1e-9
    assert abs(tr[0] + 128e-6 - 1000.0) < 1e-9


def test_ray_along_z_leaves_world_in_one_step(nested_nav):
    ns, fp, tr, hs = nested_nav.trace(np.zeros((3, 1)), np.array([[0.0], [0.0], [1.0]]))
    assert ns[0] == 1 and tr[0] == 2000.0
    assert hs[0] == fold([(-1, 1)])


def test_first_step_from_origin_reaches_innermost_shell(nested, nested_nav):
    gm, _ = nested
    p = np.zeros((3, 4))
    ang = np.array([0.1, 1.0, 2.5, 4.0])
    d = np.stack([np.cos(ang), np.sin(ang), np.zeros(4)])
    level, path = nested_nav.locate(p)
    for kind, _ in KINDS:
        lo, pa, bo, st = nested_nav.step(p, d, level, path, kind=kind)
        assert np.allclose(st, 100.0, rtol=0, atol=1e-12)
        assert (lo == 2).all() and (pa[1] == shells(gm)[0]).all() and bo.all()


def test_locate_trd_centres(nested, nested_nav):
    gm, flat = nested
    world = gm.world
    pts, want = [], []
    for s in world.logical.daughters[::9]:
        for b in s.logical.daughters[::5]:
            t = b.transformation.t
            for trd in b.logical.daughters[::3]:
                pts.append([t[0], t[1], t[2] + trd.transformation.t[2]])
                want.append([world.id, s.id, b.id, trd.id])
    p = np.array(pts).T
    level, path = nested_nav.locate(p)
    assert (level == 4).all()
    assert (path.T == np.array(want)).all()


def test_safety_at_origin(nested_nav):
    level, path = nested_nav.locate(np.zeros((3, 1)))
    assert nested_nav.safety(np.zeros((3, 1)), level, path)[0] == 100.0


def baskets_by_volume(level, path, count=4):
    """Indices of the tracks of the `count` most common current volumes: the
    basket navigator assumes a basket shares one current volume."""
    top = path[np.maximum(level - 1, 0), np.arange(level.size)]
    top = np.where(level > 0, top, -1)
    vols, counts = np.unique(top[top >= 0], return_counts=True)
    order = np.argsort(-counts)[:count]
    return [np.flatnonzero(top == v) for v in vols[order]]


@pytest.mark.parametrize("geometry", ["nested", "zoo"])
@pytest.mark.parametrize("pstep", [np.inf, 7.5])
def test_basket_navigator_against_live_reference(oracle, reference, geometry, pstep):
    """SimpleNavigator's basket FindNextBoundaryAndStep (navigator kind 3,
    SimpleNavigator.h:638-743): the oracle against the reference's own basket
    interface on baskets of tracks sharing a current volume."""
    if geometry == "nested":
        _, flat = W.nested_geometry()
        rng = np.random.Generator(np.random.Philox(1991))
        n = 4000
        r = rng.uniform(0.0, 1000.0, n)
        ph = rng.uniform(0, 2 * np.pi, n)
        p = np.stack([r * np.cos(ph), r * np.sin(ph), rng.uniform(-1900, 1900, n)])
        d = W.isotropic(n, rng)
    else:
        _, flat, placed = W.nav_zoo_geometry(names=ZOO_REF_NAMES)
        centres = [flat.placed_transform[12 * i:12 * i + 3] for _, i in placed]
        p, d = W.nav_zoo_tracks(4000, centres, seed=1992)
    a, b = po.Nav(oracle, flat), po.Nav(reference, flat)
    level, path = a.locate(p)
    for idx in baskets_by_volume(level, path):
        q, v = p[:, idx].copy(), d[:, idx].copy()
        lo, pa = level[idx].copy(), path[:, idx].copy()
        got = a.step(q, v, lo, pa, pstep=pstep, kind=3)
        want = b.step(q, v, lo, pa, pstep=pstep, kind=3)
        for x, y in zip(got, want):
            assert same(x, y).all(), (geometry, len(idx))
        assert (got[3] <= pstep).all()


def test_basket_navigator_differs_from_scalar_where_the_reference_does(oracle, reference):
    """States in the world whose points lie outside the world box: the
    world's DistanceToOut is negative there, which the scalar SimpleNavigator
    repairs (step <= 0 -> kInfinity -> leave the world with step 0) and the
    basket navigator does not (it steps by the negative distance, pushes by
    1e-9 and relocates).  Oracle and reference agree on both, and the two
    navigators differ wherever that distance is negative."""
    _, flat = W.nested_geometry()
    rng = np.random.Generator(np.random.Philox(77))
    n = 64
    p = np.stack([rng.uniform(1001, 1100, n), rng.uniform(-900, 900, n), rng.uniform(-1900, 1900, n)])
    d = W.isotropic(n, rng)
    world = flat.world
    level = np.ones(n, np.int32)
    depth = po.Nav(oracle, flat).depth
    path = np.full((depth, n), -1, np.int32)
    path[0] = world
    a, b = po.Nav(oracle, flat), po.Nav(reference, flat)
    ob, rb = a.step(p, d, level, path, kind=3), b.step(p, d, level, path, kind=3)
    for x, y in zip(ob, rb):
        assert same(x, y).all()
    os_, rs = a.step(p, d, level, path, kind=1), b.step(p, d, level, path, kind=1)
    for x, y in zip(os_, rs):
        assert same(x, y).all()
    neg = ob[3] < 0  # the world's DistanceToOut is negative on these tracks
    assert neg.sum() > 8
    assert (os_[3][neg] == 0).all() and (os_[0][neg] == 0).all()  # scalar: step 0, out of the world


def test_basket_navigator_closed_form(nested, nested_nav):
    """A basket of tracks in the world at radii below the innermost shell,
    moving radially outwards: each steps to the shell's inner radius and is
    located in it."""
    gm, _ = nested
    sh = shells(gm)
    rmin0 = 100.0
    n = 16
    ph = np.linspace(0, 2 * np.pi, n, endpoint=False)
    r0 = np.linspace(1.0, 90.0, n)
    p = np.stack([r0 * np.cos(ph), r0 * np.sin(ph), np.linspace(-500, 500, n)])
    d = np.stack([np.cos(ph), np.sin(ph), np.zeros(n)])
    level, path = nested_nav.locate(p)
    assert (level == 1).all()
    lo, pa, bo, st = nested_nav.step(p, d, level, path, kind=3)
    np.testing.assert_allclose(st, rmin0 - r0, rtol=0, atol=1e-9)
    assert (bo == 1).all() and (lo == 2).all() and (pa[1] == sh[0]).all()
