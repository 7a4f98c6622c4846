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
