"""The device navigation code, compiled for the host with bounds checks
(tests/native/navhost.cpp over vecgeom_amd/csrc/vg_nav_impl.hpp), against
the oracle.  This checks the navigator's logic and indexing on the CPU; the
GPU tests check the same code on the device."""
import ctypes
import os

import numpy as np
import pytest

import pyoracle as po
from conftest import ROOT, same
from vecgeom_amd import workloads as W

LIB = os.path.join(ROOT, "tests", "native", "libnavhost.so")


class _HostNav:
    def __init__(self):
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is missing: build it with __graft_entry__.build()")
        self.lib = ctypes.CDLL(LIB)
        self.prefix = "navhost_"


@pytest.fixture(scope="module")
def hostnav():
    return _HostNav()


def cases():
    _, flat = W.nested_geometry()
    p1, d1 = W.track_starts(6000, seed=1301)
    rng = np.random.Generator(np.random.Philox(1302))
    n = 6000
    r = rng.uniform(0, 1000, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    p2 = np.stack([r * np.cos(ph), r * np.sin(ph), rng.uniform(-1950, 1950, n)])
    d2 = W.isotropic(n, rng)
    _, zflat, placed = W.nav_zoo_geometry()
    centres = [zflat.placed_transform[12 * i:12 * i + 3] for _, i in placed]
    p3, d3 = W.nav_zoo_tracks(6000, centres, seed=1303)
    return {"nested_vertex": (flat, p1, d1), "nested_scatter": (flat, p2, d2), "zoo": (zflat, p3, d3)}


CASES = cases()


@pytest.mark.parametrize("case", sorted(CASES))
def test_host_build_matches_oracle(oracle, hostnav, case):
    flat, p, d = CASES[case]
    O, H = po.Nav(oracle, flat), po.Nav(hostnav, flat)
    lo, pa = O.locate(p, nthreads=8)
    lh, ph = H.locate(p, nthreads=8)
    assert (lo == lh).all() and (pa == ph).all()
    # kind 2 (the mixed-precision ABBox navigator) must equal the oracle's
    # ABBox navigator (kind 0) bit for bit.
    for kind, hkind in ((0, 0), (1, 1), (0, 2)):
        for x, y in zip(O.step(p, d, lo, pa, kind=kind, nthreads=8), H.step(p, d, lo, pa, kind=hkind, nthreads=8)):
            assert same(x, y).all(), hkind
        for x, y in zip(O.trace(p, d, kind=kind, nthreads=8), H.trace(p, d, kind=hkind, nthreads=8)):
            assert same(x, y).all(), hkind
    assert same(O.safety(p, lo, pa, nthreads=8), H.safety(p, lo, pa, nthreads=8)).all()


@pytest.mark.parametrize("case", sorted(CASES))
def test_exact_walk_matches_oracle(oracle, hostnav, case):
    """The ABBox candidate walk in exact sorted order alone (navigator kind
    9 of the host build) agrees with the oracle's sorted walk."""
    flat, p, d = CASES[case]
    O, H = po.Nav(oracle, flat), po.Nav(hostnav, flat)
    lo, pa = O.locate(p, nthreads=8)
    for x, y in zip(O.step(p, d, lo, pa, kind=0, nthreads=8), H.step(p, d, lo, pa, kind=9, nthreads=8)):
        assert same(x, y).all()
    for x, y in zip(O.trace(p, d, kind=0, nthreads=8), H.trace(p, d, kind=9, nthreads=8)):
        assert same(x, y).all()


@pytest.mark.parametrize("case", sorted(CASES))
def test_basket_navigator_host_build_matches_oracle(oracle, hostnav, case):
    """The basket SimpleNavigator (kind 3) of the product code, on baskets of
    tracks sharing a current volume, against the oracle."""
    from test_oracle_nav import baskets_by_volume
    flat, p, d = CASES[case]
    O, H = po.Nav(oracle, flat), po.Nav(hostnav, flat)
    level, path = O.locate(p, nthreads=8)
    for idx in baskets_by_volume(level, path):
        q, v, lo, pa = p[:, idx].copy(), d[:, idx].copy(), level[idx].copy(), path[:, idx].copy()
        for pstep in (np.inf, 7.5):
            for x, y in zip(O.step(q, v, lo, pa, pstep=pstep, kind=3), H.step(q, v, lo, pa, pstep=pstep, kind=3)):
                assert same(x, y).all(), (case, pstep)
