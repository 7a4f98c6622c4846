"""The device shape code (vecgeom_amd/csrc/shapes_*.hpp), compiled for the
host through tests/native/navhost.cpp, against the oracle: bit-exact on the
golden inputs and on seeded random baskets, for every shape with device
kernels, under identity and random placements.  Checks the shape logic on
the CPU; the GPU tests check the same code on the device."""
import ctypes
import os

import numpy as np
import pytest

import pyoracle as po
from conftest import ROOT, assert_agree, golden_cases, load_golden, same
from vecgeom_amd import geometry as G
from vecgeom_amd import workloads as W

LIB = os.path.join(ROOT, "tests", "native", "libnavhost.so")
DEVICE_TYPES = {1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12}
QUERIES = [("contains", po.Q_CONTAINS, 0), ("inside", po.Q_INSIDE, 1), ("dist_in", po.Q_DIST_IN, 2),
           ("dist_out", po.Q_DIST_OUT, 3), ("safe_in", po.Q_SAFE_IN, 4), ("safe_out", po.Q_SAFE_OUT, 5)]
EXTRA = [  # phi-sectioned and degenerate variants beyond the zoo
    ("ConePhi", G.UnplacedCone(5, 20, 10, 40, 30, 0.3, 4.0)),
    ("ConeNoRmin", G.UnplacedCone(0, 20, 0, 40, 30, 1.0, 2.0)),
    ("ConeHalf", G.UnplacedCone(5, 20, 10, 40, 30, 0.5, np.pi)),
    ("ConeApex", G.UnplacedCone(0, 30, 0, 1e-4, 40, 0, 2 * np.pi)),
    ("TubePhi", G.UnplacedTube(10, 50, 50, 0.4, 2.0)),
    ("PolyconePhi", G.UnplacedPolycone(0.5, 4.0, rmin=[5, 5, 10, 10, 8], rmax=[20, 30, 30, 40, 40],
                                       z=[-50, -40, -30, -20, 0])),
    ("PolyconeDupZ", G.UnplacedPolycone(0.0, 2 * np.pi, rmin=[5, 5, 8, 8], rmax=[20, 25, 25, 30], z=[-30, 0, 0, 30])),
    ("TorusPhi", G.UnplacedTorus(5, 10, 50, 0.3, 4.0)),
    ("TorusSolid", G.UnplacedTorus(0, 10, 40, 0, 2 * np.pi)),
    ("PolyconeOne", G.UnplacedPolycone(0.0, 2 * np.pi, rmin=[0, 4], rmax=[20, 30], z=[-30, 30])),
    ("PolyhedronPhi", G.UnplacedPolyhedron(20.0, 250.0, 5, z=[-40, -10, 10, 40], rmin=[5, 8, 8, 5],
                                           rmax=[25, 35, 30, 20])),
    ("PolyhedronLargeCut", G.UnplacedPolyhedron(-30.0, 120.0, 3, z=[-30, 0, 30], rmin=[0, 5, 0], rmax=[20, 25, 20])),
    ("PolyhedronSolid", G.UnplacedPolyhedron(0.0, 360.0, 6, z=[-20, 0, 20], rmin=[0, 0, 0], rmax=[15, 25, 15])),
    ("PolyhedronDupZ", G.UnplacedPolyhedron(0.0, 360.0, 7, z=[-30, -10, -10, 15, 30], rmin=[5, 5, 12, 8, 6],
                                            rmax=[20, 25, 35, 30, 22])),
    ("PolyhedronDupZPhi", G.UnplacedPolyhedron(45.0, 300.0, 4, z=[-30, -10, -10, 15, 30], rmin=[0, 2, 12, 8, 6],
                                               rmax=[20, 25, 35, 30, 22])),
    # 12 sides x 12 segments: a parameter block larger than one workgroup's LDS stage
    ("PolyhedronMany", G.UnplacedPolyhedron(10.0, 350.0, 12, z=list(np.linspace(-60, 60, 13)),
                                            rmin=[5, 6, 8, 4, 9, 12, 3, 7, 8, 10, 6, 5, 4],
                                            rmax=[20, 25, 30, 22, 35, 40, 18, 27, 33, 38, 26, 21, 19])),
]


# Distances of the Torus: the device solves the reference's quartic for its
# exact real roots, the reference (and the oracle) in closed form; they are
# held to the FP64 distance bar max(1e-9 mm, 1e-12 relative), and lanes
# beyond it to the exact first crossing (conftest.assert_agree).
TOLERANT = {("Torus", "dist_in"), ("Torus", "dist_out")}


class HostLib:
    def __init__(self):
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is missing: build it with __graft_entry__.build()")
        self.lib = ctypes.CDLL(LIB)
        self.prefix = "navhost_"
        self.lib.navhost_query.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_long] + \
            [ctypes.c_void_p] * 8
        self.lib.navhost_query.restype = ctypes.c_int


@pytest.fixture(scope="module")
def host():
    return HostLib()


def run_host(host, nav, pid, q, p, d=None, sm=None):
    ptr = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)
    p = np.ascontiguousarray(p)
    n = p.shape[1]
    out = np.empty(n, dtype=po.QUERY_DTYPE[q])
    dd = [None] * 3 if d is None else [ptr(np.ascontiguousarray(d[i])) for i in range(3)]
    rc = host.lib.navhost_query(nav.h, pid, q, n, ptr(np.ascontiguousarray(p[0])), ptr(np.ascontiguousarray(p[1])),
                                ptr(np.ascontiguousarray(p[2])), *dd, ptr(sm), ptr(out))
    assert rc == 0
    return out


def compare(oracle, host, u, xf, inputs, family=""):
    flat, ids, _ = G.single_shape_flat([u], [xf])
    nav = po.Nav(host, flat)
    o = oracle.placed(u.type, u.params, xf.array())
    for name, q, hq in QUERIES:
        if name in ("dist_out", "safe_out"):
            args = (inputs["lp"], inputs["ld"] if name == "dist_out" else None,
                    inputs["sm"] if name == "dist_out" else None)
        else:
            args = (inputs["p"], inputs["d"] if name == "dist_in" else None, inputs["sm"] if name == "dist_in" else None)
        want = o.query(q, *[a for a in args if a is not None]) if args[1] is not None else o.query(q, args[0])
        got = run_host(host, nav, ids[0], hq, *args)
        if (family, name) in TOLERANT:
            assert_agree("Torus", name, got, want, "host", torus=(u.params, inputs["lp"], inputs["ld"]))
            continue
        ok = same(got, want)
        assert ok.all(), f"{name}: {int((~ok).sum())} lanes differ, first {np.flatnonzero(~ok)[:4]}"


@pytest.mark.parametrize("case", golden_cases())
def test_host_shapes_on_golden(oracle, host, case):
    g = load_golden(case)
    if int(g["type"]) not in DEVICE_TYPES:
        pytest.skip("no device kernels for this shape")
    u = G.UnplacedVolume(int(g["type"]), g["params"])
    compare(oracle, host, u, G.Transformation3D.from_array(g["xf"]), g, case.split("_")[1].replace("C2", ""))


@pytest.mark.parametrize("name", [n for n, u in W.zoo_shapes() if u.type in DEVICE_TYPES] + [n for n, _ in EXTRA])
def test_host_shapes_random(oracle, host, name):
    u = dict(W.zoo_shapes() + EXTRA)[name]
    rng = np.random.Generator(np.random.Philox(2024))
    for xf in (G.Transformation3D(), W.random_placement(rng)):
        lo, hi = oracle.placed(u.type, u.params, np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1.0])).extent()
        inp = W.shape_inputs((lo, hi), xf, 30000, seed=77)
        inp["sm"] = np.full(30000, 1e30)
        compare(oracle, host, u, xf, inp, "Torus" if u.type == 10 else name)
