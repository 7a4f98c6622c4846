"""The Parallelepiped's extent and ABBox, pinned by closed forms.

The reference has no Parallelepiped Extent (its ABBoxManager asserts), so
the product bounds the shape by its eight vertices.  A parallelepiped with
half-lengths (dx, dy, dz) and angles (alpha, theta, phi) is the image of the
box |X| <= dx, |Y| <= dy, |Z| <= dz under
    x = X + Y tan(alpha) + Z tan(theta) cos(phi),  y = Y + Z tan(theta) sin(phi),  z = Z
(the inverse of the shear in volumes/kernel/ParallelepipedImplementation.h),
so its tight local extent is the min / max over the eight images of the
box corners, and its ABBox (source/ABBoxManager.cpp:128-161) is the
mother-frame box of the eight extent-box corners under the placement,
padded by 1e-3 mm.  Both are computed here in numpy from the parameters
alone and compared with what the geometry uploader builds (host build,
tests/native/navhost.cpp); a million points sampled inside the shape must
lie in the float box the navigator tests rays against, and the oracle's
Contains must hold at those points."""
import ctypes
import math
import os

import numpy as np
import pytest

import pyoracle as po
from conftest import ROOT
from vecgeom_amd import geometry as G
from vecgeom_amd import workloads as W

LIB = os.path.join(ROOT, "tests", "native", "libnavhost.so")

PARAS = [(20.0, 30.0, 40.0, 30.0, 15.0, 30.0), (5.0, 8.0, 3.0, -40.0, 60.0, 200.0), (10.0, 1.0, 25.0, 10.0, 80.0, -75.0),
         (7.0, 7.0, 7.0, 0.0, 0.0, 0.0)]


class _Host:
    def __init__(self):
        self.lib = ctypes.CDLL(LIB)
        self.prefix = "navhost_"
        self.lib.navhost_aabb.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 3


def _vertices(dx, dy, dz, a, t, ph):
    ta, tt = math.tan(math.radians(a)), math.tan(math.radians(t))
    c, s = math.cos(math.radians(ph)), math.sin(math.radians(ph))
    X, Y, Z = np.meshgrid([-dx, dx], [-dy, dy], [-dz, dz], indexing="ij")
    X, Y, Z = X.ravel(), Y.ravel(), Z.ravel()
    return np.stack([X + Y * ta + Z * tt * c, Y + Z * tt * s, Z])


def _to_master(xf, p):
    r = np.asarray(xf.r).reshape(3, 3)
    return np.asarray(xf.t)[:, None] + r @ p


@pytest.mark.parametrize("k", range(len(PARAS)))
def test_parallelepiped_extent_and_abbox_closed_form(oracle, k):
    host = _Host()
    prm = PARAS[k]
    rng = np.random.Generator(np.random.Philox(4242 + k))
    xf = W.random_placement(rng)
    u = G.UnplacedParallelepiped(*prm)
    flat, ids, _ = G.single_shape_flat([u], [xf])
    nav = po.Nav(host, flat)
    box = np.zeros(6)
    lo, hi = np.zeros(3), np.zeros(3)
    host.lib.navhost_aabb(nav.h, ids[0], *[a.ctypes.data_as(ctypes.c_void_p) for a in (box, lo, hi)])
    v = _vertices(*prm)
    # tight local extent: the bounds of the eight vertices
    np.testing.assert_allclose(lo, v.min(axis=1), rtol=1e-13, atol=1e-12)
    np.testing.assert_allclose(hi, v.max(axis=1), rtol=1e-13, atol=1e-12)
    # ABBox: the eight extent-box corners in the mother frame, padded 1e-3
    C = np.stack(np.meshgrid([lo[0], hi[0]], [lo[1], hi[1]], [lo[2], hi[2]], indexing="ij")).reshape(3, -1)
    m = _to_master(xf, C)
    np.testing.assert_allclose(box[:3], m.min(axis=1) - 1e-3, rtol=0, atol=1e-9)
    np.testing.assert_allclose(box[3:], m.max(axis=1) + 1e-3, rtol=0, atol=1e-9)
    # it contains the shape's vertices with at least the padding to spare
    vm = _to_master(xf, v)
    assert (vm.min(axis=1) >= box[:3] + 1e-3 - 1e-9).all() and (vm.max(axis=1) <= box[3:] - 1e-3 + 1e-9).all()
    # and a million interior points, inside the float box the navigator uses,
    # each contained by the oracle's shape
    n = 1_000_000
    dx, dy, dz = prm[:3]
    q = np.stack([rng.uniform(-dx, dx, n), rng.uniform(-dy, dy, n), rng.uniform(-dz, dz, n)])
    ta, tt = math.tan(math.radians(prm[3])), math.tan(math.radians(prm[4]))
    c, s = math.cos(math.radians(prm[5])), math.sin(math.radians(prm[5]))
    local = np.stack([q[0] + q[1] * ta + q[2] * tt * c, q[1] + q[2] * tt * s, q[2]])
    pm = _to_master(xf, local)
    fbox = box.astype(np.float32).astype(np.float64)
    assert ((pm >= fbox[:3, None]) & (pm <= fbox[3:, None])).all()
    o = oracle.placed(u.type, u.params, xf.array())
    inner = 0.999 * local  # away from the surface band
    assert o.query(po.Q_CONTAINS, _to_master(xf, inner)[:, :100000]).all()
