"""The reference-side facade (oracle/facade/vecgeom_amd_facade.h), compiled
against the reference headers: a geometry is built with the reference's own
classes (LogicalVolume, PlaceDaughter, Transformation3D, GeoManager) from
the flat arrays geometry.py produces, flattened back by the facade from the
reference's world volume, and compared with the original arrays -- the
same tree, shapes, parameters, daughter order and transforms, visited from
the world in depth-first order (ids may be numbered differently)."""
import ctypes
import os

import numpy as np
import pytest

from conftest import ROOT
from vecgeom_amd import workloads as W

LIB = os.path.join(ROOT, "oracle", "_ref", "libfacade.so")


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def roundtrip(flat):
    lib = ctypes.CDLL(LIB)
    v = ctypes.c_void_p
    lib.facade_roundtrip.argtypes = [ctypes.c_int, v, v, v, ctypes.c_int, v, v, v, ctypes.c_int, v, v, ctypes.c_int, v]
    lib.facade_arrays.argtypes = [v] * 8
    a = lambda x, t: np.ascontiguousarray(x, dtype=t)
    st, pb, pr = a(flat.shape_type, np.int32), a(flat.shape_param_begin, np.int32), a(flat.params, np.float64)
    ls, ld, da = a(flat.logical_shape, np.int32), a(flat.logical_daughter_begin, np.int32), a(flat.daughters, np.int32)
    pl, px = a(flat.placed_logical, np.int32), a(flat.placed_transform, np.float64)
    out6 = np.zeros(6, np.int32)
    rc = lib.facade_roundtrip(len(st), _p(st), _p(pb), _p(pr), len(ls), _p(ls), _p(ld), _p(da), len(pl), _p(pl),
                              _p(px), int(flat.world), _p(out6))
    assert rc == 0
    ns, npar, nl, nd, npl, world = (int(x) for x in out6)
    r = dict(shape_type=np.zeros(ns, np.int32), shape_param_begin=np.zeros(ns + 1, np.int32),
             params=np.zeros(npar), logical_shape=np.zeros(nl, np.int32),
             logical_daughter_begin=np.zeros(nl + 1, np.int32), daughters=np.zeros(nd, np.int32),
             placed_logical=np.zeros(npl, np.int32), placed_transform=np.zeros(12 * npl))
    lib.facade_arrays(*[_p(r[k]) for k in ("shape_type", "shape_param_begin", "params", "logical_shape",
                                            "logical_daughter_begin", "daughters", "placed_logical",
                                            "placed_transform")])
    r["world"] = world
    return r


def canonical(f):
    """(shape type, parameters, transform, number of daughters) of every
    placed volume in depth-first order from the world."""
    get = (lambda k: np.asarray(f[k])) if isinstance(f, dict) else (lambda k: np.asarray(getattr(f, k)))
    st, pb, pr = get("shape_type"), get("shape_param_begin"), get("params")
    ls, ld, da, pl, px = (get(k) for k in ("logical_shape", "logical_daughter_begin", "daughters", "placed_logical",
                                           "placed_transform"))
    world = int(f["world"] if isinstance(f, dict) else f.world)
    out, stack = [], [world]
    while stack:
        p = stack.pop()
        lv = int(pl[p])
        s = int(ls[lv])
        kids = [int(x) for x in da[ld[lv]:ld[lv + 1]]]
        out.append((int(st[s]), tuple(pr[pb[s]:pb[s + 1]]), tuple(px[12 * p:12 * p + 12]), len(kids)))
        stack.extend(reversed(kids))
    return out


@pytest.mark.skipif(not os.path.exists(LIB), reason="reference build (oracle/_ref) not present")
def test_facade_flattens_a_reference_tree_like_geometry_py():
    _, flat = W.nested_geometry()
    got = roundtrip(flat)
    a, b = canonical(got), canonical(flat)
    assert len(a) == len(b) and a == b
    # The same placed and logical volumes (shared logical volumes stay shared).
    assert len(got["placed_logical"]) == len(flat.placed_logical)
    assert len(got["logical_shape"]) == len(flat.logical_shape)
    assert got["world"] == 0
