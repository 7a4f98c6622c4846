"""ORACLE -- test infrastructure only.

ctypes access to the CPU restatement (``oracle/lib/liboracle*.so``) and to the
shim over the real reference build (``oracle/_ref/libvgref*.so``).  Both
export the same entry points (prefix ``vgo_`` / ``vgref_``).  Only tests/,
``__graft_entry__.smoke()`` and the cpu_baseline leg of bench.py import this
module; the product package never does.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

Q_CONTAINS, Q_INSIDE, Q_DIST_IN, Q_DIST_OUT, Q_SAFE_IN, Q_SAFE_OUT, Q_PLACED_DIST_OUT = range(7)
QUERY_DTYPE = {
    Q_CONTAINS: np.uint8,
    Q_INSIDE: np.int32,
    Q_DIST_IN: np.float64,
    Q_DIST_OUT: np.float64,
    Q_SAFE_IN: np.float64,
    Q_SAFE_OUT: np.float64,
    Q_PLACED_DIST_OUT: np.float64,
}

_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.c_void_p)


class Lib:
    """One loaded oracle flavour (restatement or reference shim)."""

    def __init__(self, path, prefix):
        self.path = path
        self.prefix = prefix
        self.lib = ctypes.CDLL(path)
        f = self._f
        f("placed_create").restype = ctypes.c_void_p
        f("placed_create").argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        f("placed_destroy").argtypes = [ctypes.c_void_p]
        f("query").restype = ctypes.c_int
        f("query").argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_long] + [ctypes.c_void_p] * 8 + [ctypes.c_int]
        f("shape_extent").argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        if self.has("basket"):
            f("basket").restype = ctypes.c_int
            f("basket").argtypes = f("query").argtypes
        if self.has("normal"):
            f("normal").restype = ctypes.c_int
            f("normal").argtypes = [ctypes.c_void_p, ctypes.c_long] + [ctypes.c_void_p] * 7
            f("distance_to_out_normal").restype = ctypes.c_int
            f("distance_to_out_normal").argtypes = [ctypes.c_void_p, ctypes.c_long] + [ctypes.c_void_p] * 13

    def _f(self, name):
        return getattr(self.lib, self.prefix + name)

    def has(self, name):
        return hasattr(self.lib, self.prefix + name)

    def placed(self, shape_type, params, xf12):
        p = np.ascontiguousarray(params, dtype=np.float64)
        x = np.ascontiguousarray(xf12, dtype=np.float64)
        h = self._f("placed_create")(int(shape_type), _ptr(p), len(p), _ptr(x))
        if not h:
            raise ValueError(f"{self.prefix}placed_create refused shape {shape_type} params {list(p)}")
        return Placed(self, h)


class Placed:
    def __init__(self, lib, handle):
        self.lib = lib
        self.h = handle

    def extent(self):
        lo = np.zeros(3)
        hi = np.zeros(3)
        self.lib._f("shape_extent")(self.h, _ptr(lo), _ptr(hi))
        return lo, hi

    def query(self, q, p, d=None, step_max=None, nthreads=1):
        """p, d: (3, n) float64 arrays (SOA).  Returns the output array."""
        p = np.ascontiguousarray(p, dtype=np.float64)
        n = p.shape[1]
        out = np.empty(n, dtype=QUERY_DTYPE[q])
        if d is not None:
            d = np.ascontiguousarray(d, dtype=np.float64)
            dx, dy, dz = _ptr(d[0]), _ptr(d[1]), _ptr(d[2])
        else:
            dx = dy = dz = None
        sm = None if step_max is None else np.ascontiguousarray(step_max, dtype=np.float64)
        rc = self.lib._f("query")(self.h, q, n, _ptr(p[0]), _ptr(p[1]), _ptr(p[2]), dx, dy, dz, _ptr(sm), _ptr(out),
                                  int(nthreads))
        if rc != 0:
            raise RuntimeError(f"{self.lib.prefix}query failed rc={rc}")
        return out

    def basket(self, q, p, d=None, step_max=None, nthreads=1):
        """The reference's SOA3D basket virtuals (reference shim only);
        queries 0..5 as in query()."""
        p = np.ascontiguousarray(p, dtype=np.float64)
        n = p.shape[1]
        out = np.empty(n, dtype=QUERY_DTYPE[q])
        if d is not None:
            d = np.ascontiguousarray(d, dtype=np.float64)
            dx, dy, dz = _ptr(d[0]), _ptr(d[1]), _ptr(d[2])
        else:
            dx = dy = dz = None
        sm = None if step_max is None else np.ascontiguousarray(step_max, dtype=np.float64)
        rc = self.lib._f("basket")(self.h, q, n, _ptr(p[0]), _ptr(p[1]), _ptr(p[2]), dx, dy, dz, _ptr(sm), _ptr(out),
                                   int(nthreads))
        if rc != 0:
            raise RuntimeError(f"{self.lib.prefix}basket failed rc={rc}")
        return out

    def normal(self, p):
        """VPlacedVolume::Normal on local points p (3, n): returns (normals
        (3, n), valid uint8 (n,)).  ValueError where the reference has no
        Normal for the shape."""
        p = np.ascontiguousarray(p, dtype=np.float64)
        n = p.shape[1]
        nrm = np.empty((3, n))
        valid = np.empty(n, dtype=np.uint8)
        rc = self.lib._f("normal")(self.h, n, _ptr(p[0]), _ptr(p[1]), _ptr(p[2]), _ptr(nrm[0]), _ptr(nrm[1]),
                                   _ptr(nrm[2]), _ptr(valid))
        if rc == -3:
            raise ValueError(f"{self.lib.prefix}normal: the reference has no Normal for this shape")
        if rc != 0:
            raise RuntimeError(f"{self.lib.prefix}normal failed rc={rc}")
        return nrm, valid

    def distance_to_out_normal(self, p, d, step_max=None):
        """The USolids DistanceToOut(point, direction, normal&, convex&,
        stepMax) on local rays: (distance, normals (3, n), valid, convex)."""
        p = np.ascontiguousarray(p, dtype=np.float64)
        d = np.ascontiguousarray(d, dtype=np.float64)
        n = p.shape[1]
        dist = np.empty(n)
        nrm = np.empty((3, n))
        valid = np.empty(n, dtype=np.uint8)
        convex = np.empty(n, dtype=np.uint8)
        sm = None if step_max is None else np.ascontiguousarray(step_max, dtype=np.float64)
        rc = self.lib._f("distance_to_out_normal")(self.h, n, _ptr(p[0]), _ptr(p[1]), _ptr(p[2]), _ptr(d[0]),
                                                   _ptr(d[1]), _ptr(d[2]), _ptr(sm), _ptr(dist), _ptr(nrm[0]),
                                                   _ptr(nrm[1]), _ptr(nrm[2]), _ptr(valid), _ptr(convex))
        if rc == -3:
            raise ValueError(f"{self.lib.prefix}distance_to_out_normal: the reference has no Normal for this shape")
        if rc != 0:
            raise RuntimeError(f"{self.lib.prefix}distance_to_out_normal failed rc={rc}")
        return dist, nrm, valid, convex

    def __del__(self):
        try:
            self.lib._f("placed_destroy")(self.h)
        except Exception:
            pass


_cache = {}


def oracle(fast=False):
    """The restatement (strict by default)."""
    key = ("vgo", fast)
    if key not in _cache:
        name = "liboracle_fast.so" if fast else "liboracle.so"
        path = os.path.join(HERE, "lib", name)
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run __graft_entry__.build()")
        _cache[key] = Lib(path, "vgo_")
    return _cache[key]


def reference(fast=False):
    """The real reference build, or None where it was not built."""
    key = ("vgref", fast)
    if key not in _cache:
        name = "libvgref_fast.so" if fast else "libvgref.so"
        path = os.path.join(HERE, "_ref", name)
        _cache[key] = Lib(path, "vgref_") if os.path.exists(path) else None
    return _cache[key]


def _truth():
    key = ("truth",)
    if key not in _cache:
        path = os.path.join(HERE, "lib", "libvgotruth.so")
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run __graft_entry__.build()")
        lib = ctypes.CDLL(path)
        lib.vgo_torus_window.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_long] + [ctypes.c_void_p] * 9
        lib.vgo_torus_on_surface.argtypes = [ctypes.c_void_p, ctypes.c_long] + [ctypes.c_void_p] * 9
        _cache[key] = lib
    return _cache[key]


def _rays(params, p, v):
    par = np.ascontiguousarray(np.asarray(params, np.float64)[:5])
    p = [np.ascontiguousarray(np.asarray(p[i], np.float64)) for i in range(3)]
    v = [np.ascontiguousarray(np.asarray(v[i], np.float64)) for i in range(3)]
    return par, p, v


def torus_window(params, p, v, out, tol=1e-9):
    """Binary128 window of acceptable first-crossing distances of rays
    (p, v) (local frame, 3 x n) for the torus (rmin, rmax, rtor, sphi, dphi),
    DistanceToIn (out False) or DistanceToOut (out True): (lo, hi) per ray,
    NaN where the start is not one the query is defined for or lies within
    tol (a scalar or one per ray) of the surface (vgo_truth.cpp)."""
    par, p, v = _rays(params, p, v)
    n = len(p[0])
    tol = np.ascontiguousarray(np.broadcast_to(np.asarray(tol, np.float64), (n,)))
    lo, hi = np.empty(n), np.empty(n)
    _truth().vgo_torus_window(_ptr(par), int(bool(out)), n, *[_ptr(a) for a in p], *[_ptr(a) for a in v],
                              _ptr(tol), _ptr(lo), _ptr(hi))
    return lo, hi


def torus_on_surface(params, p, v, t, tol=1e-9):
    """Per ray: the point at the finite distance t lies within tol of a
    boundary surface of the torus (binary128, vgo_truth.cpp)."""
    par, p, v = _rays(params, p, v)
    t = np.ascontiguousarray(np.asarray(t, np.float64))
    tol = np.ascontiguousarray(np.broadcast_to(np.asarray(tol, np.float64), t.shape))
    on = np.empty(len(t), np.uint8)
    _truth().vgo_torus_on_surface(_ptr(par), len(t), *[_ptr(a) for a in p], *[_ptr(a) for a in v], _ptr(t),
                                  _ptr(tol), _ptr(on))
    return on.astype(bool)


def torus_distance_ok(params, p, v, got, out, tol=1e-9):
    """Per lane: the Torus distance `got` of the local ray (p, v) is right to
    within the surface tolerance.  Where the query is defined (a start
    outside for DistanceToIn, out False; inside for DistanceToOut) it lies
    in the window of the exact first boundary crossing, widened by the FP64
    bar max(1e-9, 1e-12 |got|), and is +inf only where the ray never gets tol
    deep into (or out of) the solid.  Elsewhere -- starts on the surface,
    and starts the query is not defined for, whose answer is an artefact of
    the algorithm -- it is a finite distance to a point within tol + bar of
    a boundary surface, or +inf for a start off the surface whose ray never
    comes within tol of the boundary.

    The reference's quartic takes the direction as a unit vector; a
    direction of norm 1 + e (float32 inputs promoted to double: e ~ 1e-8)
    moves its roots by up to ~e times the torus' size.  The crossings are
    judged along v/|v|, and a lane passes with the tolerance tol or with
    tol widened by 4 e (rtor + rmax + |p|)."""
    got = np.asarray(got, dtype=np.float64)
    p = np.asarray(p, np.float64)
    v = np.asarray(v, np.float64)
    nv = np.sqrt((v * v).sum(axis=0))
    vh = v / nv
    e = np.abs(nv * nv - 1.0)
    wide = tol + 4.0 * e * (float(params[2]) + float(params[1]) + np.sqrt((p * p).sum(axis=0)))
    fin_got = np.isfinite(got)
    bar = np.maximum(1e-9, 1e-12 * np.where(fin_got, np.abs(got), 0.0))

    def check(t):
        lo, hi = torus_window(params, p, vh, out, t)
        with np.errstate(invalid="ignore"):
            inside = (got >= lo - bar) & (got <= hi + bar)
        defined = np.where(np.isinf(got), np.isinf(hi), inside)
        # Starts the query is not defined for: +inf only if the ray never
        # comes within t of the boundary (the other query's window opens
        # nowhere), a finite distance only onto a boundary surface.
        lo2, _ = torus_window(params, p, vh, not out, t)
        undefined = np.where(np.isinf(got), np.isinf(lo2), torus_on_surface(params, p, vh, got, t + bar))
        return np.where(np.isnan(lo), undefined, defined)

    ok = check(np.full(got.shape, float(tol)))
    if not ok.all() and (wide > tol).any():
        ok |= check(wide)
    return ok


# ---------------------------------------------------------------- navigation
_ip = ctypes.POINTER(ctypes.c_int)


class Nav:
    """Flat geometry + navigator of one oracle flavour."""

    def __init__(self, lib, flat):
        self.lib = lib
        f = lib.lib
        pre = lib.prefix
        self.pre = pre
        getattr(f, pre + "geom_create").restype = ctypes.c_void_p
        getattr(f, pre + "geom_create").argtypes = [ctypes.c_int] + [ctypes.c_void_p] * 3 + [ctypes.c_int] + \
            [ctypes.c_void_p] * 3 + [ctypes.c_int] + [ctypes.c_void_p] * 2 + [ctypes.c_int]
        getattr(f, pre + "geom_max_depth").argtypes = [ctypes.c_void_p]
        getattr(f, pre + "locate").argtypes = [ctypes.c_void_p, ctypes.c_long] + [ctypes.c_void_p] * 3 + \
            [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        getattr(f, pre + "nav_step").argtypes = [ctypes.c_void_p, ctypes.c_long] + [ctypes.c_void_p] * 6 + \
            [ctypes.c_int] + [ctypes.c_void_p] * 6 + [ctypes.c_double, ctypes.c_int, ctypes.c_int]
        getattr(f, pre + "safety").argtypes = [ctypes.c_void_p, ctypes.c_long] + [ctypes.c_void_p] * 3 + \
            [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        getattr(f, pre + "trace").argtypes = [ctypes.c_void_p, ctypes.c_long] + [ctypes.c_void_p] * 6 + \
            [ctypes.c_double, ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 6 + [ctypes.c_int]
        self.flat = flat
        self.h = getattr(f, pre + "geom_create")(
            len(flat.shape_type), _ptr(flat.shape_type), _ptr(flat.shape_param_begin), _ptr(flat.params),
            len(flat.logical_shape), _ptr(flat.logical_shape), _ptr(flat.logical_daughter_begin), _ptr(flat.daughters),
            len(flat.placed_logical), _ptr(flat.placed_logical), _ptr(flat.placed_transform), flat.world)
        if not self.h:
            raise ValueError(pre + "geom_create failed")
        self.depth = getattr(f, pre + "geom_max_depth")(self.h)

    def _f(self, n):
        return getattr(self.lib.lib, self.pre + n)

    def locate(self, p, nthreads=1):
        p = np.ascontiguousarray(p, np.float64)
        n = p.shape[1]
        level = np.zeros(n, np.int32)
        path = np.full((self.depth, n), -1, np.int32)
        self._f("locate")(self.h, n, _ptr(p[0]), _ptr(p[1]), _ptr(p[2]), self.depth, _ptr(level), _ptr(path), nthreads)
        return level, path

    def step(self, p, d, level, path, pstep=np.inf, kind=0, nthreads=1):
        p = np.ascontiguousarray(p, np.float64)
        d = np.ascontiguousarray(d, np.float64)
        n = p.shape[1]
        level = np.ascontiguousarray(level, np.int32)
        path = np.ascontiguousarray(path, np.int32)
        lo = np.zeros(n, np.int32)
        po = np.full((self.depth, n), -1, np.int32)
        bo = np.zeros(n, np.uint8)
        so = np.zeros(n, np.float64)
        self._f("nav_step")(self.h, n, _ptr(p[0]), _ptr(p[1]), _ptr(p[2]), _ptr(d[0]), _ptr(d[1]), _ptr(d[2]),
                            self.depth, _ptr(level), _ptr(path), _ptr(lo), _ptr(po), _ptr(bo), _ptr(so), float(pstep),
                            int(kind), nthreads)
        return lo, po, bo, so

    def safety(self, p, level, path, nthreads=1):
        p = np.ascontiguousarray(p, np.float64)
        n = p.shape[1]
        out = np.zeros(n)
        self._f("safety")(self.h, n, _ptr(p[0]), _ptr(p[1]), _ptr(p[2]), self.depth,
                          _ptr(np.ascontiguousarray(level, np.int32)), _ptr(np.ascontiguousarray(path, np.int32)),
                          _ptr(out), nthreads)
        return out

    def trace(self, p, d, pstep=np.inf, kind=0, max_steps=100000, nthreads=1):
        p = np.ascontiguousarray(p, np.float64)
        d = np.ascontiguousarray(d, np.float64)
        n = p.shape[1]
        ns = np.zeros(n, np.int32)
        fp = np.zeros((3, n))
        tr = np.zeros(n)
        hs = np.zeros(n, np.uint64)
        self._f("trace")(self.h, n, _ptr(p[0]), _ptr(p[1]), _ptr(p[2]), _ptr(d[0]), _ptr(d[1]), _ptr(d[2]),
                         float(pstep), int(kind), int(max_steps), _ptr(ns), _ptr(fp[0]), _ptr(fp[1]), _ptr(fp[2]),
                         _ptr(tr), _ptr(hs), nthreads)
        return ns, fp, tr, hs
