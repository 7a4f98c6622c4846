"""Python mirror of the reference's geometry-construction API, used by tests
and bench.py to build volume trees and upload them through the C-ABI.

Names and argument meaning follow the reference:
  UnplacedBox(dx, dy, dz) ...            volumes/Unplaced*.h
  LogicalVolume(label, unplaced)         volumes/LogicalVolume.h:65
  LogicalVolume.PlaceDaughter(label, logical, transformation)
                                         source/LogicalVolume.cpp:84-102
  Transformation3D(tx, ty, tz[, phi, theta, psi] | [, r0..r8])
                                         base/Transformation3D.h:63-101,
                                         source/Transformation3D.cpp:94-114
  GeoManager.SetWorld / CloseGeometry    management/GeoManager.h:129-200
The device side (DeviceGeometry) replaces CudaManager::LoadGeometry +
Synchronize (management/CudaManager.h:97-160).
"""
import ctypes
import math

import numpy as np

from . import capi

VG_BOX, VG_TUBE, VG_ORB, VG_SPHERE, VG_CONE, VG_TRD, VG_TRAPEZOID = 1, 2, 3, 4, 5, 6, 7
VG_PARALLELEPIPED, VG_PARABOLOID, VG_TORUS, VG_POLYCONE, VG_POLYHEDRON = 8, 9, 10, 11, 12
VG_BOOLEAN = 13
kUnion, kIntersection, kSubtraction = 0, 1, 2  # BooleanOperation, volumes/UnplacedBooleanVolume.h:10-14

SHAPE_NAMES = {
    VG_BOX: "Box", VG_TUBE: "Tube", VG_ORB: "Orb", VG_SPHERE: "Sphere", VG_CONE: "Cone", VG_TRD: "Trd",
    VG_TRAPEZOID: "Trapezoid", VG_PARALLELEPIPED: "Parallelepiped", VG_PARABOLOID: "Paraboloid",
    VG_TORUS: "Torus", VG_POLYCONE: "Polycone", VG_POLYHEDRON: "Polyhedron", VG_BOOLEAN: "Boolean",
}

kDegToRad = math.pi / 180.0


class UnplacedVolume:
    """Shape type + constructor parameters (the device library derives the
    reference's cached values from these)."""

    def __init__(self, shape_type, params):
        self.type = int(shape_type)
        self.params = [float(p) for p in params]

    def __repr__(self):
        return f"Unplaced{SHAPE_NAMES.get(self.type, self.type)}({', '.join(f'{p:g}' for p in self.params)})"


def UnplacedBox(dx, dy, dz):
    return UnplacedVolume(VG_BOX, [dx, dy, dz])


def UnplacedTube(rmin, rmax, dz, sphi, dphi):
    return UnplacedVolume(VG_TUBE, [rmin, rmax, dz, sphi, dphi])


def UnplacedOrb(r):
    return UnplacedVolume(VG_ORB, [r])


def UnplacedSphere(rmin, rmax, sphi, dphi, stheta, dtheta):
    return UnplacedVolume(VG_SPHERE, [rmin, rmax, sphi, dphi, stheta, dtheta])


def UnplacedCone(rmin1, rmax1, rmin2, rmax2, dz, sphi, dphi):
    return UnplacedVolume(VG_CONE, [rmin1, rmax1, rmin2, rmax2, dz, sphi, dphi])


def UnplacedTrd(dx1, dx2, dy1, dy2, dz):
    return UnplacedVolume(VG_TRD, [dx1, dx2, dy1, dy2, dz])


def UnplacedTrapezoid(dz, theta, phi, dy1, dx1, dx2, tanalpha1, dy2, dx3, dx4, tanalpha2):
    return UnplacedVolume(VG_TRAPEZOID, [dz, theta, phi, dy1, dx1, dx2, tanalpha1, dy2, dx3, dx4, tanalpha2])


def UnplacedParallelepiped(x, y, z, alpha, theta, phi):
    return UnplacedVolume(VG_PARALLELEPIPED, [x, y, z, alpha, theta, phi])


def UnplacedParaboloid(rlo, rhi, dz):
    return UnplacedVolume(VG_PARABOLOID, [rlo, rhi, dz])


def UnplacedBooleanVolume(op, left, left_xf, right, right_xf):
    """UnplacedBooleanVolume(op, leftPlaced, rightPlaced)
    (volumes/UnplacedBooleanVolume.h:39-50): the operands are primitive
    shapes placed in the Boolean's own frame, given here as (unplaced,
    Transformation3D) pairs.  Parameter block: op, then per operand type,
    number of parameters, parameters, placement (12 doubles)."""
    if op not in (kUnion, kIntersection, kSubtraction):
        raise ValueError("op must be kUnion, kIntersection or kSubtraction")
    params = [op]
    for u, xf in ((left, left_xf), (right, right_xf)):
        if u.type == VG_BOOLEAN:
            raise ValueError("Boolean operands must be primitive shapes")
        params += [u.type, len(u.params)] + list(u.params) + list(xf.array())
    return UnplacedVolume(VG_BOOLEAN, params)


def UnplacedTorus(rmin, rmax, rtor, sphi, dphi):
    return UnplacedVolume(VG_TORUS, [rmin, rmax, rtor, sphi, dphi])


def UnplacedPolycone(sphi, dphi, rmin, rmax, z):
    nz = len(z)
    assert len(rmin) == nz and len(rmax) == nz
    return UnplacedVolume(VG_POLYCONE, [sphi, dphi, nz] + list(rmin) + list(rmax) + list(z))


def UnplacedPolyhedron(phistart_deg, phidelta_deg, sides, z, rmin, rmax):
    nz = len(z)
    assert len(rmin) == nz and len(rmax) == nz
    return UnplacedVolume(VG_POLYHEDRON, [phistart_deg, phidelta_deg, sides, nz] + list(z) + list(rmin) + list(rmax))


class Transformation3D:
    """12 numbers tx ty tz r0..r8; local = R^T (master - t) as in
    base/Transformation3D.h:469-525."""

    def __init__(self, tx=0.0, ty=0.0, tz=0.0, *rot):
        self.t = [float(tx), float(ty), float(tz)]
        if len(rot) == 0:
            self.r = [1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0]
        elif len(rot) == 3:
            self.r = euler_rotation(*rot)
        elif len(rot) == 9:
            self.r = [float(v) for v in rot]
        else:
            raise ValueError("Transformation3D takes 0, 3 (phi theta psi, degrees) or 9 rotation values")

    def array(self):
        return np.array(self.t + self.r, dtype=np.float64)

    @staticmethod
    def from_array(a):
        a = [float(v) for v in a]
        return Transformation3D(a[0], a[1], a[2], *a[3:12])


def euler_rotation(phi, theta, psi):
    """Transformation3D::SetRotation(phi, theta, psi) in degrees,
    source/Transformation3D.cpp:94-114."""
    sinphi, cosphi = math.sin(kDegToRad * phi), math.cos(kDegToRad * phi)
    sinthe, costhe = math.sin(kDegToRad * theta), math.cos(kDegToRad * theta)
    sinpsi, cospsi = math.sin(kDegToRad * psi), math.cos(kDegToRad * psi)
    return [
        cospsi * cosphi - costhe * sinphi * sinpsi,
        -sinpsi * cosphi - costhe * sinphi * cospsi,
        sinthe * sinphi,
        cospsi * sinphi + costhe * cosphi * sinpsi,
        -sinpsi * sinphi + costhe * cosphi * cospsi,
        -sinthe * cosphi,
        sinpsi * sinthe,
        cospsi * sinthe,
        costhe,
    ]


class PlacedVolume:
    def __init__(self, manager, pid, label, logical, transformation):
        self.manager = manager
        self.id = pid
        self.label = label
        self.logical = logical
        self.transformation = transformation

    def __repr__(self):
        return f"PlacedVolume({self.id}, {self.label!r}, {self.logical.unplaced!r})"


class LogicalVolume:
    def __init__(self, label, unplaced, manager=None):
        self.label = label
        self.unplaced = unplaced
        self.daughters = []
        self.manager = manager if manager is not None else GeoManager.Instance()
        self.id = self.manager._register_logical(self)

    def PlaceDaughter(self, label, logical, transformation=None):
        pv = self.manager._place(label, logical, transformation or Transformation3D())
        self.daughters.append(pv)
        return pv

    def Place(self, label="", transformation=None):
        return self.manager._place(label, self, transformation or Transformation3D())


class GeoManager:
    """Registry of volumes; CloseGeometry() flattens the tree for upload."""

    _instance = None

    @classmethod
    def Instance(cls):
        if cls._instance is None:
            cls._instance = GeoManager()
        return cls._instance

    def __init__(self):
        self.logical = []
        self.placed = []
        self.world = None

    def _register_logical(self, lv):
        self.logical.append(lv)
        return len(self.logical) - 1

    def _place(self, label, logical, transformation):
        pv = PlacedVolume(self, len(self.placed), label, logical, transformation)
        self.placed.append(pv)
        return pv

    def SetWorld(self, world_placed):
        self.world = world_placed

    def GetWorld(self):
        return self.world

    def CloseGeometry(self):
        """Returns the flat arrays of vg_flat_geometry (numpy)."""
        if self.world is None:
            raise RuntimeError("GeoManager: SetWorld was not called")
        shapes, sbegin, params = [], [0], []
        shape_of = {}
        lshape, lbegin, daughters = [], [0], []
        for lv in self.logical:
            key = id(lv.unplaced)
            if key not in shape_of:
                shape_of[key] = len(shapes)
                shapes.append(lv.unplaced.type)
                params.extend(lv.unplaced.params)
                sbegin.append(len(params))
            lshape.append(shape_of[key])
            daughters.extend(pv.id for pv in lv.daughters)
            lbegin.append(len(daughters))
        plog = [pv.logical.id for pv in self.placed]
        pxf = np.concatenate([pv.transformation.array() for pv in self.placed]) if self.placed else np.zeros(0)
        return FlatArrays(
            np.array(shapes, np.int32), np.array(sbegin, np.int32), np.array(params, np.float64),
            np.array(lshape, np.int32), np.array(lbegin, np.int32), np.array(daughters, np.int32),
            np.array(plog, np.int32), pxf.astype(np.float64), int(self.world.id))


class FlatArrays:
    def __init__(self, shape_type, shape_param_begin, params, logical_shape, logical_daughter_begin, daughters,
                 placed_logical, placed_transform, world):
        self.shape_type = np.ascontiguousarray(shape_type)
        self.shape_param_begin = np.ascontiguousarray(shape_param_begin)
        self.params = np.ascontiguousarray(params) if len(params) else np.zeros(1)
        self.logical_shape = np.ascontiguousarray(logical_shape)
        self.logical_daughter_begin = np.ascontiguousarray(logical_daughter_begin)
        self.daughters = np.ascontiguousarray(daughters) if len(daughters) else np.zeros(1, np.int32)
        self.placed_logical = np.ascontiguousarray(placed_logical)
        self.placed_transform = np.ascontiguousarray(placed_transform)
        self.world = world

    def struct(self):
        p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        return capi.FlatGeometry(
            len(self.shape_type), p(self.shape_type), p(self.shape_param_begin), p(self.params),
            len(self.logical_shape), p(self.logical_shape), p(self.logical_daughter_begin), p(self.daughters),
            len(self.placed_logical), p(self.placed_logical), p(self.placed_transform), self.world)


class DeviceGeometry:
    """An uploaded geometry (vg_geometry*).  Replaces CudaManager."""

    def __init__(self, flat):
        self.flat = flat
        self._struct = flat.struct()
        h = ctypes.c_void_p()
        capi.call("vg_geometry_upload", ctypes.byref(self._struct), ctypes.byref(h))
        self.handle = h

    def max_depth(self):
        d = ctypes.c_int32()
        capi.call("vg_geometry_max_depth", self.handle, ctypes.byref(d))
        return d.value

    def placed_extent(self, pid):
        lo = np.zeros(3)
        hi = np.zeros(3)
        capi.call("vg_geometry_placed_extent", self.handle, int(pid), lo.ctypes.data_as(ctypes.c_void_p),
                  hi.ctypes.data_as(ctypes.c_void_p))
        return lo, hi

    def placed_aabb(self, pid):
        lo = np.zeros(3)
        hi = np.zeros(3)
        capi.call("vg_geometry_placed_aabb", self.handle, int(pid), lo.ctypes.data_as(ctypes.c_void_p),
                  hi.ctypes.data_as(ctypes.c_void_p))
        return lo, hi

    def free(self):
        if self.handle:
            capi.call("vg_geometry_free", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def single_shape_flat(unplaced_list, transforms, world_half=1e4):
    """World box holding one placement per (shape, transform); returns
    (flat geometry, [placed ids], GeoManager) on the host."""
    gm = GeoManager()
    world_lv = LogicalVolume("world", UnplacedBox(world_half, world_half, world_half), gm)
    ids = []
    for k, (u, xf) in enumerate(zip(unplaced_list, transforms)):
        lv = LogicalVolume(f"shape{k}", u, gm)
        pv = world_lv.PlaceDaughter(f"shape{k}", lv, xf)
        ids.append(pv.id)
    world = world_lv.Place("world")
    gm.SetWorld(world)
    return gm.CloseGeometry(), ids, gm


def single_shape_geometry(unplaced_list, transforms, world_half=1e4):
    """single_shape_flat uploaded: (DeviceGeometry, [placed ids], GeoManager).
    Used by shape benchmarks and tests."""
    flat, ids, gm = single_shape_flat(unplaced_list, transforms, world_half)
    return DeviceGeometry(flat), ids, gm
