"""Synthetic workloads: shape sets, placements and seeded SOA3D inputs.

Shape-benchmark protocol (after source/benchmarking/Benchmarker.cpp:637-760,
with this project's own generator): points are uniform in 1.5x the local
bounding extent of each shape, expressed in the mother frame through the
placement; directions are isotropic (cos(theta) uniform, phi uniform).
Contains and Inside run on all points; points that the shape contains feed
DistanceToOut / SafetyToOut in the local frame, the others feed
DistanceToIn / SafetyToIn (step max 1e30).
"""
import math

import numpy as np

from . import geometry as G

TWO_PI = 2.0 * math.pi
SEED_BASE = 0xC0FFEE00


def config1_shapes():
    """Box (10,15,20) and full tube (5,20,30), identity placements."""
    return [
        ("Box", G.UnplacedBox(10.0, 15.0, 20.0)),
        ("Tube", G.UnplacedTube(5.0, 20.0, 30.0, 0.0, TWO_PI)),
    ]


def zoo_shapes():
    """The 12-shape zoo (millimetres, radians unless noted)."""
    return [
        ("Box", G.UnplacedBox(30.0, 40.0, 50.0)),
        ("Orb", G.UnplacedOrb(50.0)),
        ("Sphere", G.UnplacedSphere(10.0, 50.0, 0.2, 4.0, 0.3, 2.0)),
        ("Tube", G.UnplacedTube(10.0, 50.0, 50.0, 0.0, 1.5 * math.pi)),
        ("Cone", G.UnplacedCone(5.0, 20.0, 10.0, 40.0, 30.0, 0.0, TWO_PI)),
        ("Trd", G.UnplacedTrd(15.0, 20.0, 25.0, 30.0, 35.0)),
        ("Trapezoid", G.UnplacedTrapezoid(40.0, 0.1, 0.3, 20.0, 15.0, 25.0, 0.1, 30.0, 20.0, 35.0, 0.1)),
        ("Parallelepiped", G.UnplacedParallelepiped(20.0, 30.0, 40.0, 30.0, 15.0, 30.0)),  # degrees
        ("Paraboloid", G.UnplacedParaboloid(10.0, 30.0, 40.0)),
        ("Torus", G.UnplacedTorus(5.0, 10.0, 50.0, 0.0, TWO_PI)),
        ("Polycone", G.UnplacedPolycone(
            0.0, TWO_PI,
            rmin=[5, 5, 10, 10, 8, 8, 12, 12, 6, 6, 5],
            rmax=[20, 30, 30, 40, 40, 35, 35, 45, 45, 30, 20],
            z=[-50, -40, -30, -20, -10, 0, 10, 20, 30, 40, 50])),
        ("Polyhedron", G.UnplacedPolyhedron(
            0.0, 360.0, 8, z=[-50, -20, 0, 20, 50], rmin=[10, 10, 15, 10, 10], rmax=[30, 40, 45, 40, 30])),
    ]


def random_rotation(rng):
    """Uniform random rotation (Shoemake's quaternion method); returns
    r0..r8 with master = t + R local."""
    u1, u2, u3 = rng.random(3)
    a = math.sqrt(1 - u1) * math.sin(TWO_PI * u2)
    b = math.sqrt(1 - u1) * math.cos(TWO_PI * u2)
    c = math.sqrt(u1) * math.sin(TWO_PI * u3)
    w = math.sqrt(u1) * math.cos(TWO_PI * u3)
    x, y, z = a, b, c
    return [
        1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
        2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
        2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y),
    ]


def random_placement(rng):
    t = rng.uniform(-100.0, 100.0, 3)
    return G.Transformation3D(t[0], t[1], t[2], *random_rotation(rng))


def isotropic(n, rng):
    cost = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, TWO_PI, n)
    sint = np.sqrt(np.maximum(0.0, 1.0 - cost * cost))
    return np.stack([sint * np.cos(phi), sint * np.sin(phi), cost])


def local_to_master(xf, p):
    """InverseTransform (base/Transformation3D.h:601-632), elementwise."""
    t, r = xf.t, xf.r
    out = np.empty_like(p)
    for i in range(3):
        m = np.full(p.shape[1], t[i])
        m = m + p[0] * r[3 * i + 0]
        m = m + p[1] * r[3 * i + 1]
        m = m + p[2] * r[3 * i + 2]
        out[i] = m
    return out


def dir_local_to_master(xf, d):
    r = xf.r
    out = np.empty_like(d)
    for i in range(3):
        m = d[0] * r[3 * i + 0]
        m = m + d[1] * r[3 * i + 1]
        m = m + d[2] * r[3 * i + 2]
        out[i] = m
    return out


def master_to_local(xf, p, direction=False):
    """Transform<generic> (base/Transformation3D.h:469-562), elementwise in
    the reference's operation order."""
    t, r = xf.t, xf.r
    q = p if direction else np.stack([p[0] - t[0], p[1] - t[1], p[2] - t[2]])
    lx = q[0] * r[0]
    ly = q[0] * r[1]
    lz = q[0] * r[2]
    lx = lx + q[1] * r[3]
    ly = ly + q[1] * r[4]
    lz = lz + q[1] * r[5]
    lx = lx + q[2] * r[6]
    ly = ly + q[2] * r[7]
    lz = lz + q[2] * r[8]
    return np.stack([lx, ly, lz])


def sample_points(lo, hi, n, rng, scale=1.5):
    """Uniform points in `scale` x the box [lo, hi] about its centre."""
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    c = 0.5 * (lo + hi)
    h = 0.5 * (hi - lo) * scale
    return rng.uniform(c - h, c + h, size=(n, 3)).T.copy()


def shape_inputs(extent, xf, n, seed, surface_fraction=0.0, surface_sampler=None):
    """Master-frame points/directions for one placed shape.

    Returns dict(p=(3,n) master points, d=(3,n) master directions,
    lp=(3,n) local points = Transform(master), ld=(3,n) local directions).
    """
    rng = np.random.Generator(np.random.Philox(seed))
    lo, hi = extent
    loc = sample_points(lo, hi, n, rng)
    if surface_fraction > 0 and surface_sampler is not None:
        k = int(round(surface_fraction * n))
        loc[:, :k] = surface_sampler(k, rng)
    d = isotropic(n, rng)
    p = local_to_master(xf, loc)
    return {
        "p": p,
        "d": d,
        "lp": master_to_local(xf, p),
        "ld": master_to_local(xf, d, direction=True),
    }
