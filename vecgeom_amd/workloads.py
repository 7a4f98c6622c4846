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


def boolean_shapes():
    """Boolean volumes (UnplacedBooleanVolume): union, intersection and
    subtraction with partially overlapping operands, a subtraction that cuts
    a hole, and an intersection with a Polyhedron operand (an operand with a
    variable-length parameter block).  The left operand sits at the
    Boolean's origin, as in the reference's examples; the right one is
    translated and rotated."""
    shapes = dict(zoo_shapes())
    ident = G.Transformation3D()
    rot = G.Transformation3D.from_array(
        [10.0, -5.0, 4.0] + list(random_rotation(np.random.Generator(np.random.Philox(SEED_BASE + 40)))))
    box = G.UnplacedBox(30.0, 20.0, 25.0)
    return [
        ("BoolUnion", G.UnplacedBooleanVolume(G.kUnion, box, ident, G.UnplacedTube(5.0, 20.0, 30.0, 0.0, TWO_PI),
                                              G.Transformation3D(15.0, 5.0, -3.0))),
        ("BoolInter", G.UnplacedBooleanVolume(G.kIntersection, box, ident, G.UnplacedOrb(28.0), rot)),
        ("BoolSub", G.UnplacedBooleanVolume(G.kSubtraction, box, ident, G.UnplacedTube(5.0, 20.0, 30.0, 0.0, TWO_PI),
                                            rot)),
        ("BoolHole", G.UnplacedBooleanVolume(G.kSubtraction, G.UnplacedCone(0.0, 30.0, 0.0, 20.0, 30.0, 0.0, TWO_PI),
                                             ident, G.UnplacedTube(0.0, 6.0, 8.0, 0.0, TWO_PI),
                                             G.Transformation3D(3.0, 2.0, -4.0))),
        ("BoolPolyInter", G.UnplacedBooleanVolume(G.kIntersection, shapes["Polyhedron"], ident,
                                                  G.UnplacedBox(25.0, 25.0, 25.0), G.Transformation3D(5.0, 5.0, 5.0))),
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
        loc = loc[:, rng.permutation(n)]
    d = isotropic(n, rng)
    p = local_to_master(xf, loc)
    return {
        "p": p,
        "d": d,
        "lp": master_to_local(xf, p),
        "ld": master_to_local(xf, d, direction=True),
    }


# ------------------------------------------------------- surface-band inputs
SURFACE_FRACTION = 0.25
SURFACE_SIGMA = 1e-7  # mm


def surface_sampler(extent, contains, distance_to_out, sigma=SURFACE_SIGMA, seed_points=None):
    """A sampler of local-frame points on a shape's surface, jittered by
    N(0, sigma^2) per coordinate: interior points (rejection-sampled
    uniformly in the extent with `contains`) moved along isotropic directions
    by `distance_to_out`.  contains(lp) -> bool array and
    distance_to_out(lp, ld) -> float array take (3, m) local points (and
    directions); any implementation of the two queries will do (the FP64
    kernels, or the oracle)."""
    lo, hi = extent

    def sample(k, rng):
        got = []
        need = k
        while need > 0:
            m = max(4 * need, 1024)
            cand = rng.uniform(np.asarray(lo, np.float64), np.asarray(hi, np.float64), size=(m, 3)).T.copy()
            inside = np.asarray(contains(cand), dtype=bool)
            cand = cand[:, inside][:, :need]
            got.append(cand)
            need -= cand.shape[1]
        lp = np.concatenate(got, axis=1)
        ld = isotropic(k, rng)
        t = np.asarray(distance_to_out(lp, ld), np.float64)
        t = np.where(np.isfinite(t) & (t >= 0), t, 0.0)
        return lp + ld * t + rng.normal(0.0, sigma, size=lp.shape)

    return sample


def mixed_inputs(extent, xf, n, seed, sampler, surface_fraction=SURFACE_FRACTION):
    """The mixed-precision workload of one placed shape: like shape_inputs,
    with `surface_fraction` of the points on the surface (sampler above) and
    points and directions stored as float32.  Returns dict(p32, d32 (3, n)
    float32 master-frame arrays; p, d, lp, ld: the float32 values promoted
    to double, and their local-frame images, for evaluating the FP64 form
    and the oracle on exactly the same inputs).  Float rounding moves a
    surface point by up to half a float ulp of its master coordinates
    (about 4e-6 mm at 100 mm), which dominates the jitter."""
    inp = shape_inputs(extent, xf, n, seed, surface_fraction, sampler)
    p32 = inp["p"].astype(np.float32)
    d32 = inp["d"].astype(np.float32)
    p = p32.astype(np.float64)
    d = d32.astype(np.float64)
    return {"p32": p32, "d32": d32, "p": p, "d": d, "lp": master_to_local(xf, p),
            "ld": master_to_local(xf, d, direction=True)}


# ---------------------------------------------------------------- navigation
# Nested benchmark geometry.  The nominal description (world box
# 2000x2000x4000 mm, 64 tube shells with rmin 100..900 mm and dz 1800 mm,
# 32 boxes of 20x40x200 mm per shell, 8 Trds per box) cannot be realised
# without overlaps: 64 disjoint shells between r=100 and r=900 leave at most
# 12.7 mm of radial thickness, and 32 boxes 40 mm wide do not fit around the
# innermost shells.  The realisation used here, recorded so that every run
# builds the same tree:
#   world   box half-lengths (1000, 1000, 2000)
#   shell k tube rmin_k = 100 + 800 k / 63, rmax_k = rmin_k + 12,
#           dz = 1800, full phi, identity placement (k = 0..63)
#   box     half-lengths (4, 20, 100): thin side radial, 40 mm tangential,
#           200 mm along z; 32 per shell = 8 azimuths x 4 z-slices
#           (z = -600, -200, 200, 600); azimuth a*45 deg + a random offset
#           per shell; centre radius rmin_k + 6; rotated about z by its
#           azimuth (a "rotated box")
#   trd     dx1 3, dx2 2, dy1 15, dy2 10, dz 12; 8 per box along z at
#           z = -87.5 + 25 j, translation only
# All boxes share one logical volume and so do the Trds, as placements of
# one logical volume do in the reference.
NESTED = {"world": (1000.0, 1000.0, 2000.0), "shells": 64, "rmin0": 100.0, "rmin1": 900.0, "thick": 12.0,
          "dz": 1800.0, "azimuths": 8, "zslices": (-600.0, -200.0, 200.0, 600.0), "box": (4.0, 20.0, 100.0),
          "trd": (3.0, 2.0, 15.0, 10.0, 12.0), "trds": 8}


def nested_geometry(seed=SEED_BASE + 4, shells=None):
    """Builds the nested geometry; returns (GeoManager, flat arrays)."""
    nshell = NESTED["shells"] if shells is None else shells
    rng = np.random.Generator(np.random.Philox(seed))
    gm = G.GeoManager()
    world_lv = G.LogicalVolume("world", G.UnplacedBox(*NESTED["world"]), gm)
    box_lv = G.LogicalVolume("box", G.UnplacedBox(*NESTED["box"]), gm)
    trd_lv = G.LogicalVolume("trd", G.UnplacedTrd(*NESTED["trd"]), gm)
    for j in range(NESTED["trds"]):
        box_lv.PlaceDaughter(f"trd{j}", trd_lv, G.Transformation3D(0.0, 0.0, -87.5 + 25.0 * j))
    na = NESTED["azimuths"]
    for k in range(nshell):
        rmin = NESTED["rmin0"] + (NESTED["rmin1"] - NESTED["rmin0"]) * k / 63.0
        rmax = rmin + NESTED["thick"]
        shell_lv = G.LogicalVolume(f"shell{k}", G.UnplacedTube(rmin, rmax, NESTED["dz"], 0.0, TWO_PI), gm)
        phi0 = rng.uniform(0.0, TWO_PI / na)
        rc = rmin + 0.5 * NESTED["thick"]
        for b, zc in enumerate(NESTED["zslices"]):
            for a in range(na):
                phi = phi0 + a * TWO_PI / na
                c, s = math.cos(phi), math.sin(phi)
                xf = G.Transformation3D(rc * c, rc * s, zc, c, -s, 0.0, s, c, 0.0, 0.0, 0.0, 1.0)
                shell_lv.PlaceDaughter(f"box{k}_{b}_{a}", box_lv, xf)
        world_lv.PlaceDaughter(f"shell{k}", shell_lv, G.Transformation3D())
    world = world_lv.Place("world")
    gm.SetWorld(world)
    return gm, gm.CloseGeometry()


def track_starts(n, seed=SEED_BASE + 5, sigma=5.0):
    """Vertices N(0, sigma^2) per axis and isotropic unit directions."""
    rng = np.random.Generator(np.random.Philox(seed))
    p = rng.normal(0.0, sigma, size=(3, n))
    d = isotropic(n, rng)
    return p, d


# Shapes placed as volumes in the zoo navigation geometry, with a point of
# each shape's interior (local frame) that holds an Orb r=3 daughter.
NAV_ZOO = [
    ("Box", (0.0, 0.0, 0.0)),
    ("Orb", (10.0, 0.0, 0.0)),
    ("Sphere", (30.0 * math.cos(1.0), 30.0 * math.sin(1.0), 0.0)),
    ("Tube", (0.0, 30.0, 0.0)),
    ("Trd", (0.0, 0.0, 5.0)),
    ("Trapezoid", (0.0, 0.0, 0.0)),
    ("Parallelepiped", (0.0, 0.0, 0.0)),
    ("Paraboloid", (0.0, 0.0, 0.0)),
]


def nav_zoo_geometry(seed=SEED_BASE + 6, names=None):
    """World box of half-length 400 mm holding one placement of each shape
    in NAV_ZOO on a 3x3 grid (pitch 220 mm, z = 0) with a random rotation;
    each holds an Orb r=3 daughter at its listed interior point.  Returns
    (GeoManager, flat arrays, list of (name, placed id))."""
    rng = np.random.Generator(np.random.Philox(seed))
    shapes = dict(zoo_shapes())
    names = [n for n, _ in NAV_ZOO] if names is None else names
    inner = dict(NAV_ZOO)
    gm = G.GeoManager()
    world_lv = G.LogicalVolume("world", G.UnplacedBox(400.0, 400.0, 400.0), gm)
    orb_lv = G.LogicalVolume("orb", G.UnplacedOrb(3.0), gm)
    slots = [(x, y) for x in (-220.0, 0.0, 220.0) for y in (-220.0, 0.0, 220.0)]
    placed = []
    for k, name in enumerate(names):
        lv = G.LogicalVolume(name, shapes[name], gm)
        lv.PlaceDaughter(name + "_orb", orb_lv, G.Transformation3D(*inner[name]))
        r = random_rotation(rng)
        x, y = slots[k]
        pv = world_lv.PlaceDaughter(name, lv, G.Transformation3D(x, y, 0.0, *r))
        placed.append((name, pv))
    gm.SetWorld(world_lv.Place("world"))
    flat = gm.CloseGeometry()
    return gm, flat, [(n, pv.id) for n, pv in placed]


def nav_zoo_tracks(n, centres, seed=SEED_BASE + 7):
    """A third of the tracks start uniform in |x|,|y|,|z| < 380 with
    isotropic directions, a third start there aimed at a random shape
    centre (jittered by 20 mm), a third start at a random shape centre
    jittered by 10 mm with isotropic directions."""
    rng = np.random.Generator(np.random.Philox(seed))
    p = rng.uniform(-380.0, 380.0, size=(3, n))
    d = isotropic(n, rng)
    group = rng.integers(0, 3, n)
    cs = np.asarray(centres, dtype=np.float64)
    c = cs[rng.integers(0, len(cs), n)].T + rng.normal(0.0, 20.0, (3, n))
    v = c - p
    v /= np.sqrt((v * v).sum(axis=0))
    d[:, group == 1] = v[:, group == 1]
    c2 = cs[rng.integers(0, len(cs), n)].T + rng.normal(0.0, 10.0, (3, n))
    p[:, group == 2] = c2[:, group == 2]
    return p, d
