"""Writes the surface-normal golden vectors tests/golden/normal_<shape>.npz
from the real reference build (oracle/_ref): VPlacedVolume::Normal on local
points, and the USolids DistanceToOut(point, direction, normal&, convex&,
stepMax) on local rays (evaluated by the shim with the reference's own
virtuals and Vector3D arithmetic).

Run in a container that has the reference tree and the built shim:
    python tests/golden/make_normal_golden.py

Points per shape: N uniform in 1.5x the extent, N exactly on the surface
(interior points moved along isotropic directions by the reference's own
DistanceToOut) and N of those jittered by N(0, (2e-10 mm)^2), inside the
surface tolerances; the Box also gets its edges and corners.  Rays: N interior
points with isotropic directions.  Outputs come from the strict build
(-O2 -fno-fast-math), except where the Cone's normal kernel would assert (a
point on no surface, volumes/kernel/ConeImplementation.h:2112): those come
from the Release build (NDEBUG), and `from_release` marks them.  Which points
those are is decided by the oracle's valid flag (the strict build cannot be
asked without aborting); tests/test_oracle_normal.py checks that flag against
the strict build on every other point and the Release build on these.
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle as po  # noqa: E402
from vecgeom_amd import geometry as G  # noqa: E402
from vecgeom_amd import workloads as W  # noqa: E402

N = 512
OUT = os.path.dirname(os.path.abspath(__file__))
I = G.Transformation3D().array()


def normal_shapes():
    """The zoo shapes with a reference Normal, and variants that reach the
    branches the zoo leaves out (no inner radius, phi cuts)."""
    zoo = [(n, u) for n, u in W.zoo_shapes() if u.type in (1, 2, 3, 4, 5, 6, 7, 11)]
    return zoo + [
        ("TubeSolidPhi", G.UnplacedTube(0.0, 20.0, 30.0, 0.5, 2.0)),
        ("ConePhi", G.UnplacedCone(5.0, 20.0, 10.0, 40.0, 30.0, 0.3, 2.5)),
        ("ConeSolid", G.UnplacedCone(0.0, 20.0, 0.0, 40.0, 30.0, 0.0, 2 * np.pi)),
        ("SphereSolid", G.UnplacedSphere(0.0, 40.0, 0.0, 2 * np.pi, 0.0, np.pi)),
        ("SphereNoRmin", G.UnplacedSphere(0.0, 40.0, 0.5, 3.0, 0.4, 1.5)),
    ]


def box_edges(d, rng, n):
    """Points on the Box's edges and corners (coordinates exactly +-d)."""
    d = np.asarray(d, np.float64)
    pts = rng.uniform(-d, d, size=(n, 3))
    for k in range(n):
        fix = rng.choice(3, size=2 + (k % 2), replace=False)
        pts[k, fix] = np.where(rng.random(len(fix)) < 0.5, -d[fix], d[fix])
    return pts.T.copy()


def inputs(name, unplaced, ref, seed):
    rng = np.random.Generator(np.random.Philox(seed))
    h = ref.placed(unplaced.type, unplaced.params, I)
    lo, hi = h.extent()
    uni = W.sample_points(lo, hi, N, rng)
    got = []
    while sum(g.shape[1] for g in got) < 2 * N:
        c = W.sample_points(lo, hi, 8 * N, rng, scale=1.0)
        got.append(c[:, h.query(po.Q_CONTAINS, c).astype(bool)])
    interior = np.concatenate(got, axis=1)[:, :2 * N]
    d = W.isotropic(N, rng)
    t = h.query(po.Q_DIST_OUT, interior[:, :N], d, np.full(N, 1e30))
    t = np.where(np.isfinite(t) & (t >= 0), t, 0.0)
    surf = interior[:, :N] + d * t
    jit = surf + rng.normal(0.0, 2e-10, size=surf.shape)
    parts = [uni, surf, jit]
    if unplaced.type == G.VG_BOX:
        parts.append(box_edges(unplaced.params[:3], rng, N // 4))
    lp = np.concatenate(parts, axis=1)
    rp = interior[:, N:]
    rd = W.isotropic(rp.shape[1], rng)
    sm = np.where(rng.random(rp.shape[1]) < 0.2, rng.uniform(0.0, 20.0, rp.shape[1]), np.inf)
    return lp, rp, rd, sm


def evaluate(ref, ref_fast, unplaced, lp, rp, rd, sm):
    """The reference's outputs; Cone lanes on no surface from the Release build."""
    h, hf = ref.placed(unplaced.type, unplaced.params, I), ref_fast.placed(unplaced.type, unplaced.params, I)
    o = po.oracle().placed(unplaced.type, unplaced.params, I)
    if unplaced.type == G.VG_CONE:
        # The oracle decides which lanes lie on no surface; the Release build
        # answers those (its Assert is compiled out), the strict build the rest.
        _, vo = o.normal(lp)
        strict = vo.astype(bool)
        nrm, valid = np.zeros((3, lp.shape[1])), np.zeros(lp.shape[1], np.uint8)
        nrm[:, strict], valid[strict] = h.normal(lp[:, strict])
        nrm[:, ~strict], valid[~strict] = hf.normal(lp[:, ~strict])
        _, _, dvo, _ = o.distance_to_out_normal(rp, rd, sm)
        dstrict = dvo.astype(bool)
        dist, dn, dv, dc = (np.zeros(rp.shape[1]), np.zeros((3, rp.shape[1])), np.zeros(rp.shape[1], np.uint8),
                            np.zeros(rp.shape[1], np.uint8))
        for sel, lib in ((dstrict, h), (~dstrict, hf)):
            if sel.any():
                dist[sel], dn[:, sel], dv[sel], dc[sel] = lib.distance_to_out_normal(rp[:, sel], rd[:, sel], sm[sel])
        return nrm, valid, dist, dn, dv, dc, (~strict).astype(np.uint8), (~dstrict).astype(np.uint8)
    nrm, valid = h.normal(lp)
    dist, dn, dv, dc = h.distance_to_out_normal(rp, rd, sm)
    return nrm, valid, dist, dn, dv, dc, np.zeros(lp.shape[1], np.uint8), np.zeros(rp.shape[1], np.uint8)


def main():
    ref, ref_fast = po.reference(), po.reference(fast=True)
    if ref is None or ref_fast is None:
        sys.exit("oracle/_ref not built: run __graft_entry__.build() in a container with the reference tree")
    manifest = {}
    for k, (name, u) in enumerate(normal_shapes()):
        lp, rp, rd, sm = inputs(name, u, ref, 6100 + k)
        nrm, valid, dist, dn, dv, dc, rel, drel = evaluate(ref, ref_fast, u, lp, rp, rd, sm)
        path = os.path.join(OUT, f"normal_{name}.npz")
        np.savez_compressed(path, type=np.int32(u.type), params=np.asarray(u.params, np.float64), lp=lp,
                            normal=nrm, valid=valid, from_release=rel, rp=rp, rd=rd, step_max=sm, dist=dist,
                            dist_normal=dn, dist_valid=dv, convex=dc, dist_from_release=drel)
        manifest[name] = {"points": int(lp.shape[1]), "rays": int(rp.shape[1]), "valid": float(valid.mean()),
                          "from_release": int(rel.sum()) + int(drel.sum())}
        print(name, manifest[name])
    with open(os.path.join(OUT, "normal_manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
