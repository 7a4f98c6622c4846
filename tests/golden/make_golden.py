"""Writes the shape-query golden vectors in tests/golden/ from the real
reference build (oracle/_ref/libvgref.so, the reference's Scalar backend
compiled from its own sources with strict FP).

Run in a container that has the reference tree and the built shim:
    python tests/golden/make_golden.py
Each shape gets two cases, identity placement and a random placement, with
N points: 70% uniform in 1.5x the extent, 15% on the surface jittered by
sigma=1e-7 mm and 15% jittered by sigma=5e-10 mm (inside the 1e-9 mm
tolerance band).  Surface points are made by moving a uniform point along its
direction by the reference's own DistanceToIn/DistanceToOut.

Outputs (strict reference):  contains, inside, dist_in (mother frame, all
points), dist_out (local frame, all points), safe_in, safe_out,
placed_dist_out.  Also stored: the count of lanes where the reference built
with its Release flags (-ffast-math) disagrees, per query.
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle as po  # noqa: E402
from vecgeom_amd import workloads as W  # noqa: E402
from vecgeom_amd import geometry as G  # noqa: E402

N = 1024
OUT = os.path.dirname(os.path.abspath(__file__))
QUERIES = {
    "contains": po.Q_CONTAINS,
    "inside": po.Q_INSIDE,
    "dist_in": po.Q_DIST_IN,
    "dist_out": po.Q_DIST_OUT,
    "safe_in": po.Q_SAFE_IN,
    "safe_out": po.Q_SAFE_OUT,
    "placed_dist_out": po.Q_PLACED_DIST_OUT,
}


# The reference's Polyhedron indexes its segment array out of range in
# Contains/Inside for points inside its bounding tube but not strictly between
# the end planes (FindZSegment gives -1 or n there;
# volumes/kernel/PolyhedronImplementation.h:731-770, 772-830), which can crash
# the process.  Sampled points whose local z lies within 1e-8 mm of an end
# plane are moved 2e-8 mm beyond it (outside the bounding tube, which reaches
# kTolerance past the planes).
SKIP = set()
END_BAND = 1e-8


def screen_polyhedron_ends(unplaced, loc):
    nz = int(unplaced.params[3])
    z = unplaced.params[4:4 + nz]
    lo, hi = z[0], z[-1]
    near_lo = np.abs(loc[2] - lo) < END_BAND
    near_hi = np.abs(loc[2] - hi) < END_BAND
    loc[2, near_lo] = lo - 2 * END_BAND
    loc[2, near_hi] = hi + 2 * END_BAND
    return int(near_lo.sum() + near_hi.sum())


def same(a, b):
    if a.dtype == np.float64:
        return (a == b) | (np.isnan(a) & np.isnan(b))
    return a == b


def make_case(ref, ref_fast, name, unplaced, xf, seed):
    xf12 = xf.array()
    h = ref.placed(unplaced.type, unplaced.params, xf12)
    lo, hi = h.extent()
    rng = np.random.Generator(np.random.Philox(seed))
    loc = W.sample_points(lo, hi, N, rng)
    d = W.isotropic(N, rng)
    ld = W.master_to_local(xf, d, direction=True)
    # Surface band: move points onto the surface with the reference's own
    # distances, then jitter.
    p_tmp = W.local_to_master(xf, loc)
    lp_tmp = W.master_to_local(xf, p_tmp)
    ins = ref.placed(unplaced.type, unplaced.params, xf12).query(po.Q_CONTAINS, p_tmp)
    din = h.query(po.Q_DIST_IN, p_tmp, d, np.full(N, 1e30))
    dout = h.query(po.Q_DIST_OUT, lp_tmp, ld, np.full(N, 1e30))
    step = np.where(ins == 1, dout, din)
    ok = np.isfinite(step) & (step >= 0)
    n1 = int(0.7 * N)
    n2 = int(0.85 * N)
    for k in range(n1, N):
        if ok[k]:
            sig = 1e-7 if k < n2 else 5e-10
            loc[:, k] = lp_tmp[:, k] + step[k] * ld[:, k] + rng.normal(0.0, sig, 3)
    if unplaced.type == G.VG_POLYHEDRON:
        screen_polyhedron_ends(unplaced, loc)
    p = W.local_to_master(xf, loc)
    lp = W.master_to_local(xf, p)
    if unplaced.type == G.VG_POLYHEDRON:
        assert screen_polyhedron_ends(unplaced, lp) == 0
    sm = np.full(N, 1e30)
    sm[::7] = rng.uniform(0.0, 50.0, len(sm[::7]))  # some finite step limits
    out = {"p": p, "d": d, "lp": lp, "ld": ld, "sm": sm, "xf": xf12,
           "type": np.int32(unplaced.type), "params": np.array(unplaced.params)}
    fast_mismatch = {}
    hf = ref_fast.placed(unplaced.type, unplaced.params, xf12) if ref_fast else None
    for qn, q in QUERIES.items():
        if qn in ("dist_out", "safe_out"):
            args = (lp, ld if qn == "dist_out" else None)
        else:
            args = (p, d if qn in ("dist_in", "placed_dist_out") else None)
        stepm = sm if qn in ("dist_in", "dist_out", "placed_dist_out") else None
        r = h.query(q, args[0], args[1], stepm)
        out[qn] = r
        if hf is not None:
            rf = hf.query(q, args[0], args[1], stepm)
            fast_mismatch[qn] = int((~same(r, rf)).sum())
    path = os.path.join(OUT, f"shape_{name}_{'id' if xf12[3:].tolist() == [1,0,0,0,1,0,0,0,1] and not any(xf12[:3]) else 'rot'}.npz")
    np.savez_compressed(path, **out)
    return os.path.basename(path), fast_mismatch


def main():
    ref = po.reference()
    if ref is None:
        raise SystemExit("oracle/_ref/libvgref.so is not built (needs the reference tree)")
    ref_fast = po.reference(fast=True)
    only = set(sys.argv[1:])  # shape names to regenerate (default: all)
    manifest = {"generator": "tests/golden/make_golden.py", "oracle": "reference Scalar backend, strict FP",
                "N": N, "cases": {}}
    mpath = os.path.join(OUT, "manifest.json")
    if only and os.path.exists(mpath):
        with open(mpath) as f:
            manifest = json.load(f)
    rng = np.random.Generator(np.random.Philox(W.SEED_BASE + 100))
    shapes = W.config1_shapes() + W.zoo_shapes() + W.boolean_shapes()
    seen = set()
    for idx, (name, u) in enumerate(shapes):
        if name in SKIP:
            continue
        tag = name if name not in seen else name + "C2"
        seen.add(name)
        for kind in ("id", "rot"):
            xf = G.Transformation3D() if kind == "id" else W.random_placement(rng)
            if only and name not in only:
                continue
            fn, fm = make_case(ref, ref_fast, tag, u, xf, W.SEED_BASE + 1000 + 10 * idx + (kind == "rot"))
            manifest["cases"][fn] = {"shape": name, "params": u.params, "placement": kind,
                                     "fast_math_mismatches": fm}
            print(fn, fm)
    with open(os.path.join(OUT, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
