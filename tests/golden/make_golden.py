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


# The reference's Polyhedron dereferences out of range for points on or next
# to the outer z-planes (FindZSegmentKernel returns nz-1 there,
# volumes/kernel/PolyhedronImplementation.h:358-371) and the process
# segfaults; its vectors come from tests/golden/make_golden_polyhedron.py,
# which screens points in a child process.
SKIP = {"Polyhedron"}


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
    p = W.local_to_master(xf, loc)
    lp = W.master_to_local(xf, p)
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
    manifest = {"generator": "tests/golden/make_golden.py", "oracle": "reference Scalar backend, strict FP",
                "N": N, "cases": {}}
    rng = np.random.Generator(np.random.Philox(W.SEED_BASE + 100))
    shapes = W.config1_shapes() + W.zoo_shapes()
    seen = set()
    for idx, (name, u) in enumerate(shapes):
        if name in SKIP:
            continue
        tag = name if name not in seen else name + "C2"
        seen.add(name)
        for kind in ("id", "rot"):
            xf = G.Transformation3D() if kind == "id" else W.random_placement(rng)
            fn, fm = make_case(ref, ref_fast, tag, u, xf, W.SEED_BASE + 1000 + 10 * idx + (kind == "rot"))
            manifest["cases"][fn] = {"shape": name, "params": u.params, "placement": kind,
                                     "fast_math_mismatches": fm}
            print(fn, fm)
    with open(os.path.join(OUT, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
