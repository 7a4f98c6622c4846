"""Writes the navigation golden vectors in tests/golden/ from the real
reference build (oracle/_ref/libvgref.so: the reference's SimpleNavigator,
ABBoxNavigator and ABBoxManager compiled from its own sources, strict FP).

Run in a container that has the reference tree and the built shim:
    python tests/golden/make_golden_nav.py

Cases
  nav_nested.npz  the nested benchmark geometry (64 tube shells, 32 boxes
                  per shell, 8 Trds per box): 1024 track vertices N(0, 5 mm)
                  about the origin and 2048 points uniform in radius
                  90..920 mm, |z| < 800 mm, all with isotropic directions.
  nav_zoo.npz     the zoo navigation geometry without Parallelepiped (the
                  reference's ABBoxManager cannot box it: UnplacedParallelepiped
                  has no Extent), 3072 tracks.
For every point: LocatePoint (level, path), one FindNextBoundaryAndStep per
navigator from the located state (level, path, boundary, step), GetSafety,
and the TraceTrack loop per navigator (steps, final point, summed step,
hash of the (top id, boundary) sequence).
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

OUT = os.path.dirname(os.path.abspath(__file__))


def nested_inputs():
    p1, d1 = W.track_starts(1024, seed=W.SEED_BASE + 21)
    rng = np.random.Generator(np.random.Philox(W.SEED_BASE + 22))
    n = 2048
    r = rng.uniform(90.0, 920.0, n)
    ph = rng.uniform(0.0, 2 * np.pi, n)
    z = rng.uniform(-800.0, 800.0, n)
    p2 = np.stack([r * np.cos(ph), r * np.sin(ph), z])
    d2 = W.isotropic(n, rng)
    return np.concatenate([p1, p2], axis=1), np.concatenate([d1, d2], axis=1)


def zoo_inputs(flat, placed):
    centres = [flat.placed_transform[12 * i:12 * i + 3] for _, i in placed]
    return W.nav_zoo_tracks(3072, centres, seed=W.SEED_BASE + 23)


def record(ref, flat, p, d):
    nav = po.Nav(ref, flat)
    out = {"x": p, "d": d}
    level, path = nav.locate(p)
    out["level"], out["path"] = level, path
    for kind, tag in ((0, "abbox"), (1, "simple")):
        lo, pa, bo, st = nav.step(p, d, level, path, kind=kind)
        out[f"step_{tag}_level"], out[f"step_{tag}_path"] = lo, pa
        out[f"step_{tag}_boundary"], out[f"step_{tag}_step"] = bo, st
        ns, fp, tr, hs = nav.trace(p, d, kind=kind)
        out[f"trace_{tag}_nsteps"], out[f"trace_{tag}_final"] = ns, fp
        out[f"trace_{tag}_travelled"], out[f"trace_{tag}_hash"] = tr, hs
    out["safety"] = nav.safety(p, level, path)
    return out, nav.depth


def main():
    ref = po.reference()
    if ref is None:
        raise SystemExit("oracle/_ref/libvgref.so is not built")
    manifest = {}
    _, flat = W.nested_geometry()
    p, d = nested_inputs()
    out, depth = record(ref, flat, p, d)
    np.savez_compressed(os.path.join(OUT, "nav_nested.npz"), **out)
    manifest["nav_nested.npz"] = {"geometry": "nested", "n": int(p.shape[1]), "depth": depth,
                                  "mean_steps_abbox": float(out["trace_abbox_nsteps"].mean())}
    names = [n for n, _ in W.NAV_ZOO if n != "Parallelepiped"]
    _, flat, placed = W.nav_zoo_geometry(names=names)
    p, d = zoo_inputs(flat, placed)
    out, depth = record(ref, flat, p, d)
    np.savez_compressed(os.path.join(OUT, "nav_zoo.npz"), **out)
    manifest["nav_zoo.npz"] = {"geometry": "zoo", "shapes": names, "n": int(p.shape[1]), "depth": depth,
                               "mean_steps_abbox": float(out["trace_abbox_nsteps"].mean())}
    with open(os.path.join(OUT, "manifest_nav.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    print(json.dumps(manifest, indent=1))


if __name__ == "__main__":
    main()
