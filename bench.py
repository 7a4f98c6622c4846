"""Shape-query benchmark on one MI355X.

Workload (the full-zoo FP64 configuration): every shape of the zoo that has
device kernels, each placed with a random rotation and a translation in
+-100 mm, queried with 2^22 seeded points and isotropic directions (SOA3D,
float64).  One step = Contains and Inside on all points, DistanceToIn and
SafetyToIn on the points outside the shape (step max 1e30), DistanceToOut and
SafetyToOut on the points inside (local frame), for every shape.  Inputs are
resident in HBM before the timed region starts.

Prints one JSON line: throughput (Gqueries/s), the roofline of the dominant
kernel (HBM bytes / launch duration from HIP events recorded on the launch
stream inside the timed region), the plain-FP64-GPU denominator, and the
oracle (C++ restatement of the reference Scalar path) timed on host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "shape Gqueries/s + nav steps/s vs plain FP64 GPU at 1e-9 mm; p99 basket latency"
HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per query: each input read once, each output written once.
BYTES = {"contains": 25, "inside": 28, "dist_in": 64, "dist_out": 64, "safe_in": 32, "safe_out": 32}
# The same with float32 inputs (step_max and outputs stay as they are).
BYTES_F32 = {"contains": 13, "inside": 16, "dist_in": 40, "dist_out": 40, "safe_in": 20, "safe_out": 20}


# (shape, query) pairs left out of the step: the reference's Sphere
# DistanceToIn has an empty body (volumes/kernel/SphereImplementation.h:88-94)
# and returns kInfinity without reading its inputs, so timing it measures
# nothing but a store.
EXCLUDED = {("Sphere", "dist_in")}

PROFILE = os.path.join(ROOT, "profiles", "shape_zoo_fp64_kernels.md")
BATCH_PROFILE = os.path.join(ROOT, "profiles", "shape_zoo_batched_kernels.md")


# Weight classes of the batched launch (vg_batch.hip ClassOf / TaskClass; the
# variable-length shapes' blocks of this workload fit in LDS: class 3; the
# Torus's Contains, Inside and safeties run in the medium class).
SHAPE_CLASS = {"Cone": 1, "Sphere": 1, "Torus": 2, "Polycone": 3, "Polyhedron": 3}


def shape_class(name, query=None):
    """Weight class of a (shape, batch query code) task."""
    from vecgeom_amd import batch as B
    if name == "Torus" and query is not None and query not in B.DIRECTIONAL:
        return 1
    return SHAPE_CLASS.get(name, 0)


def ClassIT(c, it):
    """Items per lane of weight class c for a batch launched with `it`."""
    return it if c == 0 else min(it, 2)


def profiled_class_traffic(cls, items):
    """HBM bytes per launch of a batch class kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (bytes per query, scaled by a
    device copy of known size) times its queries; None if not profiled."""
    try:
        with open(BATCH_PROFILE) as f:
            for line in f:
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cells) >= 6 and cells[0] == f"class {cls}":
                    return int(round(float(cells[5]) * items))
    except (OSError, ValueError):
        pass
    return None
KERNEL_SHAPE = {"Parallelepiped": "Para", "Paraboloid": "Parab", "Trapezoid": "Trap"}


def profiled_traffic(shape, query, n):
    """HBM bytes per launch of a shape kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (bytes per query, scaled by a device copy
    of known size) times this launch's queries; None if not profiled."""
    key = f"{KERNEL_SHAPE.get(shape, shape)}/{query}"
    try:
        with open(PROFILE) as f:
            for line in f:
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cells) == 7 and cells[1] == key:
                    return int(round(float(cells[6]) * n))
    except OSError:
        pass
    return None


# Kernels other than the per-lane ShapeQueryKernel.
KERNEL = {("Polyhedron", "dist_in"): "WalkKernel<Polyhedron,DistInOps>",
          ("Polyhedron", "dist_out"): "WalkKernel<Polyhedron,DistOutOps>",
          ("Torus", "dist_in"): "CompactDistInKernel<Torus>"}


def kernel_name(shape, query):
    """The device kernel that serves one (shape, query) launch of the step."""
    return KERNEL.get((shape, query), f"ShapeQueryKernel<{shape},{query}>")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--points", type=int, default=1 << 22)
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="target CPU-baseline duration (shapes)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-FP64-form denominator")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched and FP32-storage forms")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision workload and forms")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-storage batched form")
    ap.add_argument("--no-boolean", action="store_true", help="skip the Boolean-volume queries")
    ap.add_argument("--no-normal", action="store_true", help="skip the surface-normal queries")
    ap.add_argument("--batch-launch", default=None, metavar="IT,BLOCKS",
                    help="launch shape of the FP64 batched form instead of autotuning it (items per lane, "
                         "persistent workgroups or 0 for one per tile): profiling runs use the shape the "
                         "autotuner picked in the timed run")
    ap.add_argument("--nav-tracks", type=int, default=1 << 20, help="tracks of the nested-geometry navigation run")
    ap.add_argument("--nav-steps", type=int, default=3, help="timed trace launches")
    ap.add_argument("--no-nav", action="store_true")
    ap.add_argument("--serve-baskets", type=int, default=2000, help="baskets per serving run")
    ap.add_argument("--serve-steps", type=int, default=8, help="navigation steps per served basket")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run a device-to-device copy of MIB MiB (for HBM counter scaling)")
    return ap.parse_args()


class Workload:
    def __init__(self, n, device):
        import torch

        from vecgeom_amd import capi
        from vecgeom_amd import geometry as G
        from vecgeom_amd import queries as Q
        from vecgeom_amd import workloads as W

        self.Q = Q
        rng = np.random.Generator(np.random.Philox(W.SEED_BASE + 2))
        all_shapes = W.zoo_shapes()
        self.shapes = [(nm, u) for nm, u in all_shapes if capi.lib().vg_shape_supported(u.type)]
        self.missing = [nm for nm, u in all_shapes if not capi.lib().vg_shape_supported(u.type)]
        xfs_all = [W.random_placement(rng) for _ in all_shapes]
        self.xfs = [xf for (nm, u), xf in zip(all_shapes, xfs_all) if capi.lib().vg_shape_supported(u.type)]
        self.geom, self.ids, _ = G.single_shape_geometry([u for _, u in self.shapes], self.xfs)
        self.n = n
        self.items = []
        t = lambda a: tuple(torch.from_numpy(np.ascontiguousarray(a[i])).to(device) for i in range(3))
        for k, ((name, u), xf, pid) in enumerate(zip(self.shapes, self.xfs, self.ids)):
            inp = W.shape_inputs(self.geom.placed_extent(pid), xf, n, seed=W.SEED_BASE + 2 + 17 * k)
            p, d, lp, ld = t(inp["p"]), t(inp["d"]), t(inp["lp"]), t(inp["ld"])
            c = Q.contains(self.geom, pid, p).bool()
            iin = torch.nonzero(c).flatten()
            iout = torch.nonzero(~c).flatten()
            pick = lambda s, idx: tuple(a[idx].contiguous() for a in s)
            it = {
                "name": name, "pid": pid, "p": p,
                "pout": pick(p, iout), "dout": pick(d, iout), "smout": torch.full((iout.numel(),), 1e30,
                                                                                   dtype=torch.float64, device=device),
                "pin": pick(lp, iin), "din": pick(ld, iin), "smin": torch.full((iin.numel(),), 1e30,
                                                                                dtype=torch.float64, device=device),
                "host": {k2: inp[k2] for k2 in ("p", "d", "lp", "ld")}, "inside_idx": iin.cpu().numpy(),
                "outside_idx": iout.cpu().numpy(),
            }
            n_in, n_out = iin.numel(), iout.numel()
            it["out"] = {
                "contains": torch.empty(n, dtype=torch.uint8, device=device),
                "inside": torch.empty(n, dtype=torch.int32, device=device),
                "dist_in": torch.empty(n_out, dtype=torch.float64, device=device),
                "safe_in": torch.empty(n_out, dtype=torch.float64, device=device),
                "dist_out": torch.empty(n_in, dtype=torch.float64, device=device),
                "safe_out": torch.empty(n_in, dtype=torch.float64, device=device),
            }
            it["count"] = {"contains": n, "inside": n, "dist_in": n_out, "safe_in": n_out, "dist_out": n_in,
                           "safe_out": n_in}
            for q in list(it["count"]):
                if (name, q) in EXCLUDED:
                    del it["count"][q]
            self.items.append(it)
        self.names = [it["name"] for it in self.items]
        self.names_of_pid = {it["pid"]: it["name"] for it in self.items}
        self.queries_per_step = sum(sum(it["count"].values()) for it in self.items)
        self.bytes_per_step = sum(c * BYTES[q] for it in self.items for q, c in it["count"].items())
        self.bytes_per_step_f32 = sum(c * BYTES_F32[q] for it in self.items for q, c in it["count"].items())

    def tasks(self, fp32=False):
        """The step as a table of tasks (vg_batch): the same inputs, or float32
        copies of them, and the same output tensors."""
        import torch

        from vecgeom_amd import batch as B

        cv = (lambda s: tuple(a.to(torch.float32) for a in s)) if fp32 else (lambda s: s)
        out = []
        for it in self.items:
            o, pid = it["out"], it["pid"]
            p, pout, dout, pin, din = cv(it["p"]), cv(it["pout"]), cv(it["dout"]), cv(it["pin"]), cv(it["din"])
            tasks = {"contains": B.Task(pid, B.CONTAINS, p, out=o["contains"]),
                     "inside": B.Task(pid, B.INSIDE, p, out=o["inside"]),
                     "dist_in": B.Task(pid, B.DISTANCE_TO_IN, pout, dout, it["smout"], out=o["dist_in"]),
                     "safe_in": B.Task(pid, B.SAFETY_TO_IN, pout, out=o["safe_in"]),
                     "dist_out": B.Task(pid, B.DISTANCE_TO_OUT, pin, din, it["smin"], out=o["dist_out"]),
                     "safe_out": B.Task(pid, B.SAFETY_TO_OUT, pin, out=o["safe_out"])}
            out += [t for q, t in tasks.items() if q in it["count"]]
        return out

    def launches(self):
        """(key, callable) for every launch of one step."""
        return [(key, fn) for key, fn in self._launches() if key[1] in self.items[self.names.index(key[0])]["count"]]

    def _launches(self):
        Q, g = self.Q, self.geom
        out = []
        for it in self.items:
            pid, o = it["pid"], it["out"]
            out += [
                ((it["name"], "contains"), lambda it=it: Q.contains(g, it["pid"], it["p"], out=it["out"]["contains"])),
                ((it["name"], "inside"), lambda it=it: Q.inside(g, it["pid"], it["p"], out=it["out"]["inside"])),
                ((it["name"], "dist_in"), lambda it=it: Q.distance_to_in(g, it["pid"], it["pout"], it["dout"],
                                                                         it["smout"], out=it["out"]["dist_in"])),
                ((it["name"], "safe_in"), lambda it=it: Q.safety_to_in(g, it["pid"], it["pout"],
                                                                       out=it["out"]["safe_in"])),
                ((it["name"], "dist_out"), lambda it=it: Q.distance_to_out(g, it["pid"], it["pin"], it["din"],
                                                                           it["smin"], out=it["out"]["dist_out"])),
                ((it["name"], "safe_out"), lambda it=it: Q.safety_to_out(g, it["pid"], it["pin"],
                                                                         out=it["out"]["safe_out"])),
            ]
        return out


class MixedWorkload:
    """The mixed-precision workload of the zoo: the same shapes and
    placements, 2^22 float32 points per shape of which a quarter lie on the
    surface, jittered by N(0, (1e-7 mm)^2) (interior points moved along
    isotropic directions by the FP64 DistanceToOut, as
    workloads.surface_sampler describes), the rest uniform in 1.5x the
    extent, with isotropic float32 directions.  The step is the FP64 step's
    set of queries; inside / outside routing by the FP64 Contains of the
    float points promoted to double."""

    def __init__(self, wl, device):
        import torch

        from vecgeom_amd import workloads as W

        Q = wl.Q
        self.wl = wl
        self.items = []
        t64 = lambda a: tuple(torch.from_numpy(np.ascontiguousarray(a[i])).to(device) for i in range(3))
        for k, ((name, u), xf, pid) in enumerate(zip(wl.shapes, wl.xfs, wl.ids)):
            ext = wl.geom.placed_extent(pid)

            def contains(lp, xf=xf, pid=pid):
                m = W.local_to_master(xf, lp)
                return Q.contains(wl.geom, pid, t64(m)).bool().cpu().numpy()

            def dist_out(lp, ld, pid=pid):
                return Q.distance_to_out(wl.geom, pid, t64(lp), t64(ld)).cpu().numpy()

            inp = W.mixed_inputs(ext, xf, wl.n, W.SEED_BASE + 3 + 17 * k, W.surface_sampler(ext, contains, dist_out))
            p32, d32 = inp["p32"], inp["d32"]
            lp32, ld32 = inp["lp"].astype(np.float32), inp["ld"].astype(np.float32)
            f = lambda a: tuple(torch.from_numpy(np.ascontiguousarray(a[i])).to(device) for i in range(3))
            P, D, LP, LD = f(p32), f(d32), f(lp32), f(ld32)
            c = Q.contains(wl.geom, pid, tuple(a.double() for a in P)).bool()
            iin, iout = torch.nonzero(c).flatten(), torch.nonzero(~c).flatten()
            pick = lambda s_, idx: tuple(a[idx].contiguous() for a in s_)
            it = {"name": name, "pid": pid, "p": P, "pout": pick(P, iout), "dout": pick(D, iout),
                  "pin": pick(LP, iin), "din": pick(LD, iin),
                  "smout": torch.full((iout.numel(),), 1e30, dtype=torch.float64, device=device),
                  "smin": torch.full((iin.numel(),), 1e30, dtype=torch.float64, device=device),
                  "host": {"p": inp["p"], "d": inp["d"], "lp": lp32.astype(np.float64), "ld": ld32.astype(np.float64)},
                  "inside_idx": iin.cpu().numpy(), "outside_idx": iout.cpu().numpy()}
            n, n_in, n_out = wl.n, iin.numel(), iout.numel()
            it["count"] = {"contains": n, "inside": n, "dist_in": n_out, "safe_in": n_out, "dist_out": n_in,
                           "safe_out": n_in}
            for q in list(it["count"]):
                if (name, q) in EXCLUDED:
                    del it["count"][q]
            self.items.append(it)
        self.queries_per_step = sum(sum(it["count"].values()) for it in self.items)
        self.bytes_per_step = sum(c * BYTES_F32[q] for it in self.items for q, c in it["count"].items())
        self.bytes_per_step_f64 = sum(c * BYTES[q] for it in self.items for q, c in it["count"].items())

    def tasks(self, form, counts=None):
        """The step as a task table: form "mixed" (float32 inputs, FP32
        filter), "fp32" (float32 inputs promoted), or "fp64" (the float32
        values promoted to float64 arrays).  Fresh output tensors; counts:
        {(shape, query): one-element int64 tensor} for FP64-lane counts."""
        import torch

        from vecgeom_amd import batch as B

        cv = (lambda s_: tuple(a.double() for a in s_)) if form == "fp64" else (lambda s_: s_)
        mixed = form == "mixed"
        out = {}
        for it in self.items:
            pid = it["pid"]
            p, pout, dout, pin, din = cv(it["p"]), cv(it["pout"]), cv(it["dout"]), cv(it["pin"]), cv(it["din"])
            c = (lambda q: counts.get((it["name"], q))) if counts is not None else (lambda q: None)
            mk = {"contains": lambda: B.Task(pid, B.CONTAINS, p, mixed=mixed, fp64_count=c("contains")),
                  "inside": lambda: B.Task(pid, B.INSIDE, p, mixed=mixed, fp64_count=c("inside")),
                  "dist_in": lambda: B.Task(pid, B.DISTANCE_TO_IN, pout, dout, it["smout"], mixed=mixed,
                                            fp64_count=c("dist_in")),
                  "safe_in": lambda: B.Task(pid, B.SAFETY_TO_IN, pout, mixed=mixed, fp64_count=c("safe_in")),
                  "dist_out": lambda: B.Task(pid, B.DISTANCE_TO_OUT, pin, din, it["smin"], mixed=mixed,
                                             fp64_count=c("dist_out")),
                  "safe_out": lambda: B.Task(pid, B.SAFETY_TO_OUT, pin, mixed=mixed, fp64_count=c("safe_out"))}
            for q in it["count"]:
                out[(it["name"], q)] = mk[q]()
        return out


def time_steps(launches, steps, stream, watch=None):
    """Runs `steps` steps and returns (elapsed seconds, {launch index: [ms]}).

    HIP events are recorded on the launch stream around the launches whose
    indices are in `watch` (None = every launch); keeping them off the other
    launches keeps the event records out of the throughput measurement."""
    import torch

    idx = range(len(launches)) if watch is None else list(watch)
    ev = {k: [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
          for k in idx}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        for k, (key, fn) in enumerate(launches):
            e = ev.get(k)
            if e is not None:
                e[s][0].record(stream)
            fn()
            if e is not None:
                e[s][1].record(stream)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    return elapsed, {k: [a.elapsed_time(b) for a, b in v] for k, v in ev.items()}


def boolean_bench(args, dev, stream, with_cpu):
    """Boolean volumes (union, intersection, subtraction; see
    workloads.boolean_shapes): each query on 2^20 points uniform in a 120 mm
    cube around the volume, random placement, per-launch kernels timed with
    HIP events (median of five launches); Contains / Inside / SafetyToIn /
    DistanceToIn on every point, DistanceToOut / SafetyToOut on the points
    inside.  With the CPU leg: the reference's scalar virtuals on one host
    thread over a strided sample of the same queries."""
    import torch

    from vecgeom_amd import geometry as G
    from vecgeom_amd import queries as Q
    from vecgeom_amd import workloads as W

    n = 1 << 20
    shapes = W.boolean_shapes()
    rng = np.random.Generator(np.random.Philox(W.SEED_BASE + 77))
    xfs = [W.random_placement(rng) for _ in shapes]
    geom, ids, _ = G.single_shape_geometry([u for _, u in shapes], xfs)
    t = lambda a: tuple(torch.from_numpy(np.ascontiguousarray(a[i])).to(dev) for i in range(3))
    rows = []
    ref = None
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po
        ref = po.reference(fast=True)
    for (name, u), xf, pid in zip(shapes, xfs, ids):
        inp = W.shape_inputs(((-60, -60, -60), (60, 60, 60)), xf, n, seed=W.SEED_BASE + 700 + pid)
        p, d = t(inp["p"]), t(inp["d"])
        c = Q.contains(geom, pid, p)
        torch.cuda.synchronize()
        inside = np.flatnonzero(c.cpu().numpy() == 1)
        lp, ld = t(inp["lp"][:, inside]), t(inp["ld"][:, inside])
        forms = [("contains", lambda: Q.contains(geom, pid, p, out=c), n),
                 ("inside", lambda: Q.inside(geom, pid, p), n),
                 ("dist_in", lambda: Q.distance_to_in(geom, pid, p, d), n),
                 ("safe_in", lambda: Q.safety_to_in(geom, pid, p), n),
                 ("dist_out", lambda: Q.distance_to_out(geom, pid, lp, ld), len(inside)),
                 ("safe_out", lambda: Q.safety_to_out(geom, pid, lp), len(inside))]
        for qn, fn, m in forms:
            if m == 0:
                continue
            for _ in range(2):
                fn()
            _, ms = time_steps([((name, qn), fn)], 5, stream)
            us = 1e3 * float(np.median(ms[0]))  # the first run holds the host's start-up gap
            rows.append({"shape": name, "query": qn, "n": m, "avg_us": round(us, 3),
                         "gqps": round(m / us / 1e3, 3)})
        if ref is not None:
            h = ref.placed(u.type, u.params, xf.array())
            sel = np.arange(0, n, 64)
            t0 = time.perf_counter()
            h.query(po.Q_CONTAINS, inp["p"][:, sel])
            h.query(po.Q_DIST_IN, inp["p"][:, sel], inp["d"][:, sel], np.full(len(sel), 1e30))
            h.query(po.Q_SAFE_IN, inp["p"][:, sel])
            dt = time.perf_counter() - t0
            rows[-1]["cpu_ref_1thread_gqps_contains_distin_safein"] = round(3 * len(sel) / dt / 1e9, 5)
    geom.free()
    tot_q = sum(r["n"] for r in rows)
    tot_us = sum(r["avg_us"] for r in rows)
    return {"workload": "Boolean volumes (union, intersection, subtraction, hole, Polyhedron intersection), "
                        "random placements, 2^20 points per volume, six queries, one launch per (volume, query)",
            "value": round(tot_q / tot_us / 1e3, 3), "unit": "Gqueries/s", "per_kernel": rows}


def normal_bench(args, dev, stream, with_cpu):
    """Surface normals (vg_normal, vg_distance_to_out_normal) on the zoo shapes
    whose reference has a Normal: per shape, 2^22 interior points (uniform in
    the extent, kept where the device Contains holds) with isotropic
    directions; DistanceToOut with the exit normal on those rays, then Normal
    on the exit points (surface points).  Per-launch kernels timed with HIP
    events, median of five launches; algorithmic bytes: Normal 49 B per
    query (24 in, 3 doubles and a flag out), DistanceToOut with normal 82 B
    (48 in, a distance, 3 doubles and two flags out).  With the CPU leg: the
    reference's own Normal and DistanceToOut-then-Normal (its Release build,
    oracle/_ref) on one host thread over every 64th point and ray."""
    import torch

    from vecgeom_amd import geometry as G
    from vecgeom_amd import queries as Q
    from vecgeom_amd import workloads as W

    n = 1 << 22
    shapes = [(nm, u) for nm, u in W.zoo_shapes() if u.type in (G.VG_BOX, G.VG_TUBE, G.VG_ORB, G.VG_SPHERE, G.VG_CONE,
                                                                  G.VG_TRD, G.VG_TRAPEZOID, G.VG_POLYCONE)]
    I = G.Transformation3D()
    geom, ids, _ = G.single_shape_geometry([u for _, u in shapes], [I] * len(shapes))
    rows = []
    ref = None
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as po
        ref = po.reference(fast=True)
    for (name, u), pid in zip(shapes, ids):
        rng = np.random.Generator(np.random.Philox(W.SEED_BASE + 900 + pid))
        lo, hi = geom.placed_extent(pid)  # identity placement: the local extent
        kept = []
        while sum(k.shape[1] for k in kept) < n:
            cand = rng.uniform(np.asarray(lo), np.asarray(hi), size=(n, 3)).T.copy()
            ct = tuple(torch.from_numpy(cand[i]).to(dev) for i in range(3))
            inside = Q.contains(geom, pid, ct).cpu().numpy().astype(bool)
            kept.append(cand[:, inside])
        p = np.concatenate(kept, axis=1)[:, :n]
        d = W.isotropic(n, rng)
        pt = tuple(torch.from_numpy(np.ascontiguousarray(p[i])).to(dev) for i in range(3))
        dt = tuple(torch.from_numpy(np.ascontiguousarray(d[i])).to(dev) for i in range(3))
        dist, _, _, _ = Q.distance_to_out_normal(geom, pid, pt, dt)
        hit = tuple((pt[i] + dt[i] * dist).contiguous() for i in range(3))
        forms = [("normal", lambda: Q.normal(geom, pid, hit), 49),
                 ("dist_out_normal", lambda: Q.distance_to_out_normal(geom, pid, pt, dt), 82)]
        for qn, fn, bpq in forms:
            for _ in range(2):
                fn()
            _, ms = time_steps([((name, qn), fn)], 5, stream)
            us = 1e3 * float(np.median(ms[0]))  # the first run holds the host's start-up gap
            rows.append({"shape": name, "query": qn, "n": n, "avg_us": round(us, 3),
                         "gqps": round(n / us / 1e3, 3), "gbs": round(n * bpq / us / 1e3, 1)})
        if ref is not None:
            h = ref.placed(u.type, u.params, I.array())
            sel = np.arange(0, n, 64)
            hp = np.stack([a.cpu().numpy()[sel] for a in hit])
            t0 = time.perf_counter()
            h.normal(hp)
            rows[-2]["cpu_ref_1thread_gqps"] = round(len(sel) / (time.perf_counter() - t0) / 1e9, 5)
            t0 = time.perf_counter()
            h.distance_to_out_normal(p[:, sel], d[:, sel])
            rows[-1]["cpu_ref_1thread_gqps"] = round(len(sel) / (time.perf_counter() - t0) / 1e9, 5)
    geom.free()
    tot_q = sum(r["n"] for r in rows)
    tot_us = sum(r["avg_us"] for r in rows)
    return {"workload": "Normal on 2^22 exit points and DistanceToOut with the exit normal on 2^22 interior rays "
                        "per shape (Box, Orb, Sphere, Tube, Cone, Trd, Trapezoid, Polycone), one launch per "
                        "(shape, query)",
            "value": round(tot_q / tot_us / 1e3, 3), "unit": "Gqueries/s", "per_kernel": rows}


def host_cpu():
    """Host CPU description from lscpu (logical CPUs, cores, model)."""
    import subprocess
    info = {"logical_cpus": os.cpu_count()}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        kv = {}
        for line in txt.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
        info["model"] = kv.get("Model name")
        if "CPU(s)" in kv:
            info["logical_cpus"] = int(kv["CPU(s)"])
        if "Core(s) per socket" in kv and "Socket(s)" in kv:
            info["physical_cores"] = int(kv["Core(s) per socket"]) * int(kv["Socket(s)"])
        info["threads_per_core"] = int(kv.get("Thread(s) per core", "1"))
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass
    return info


def cpu_baseline(wl, target_s):
    """The reference's CPU basket path on the host: its SOA3D basket
    virtuals, built from the reference's sources with its Release flags
    (oracle/_ref/libvgref_fast.so), over contiguous ranges of each basket;
    the oracle restatement stands in where that build is absent.  The same
    queries on a bounded, evenly strided sample of the same inputs, timed on
    1 thread and on every CPU this process may run on: one warm-up pass, then
    the median of 5 passes (steady clock)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    ref = po.reference(fast=True)
    o = ref if ref is not None and ref.has("basket") else po.oracle()
    kind = "reference" if o is ref else "port"
    host = host_cpu()
    nall = max(1, host.get("affinity_cpus") or host.get("logical_cpus") or 1)
    handles = [o.placed(u.type, u.params, xf.array()) for (name, u), xf in zip(wl.shapes, wl.xfs)]

    def one_pass(stride, nthreads):
        total = 0
        t0 = time.perf_counter()
        for it, h in zip(wl.items, handles):
            run = h.basket if kind == "reference" else h.query
            hp = it["host"]
            sel_all = slice(0, wl.n, stride)
            iout = it["outside_idx"][::stride]
            iin = it["inside_idx"][::stride]
            P, D, LP, LD = hp["p"], hp["d"], hp["lp"], hp["ld"]
            run(po.Q_CONTAINS, P[:, sel_all], nthreads=nthreads)
            run(po.Q_INSIDE, P[:, sel_all], nthreads=nthreads)
            if (it["name"], "dist_in") not in EXCLUDED:
                run(po.Q_DIST_IN, P[:, iout], D[:, iout], np.full(len(iout), 1e30), nthreads=nthreads)
                total += len(iout)
            run(po.Q_SAFE_IN, P[:, iout], nthreads=nthreads)
            run(po.Q_DIST_OUT, LP[:, iin], LD[:, iin], np.full(len(iin), 1e30), nthreads=nthreads)
            run(po.Q_SAFE_OUT, LP[:, iin], nthreads=nthreads)
            total += 2 * len(range(0, wl.n, stride)) + len(iout) + 2 * len(iin)
        return total, time.perf_counter() - t0

    # Size the sample so that one single-thread pass takes about target_s / 7
    # (6 single-thread passes and 6 shorter multi-thread ones fit target_s).
    stride = 1024
    want = target_s / 7
    while True:
        _, dt = one_pass(stride, 1)
        if dt >= 0.6 * want or stride == 1:
            break
        nxt = int(stride * dt / want) if dt > 0.05 else stride // 16
        stride = max(1, min(stride - 1, nxt))
    res = {}
    for nt in (1, nall):
        one_pass(stride, nt)  # warm-up
        runs = [one_pass(stride, nt) for _ in range(5)]
        q = runs[0][0]
        med = float(np.median([dt for _, dt in runs]))
        res[nt] = (q, med)
    q, med_all = res[nall]
    _, med_1 = res[1]
    what = ("reference SOA3D basket virtuals, Scalar backend, Release flags (-O3 -ffast-math)" if kind == "reference"
            else "oracle C++ restatement of the reference Scalar backend")
    return {
        "value": round(q / med_all / 1e9, 5),
        "unit": "Gqueries/s",
        "cores": nall,
        "kind": kind,
        "single_thread": {"value": round(q / med_1 / 1e9, 6), "unit": "Gqueries/s", "cores": 1,
                          "median_s": round(med_1, 4)},
        "median_s": round(med_all, 4),
        "host_cpu": host,
        "sample": f"every {stride}th query of the same workload ({q} queries per pass), {what}, std::thread over "
                  f"contiguous ranges; one warm-up pass, median of 5 passes, on 1 thread and on all {nall} CPUs",
    }


SAMPLE_STRIDE = 61


def capture(items, outputs, stride=SAMPLE_STRIDE):
    """Every stride-th output of each (shape, query) of a form, as numpy:
    classifications over all points, distances and safeties over their
    basket.  outputs: {(shape, query): tensor}."""
    return {k: v.cpu().numpy()[::stride].copy() for k, v in outputs.items()}


def sample_check(items, shapes, xfs, n, captured, promote, stride=SAMPLE_STRIDE):
    """Part of the CPU-baseline leg: compares sampled device outputs
    (capture) with the oracle evaluated on the inputs the device saw (float32
    values promoted to double when promote).  Returns per shape:
    classification mismatches, distance/safety lanes beyond max(1e-9 mm,
    1e-12 relative) -- for the Torus also how many of those lie within the
    surface tolerance of the exact first crossing -- the largest |difference|
    and the lanes compared."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    from vecgeom_amd import geometry as G
    from vecgeom_amd import workloads as W

    o = po.oracle()
    cv = (lambda a: a.astype(np.float32).astype(np.float64)) if promote else (lambda a: a)
    qcode = {"contains": po.Q_CONTAINS, "inside": po.Q_INSIDE, "dist_in": po.Q_DIST_IN, "safe_in": po.Q_SAFE_IN,
             "dist_out": po.Q_DIST_OUT, "safe_out": po.Q_SAFE_OUT}
    out = {}
    for it, (name, u), xf in zip(items, shapes, xfs):
        h = o.placed(u.type, u.params, xf.array())
        hp = it["host"]
        r = {"class_mismatch": 0, "beyond_bar": 0, "max_abs_err": 0.0, "lanes": 0}
        for q in it["count"]:
            got = captured[(name, q)]
            if q in ("contains", "inside"):
                idx = np.arange(0, n, stride)
                want = h.query(qcode[q], cv(hp["p"][:, idx]))
                r["class_mismatch"] += int((got.astype(np.int64) != want.astype(np.int64)).sum())
            else:
                src_idx = it["outside_idx"] if q in ("dist_in", "safe_in") else it["inside_idx"]
                sel = src_idx[::stride]
                P = cv(hp["p"][:, sel]) if q in ("dist_in", "safe_in") else cv(hp["lp"][:, sel])
                if q in ("dist_in", "dist_out"):
                    D = cv(hp["d"][:, sel]) if q == "dist_in" else cv(hp["ld"][:, sel])
                    want = h.query(qcode[q], P, D, np.full(len(sel), 1e30))
                else:
                    want = h.query(qcode[q], P)
                eq = (got == want) | (np.isnan(got) & np.isnan(want))
                with np.errstate(invalid="ignore"):
                    diff = np.where(eq, 0.0, np.abs(got - want))
                    ok = eq | (diff <= np.maximum(1e-9, 1e-12 * np.abs(want)))
                r["beyond_bar"] += int((~ok).sum())
                if u.type == G.VG_TORUS and q in ("dist_in", "dist_out") and not ok.all():
                    # The device's Torus roots are exact where the reference's
                    # closed form is not: lanes beyond the bar are checked
                    # against the exact first crossing.
                    bad = np.flatnonzero(~ok)
                    lp, ld = (W.master_to_local(xf, P), W.master_to_local(xf, D, direction=True)) \
                        if q == "dist_in" else (P, D)
                    r["beyond_bar_exact_ok"] = r.get("beyond_bar_exact_ok", 0) + int(
                        po.torus_distance_ok(u.params, lp[:, bad], ld[:, bad], got[bad], q == "dist_out").sum())
                fin = np.isfinite(diff)
                if fin.any():
                    r["max_abs_err"] = max(r["max_abs_err"], float(diff[fin].max()))
                if (~fin & ~eq).any():
                    r["max_abs_err"] = float("inf")
            r["lanes"] += got.size
        out[name] = r
    return out


def batch_forms(args, wl, stream):
    """The step as one batch (one launch per shape weight class), autotuned;
    the same batch replayed from a captured hipGraph; and the FP32-storage
    form (float32 inputs promoted to double in registers)."""
    import torch

    from vecgeom_amd import batch as B

    res = {}
    bt = B.Batch(wl.geom, wl.tasks())
    if args.batch_launch:
        it_, blocks_ = (int(v) for v in args.batch_launch.split(","))
        bt.set_launch(it_, blocks_)
        best_ms = None
    else:
        best_ms = bt.autotune(3)
    it, blocks, items = bt.launch()
    # Timed region: HIP events around each class kernel on the stream it runs
    # on (vg_batch_set_timing); a handful of events per step.
    bt.set_timing(True)
    el, _ = time_steps([(("batch", "all"), bt.run)], args.steps, stream, watch=[])
    classes = [c for c in bt.class_stats() if c["runs"] > 0]
    bt.set_timing(False)
    res = {"value": round(wl.queries_per_step * args.steps / el / 1e9, 3), "ms_per_step": round(1e3 * el / args.steps, 4),
           "items_per_lane": it, "blocks": blocks, "autotune_ms": None if best_ms is None else round(best_ms, 4),
           "achieved_gbs": round(wl.bytes_per_step / (el / args.steps) / 1e9, 1)}
    for c in classes:
        c["achieved_gbs"] = round(c["bytes"] / (c["mean_ms"] * 1e-3) / 1e9, 1) if c["mean_ms"] > 0 else None
        c["mean_us"] = round(1e3 * c.pop("mean_ms"), 3)
    res["classes_concurrent"] = classes
    # The class kernels of the step run concurrently, so the events around
    # each measure its span inside the step, not its own speed.  Each class
    # is therefore also run alone (a batch of its tasks only, the same
    # launch shape), timed with HIP events on its launch stream over
    # args.steps runs; the class that takes longest alone is the dominant
    # kernel of the step.
    alone = []
    for c in classes:
        sub = [t for t in wl.tasks() if shape_class(wl.names_of_pid[t.placed], t.query) == c["class"]]
        if not sub:
            continue
        cb = B.Batch(wl.geom, sub)
        cb.set_launch(it, blocks)
        cb.set_timing(True)
        for _ in range(max(1, args.warmup)):
            cb.run()
        cb.set_timing(True)
        time_steps([(("class", c["class"]), cb.run)], args.steps, stream, watch=[])
        st = cb.class_stats()[c["class"]]
        cb.free()
        ms = st["mean_ms"]
        alone.append({"class": c["class"], "name": c["name"], "items": st["items"], "bytes": st["bytes"],
                      "runs": st["runs"], "mean_us": round(1e3 * ms, 3),
                      "achieved_gbs": round(st["bytes"] / (ms * 1e-3) / 1e9, 1) if ms > 0 else None})
    res["classes_alone"] = alone
    dom = max(alone, key=lambda c: c["mean_us"])
    res["roofline"] = {"bound": "hbm", "achieved": dom["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(dom["achieved_gbs"] / HBM_PEAK_GBS, 4),
                       "traffic": profiled_class_traffic(dom["class"], dom["items"]),
                       "kernel": f"BatchKernel<{ClassIT(dom['class'], it)},{dom['class']}> ({dom['name']})",
                       "avg_launch_us": dom["mean_us"], "queries_per_launch": dom["items"],
                       "bytes_per_launch": dom["bytes"], "timed": "alone (its class's tasks only), HIP events on its "
                                                                  "launch stream, mean of the timed runs",
                       "bytes_rule": "each input element read once, each output element written once (FP64: "
                                     "Contains 25 B, Inside 28, distances 64, safeties 32 per query)"}
    res["step_roofline"] = {"bound": "hbm", "achieved": res["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(res["achieved_gbs"] / HBM_PEAK_GBS, 4),
                            "what": "the whole batched step (all class kernels, concurrent): its algorithmic bytes "
                                    "over its time per step"}
    res["_captured"] = capture(wl.items, {(it["name"], q): it["out"][q] for it in wl.items for q in it["count"]})
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        bt.run()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        bt.run()
    torch.cuda.synchronize()
    for _ in range(max(1, args.warmup)):
        graph.replay()
    gel, _ = time_steps([(("graph", "all"), graph.replay)], args.steps, stream, watch=[])
    res["graph_value"] = round(wl.queries_per_step * args.steps / gel / 1e9, 3)
    res["graph_ms_per_step"] = round(1e3 * gel / args.steps, 4)
    del graph
    bt.free()
    if args.no_fp32:
        return res, None
    f32 = B.Batch(wl.geom, wl.tasks(fp32=True))
    f32.autotune(3)
    it32, blocks32, _ = f32.launch()
    el32, _ = time_steps([(("batch32", "all"), f32.run)], args.steps, stream, watch=[])
    fp32 = {"value": round(wl.queries_per_step * args.steps / el32 / 1e9, 3), "unit": "Gqueries/s",
            "ms_per_step": round(1e3 * el32 / args.steps, 4), "items_per_lane": it32, "blocks": blocks32,
            "achieved_gbs": round(wl.bytes_per_step_f32 / (el32 / args.steps) / 1e9, 1),
            "note": "float32 inputs promoted to double before any arithmetic; check_vs_oracle compares a sample "
                    "of the outputs with the oracle evaluated on the same promoted inputs"}
    fp32["_captured"] = capture(wl.items, {(it["name"], q): it["out"][q] for it in wl.items for q in it["count"]})
    f32.free()
    return res, fp32


def mixed_forms(args, mwl, stream):
    """The mixed-precision workload in four forms, each autotuned and timed
    over the same number of steps: the mixed form (FP32 filter, FP64 on the
    undecided lanes), FP32 storage (FP64 arithmetic on the promoted float
    inputs), the FP64 batched form and the plain FP64 form on float64 copies
    of the same values.  Per shape: mismatches and max / RMS |difference| of
    the mixed outputs against the FP64 batched outputs (every lane), and the
    fraction of lanes the mixed form evaluated in FP64 (one counted run)."""
    import torch

    from vecgeom_amd import batch as B
    from vecgeom_amd import capi

    res = {"workload": "shape zoo, random placements, 2^%d float32 points per shape, %d%% on the surface jittered "
                       "by N(0, (1e-7 mm)^2), the rest uniform in 1.5x the extent" % (
                           int(np.log2(mwl.wl.n)), 25),
           "queries_per_step": mwl.queries_per_step}
    outs = {}
    for form in ("mixed", "mixed_all", "fp32", "fp64"):
        # mixed: the FP32 filter where it was measured to pay; mixed_all: on
        # every Contains, Inside and DistanceToIn task.
        B.set_mixed_policy(B.POLICY_ALL if form == "mixed_all" else B.POLICY_MEASURED)
        tasks = mwl.tasks("mixed" if form.startswith("mixed") else form)
        bt = B.Batch(mwl.wl.geom, list(tasks.values()))
        bt.autotune(3)
        it, blocks, _ = bt.launch()
        for _ in range(max(1, args.warmup)):
            bt.run()
        el, _ = time_steps([(("batch", form), bt.run)], args.steps, stream, watch=[])
        nbytes = mwl.bytes_per_step if form != "fp64" else mwl.bytes_per_step_f64
        res[form] = {"value": round(mwl.queries_per_step * args.steps / el / 1e9, 3), "unit": "Gqueries/s",
                     "ms_per_step": round(1e3 * el / args.steps, 4), "items_per_lane": it, "blocks": blocks,
                     "achieved_gbs": round(nbytes / (el / args.steps) / 1e9, 1)}
        torch.cuda.synchronize()
        outs[form] = {k: t.out for k, t in tasks.items()}
        bt.free()
    B.set_mixed_policy(B.POLICY_MEASURED)
    # The plain FP64 form on the float64 copies: one launch per (shape, query).
    Q, g = mwl.wl.Q, mwl.wl.geom
    plain = []
    for it in mwl.items:
        d64 = lambda s_: tuple(a.double() for a in s_)
        P, PO, DO, PI, DI = d64(it["p"]), d64(it["pout"]), d64(it["dout"]), d64(it["pin"]), d64(it["din"])
        fns = {"contains": lambda P=P, pid=it["pid"]: Q.contains(g, pid, P),
               "inside": lambda P=P, pid=it["pid"]: Q.inside(g, pid, P),
               "dist_in": lambda PO=PO, DO=DO, it=it: Q.distance_to_in(g, it["pid"], PO, DO, it["smout"]),
               "safe_in": lambda PO=PO, it=it: Q.safety_to_in(g, it["pid"], PO),
               "dist_out": lambda PI=PI, DI=DI, it=it: Q.distance_to_out(g, it["pid"], PI, DI, it["smin"]),
               "safe_out": lambda PI=PI, it=it: Q.safety_to_out(g, it["pid"], PI)}
        plain += [((it["name"], q), fns[q]) for q in it["count"]]
    capi.call("vg_set_kernel_variant", 1)
    for _ in range(max(1, args.warmup)):
        for _, fn in plain:
            fn()
    pel, _ = time_steps(plain, args.steps, stream, watch=[])
    capi.call("vg_set_kernel_variant", 0)
    res["plain_fp64_gpu"] = {"value": round(mwl.queries_per_step * args.steps / pel / 1e9, 3), "unit": "Gqueries/s",
                             "ms_per_step": round(1e3 * pel / args.steps, 4)}
    res["speedup_vs_plain_fp64"] = round(res["mixed"]["value"] / res["plain_fp64_gpu"]["value"], 3)
    res["speedup_vs_fp64_batched"] = round(res["mixed"]["value"] / res["fp64"]["value"], 3)
    res["policy"] = ("mixed: FP32 filter on the (shape, query) pairs where it was measured to pay (Polycone "
                     "DistanceToIn), FP32 storage elsewhere; mixed_all: FP32 filter on every Contains, Inside and "
                     "DistanceToIn task.  On gfx950 an FP64 FMA issues at the unpacked FP32 rate.")
    # Errors of the mixed forms against the FP64 form on the same values
    # (every lane).
    for form in ("mixed", "mixed_all"):
        err = {}
        for it in mwl.items:
            r = {"class_mismatch": 0, "max_abs_err": 0.0, "rms_err": 0.0, "lanes": 0}
            sq, nd = 0.0, 0
            for q in it["count"]:
                a = outs[form][(it["name"], q)].cpu().numpy()
                b = outs["fp64"][(it["name"], q)].cpu().numpy()
                r["lanes"] += a.size
                if q in ("contains", "inside"):
                    r["class_mismatch"] += int((a != b).sum())
                    continue
                eq = (a == b) | (np.isnan(a) & np.isnan(b))
                with np.errstate(invalid="ignore"):
                    dif = np.where(eq, 0.0, np.abs(a - b))
                if not np.isfinite(dif).all():
                    r["max_abs_err"] = float("inf")
                else:
                    r["max_abs_err"] = max(r["max_abs_err"], float(dif.max(initial=0.0)))
                    sq += float((dif * dif).sum())
                    nd += dif.size
            r["rms_err"] = (sq / nd) ** 0.5 if nd else 0.0
            err[it["name"]] = r
        res[form]["error_vs_fp64"] = err
    for form, policy in (("mixed", B.POLICY_MEASURED), ("mixed_all", B.POLICY_ALL)):
        B.set_mixed_policy(policy)
        counts = {(it["name"], q): torch.zeros(1, dtype=torch.int64, device=mwl.items[0]["p"][0].device)
                  for it in mwl.items for q in it["count"]}
        tasks = mwl.tasks("mixed", counts)
        bt = B.Batch(mwl.wl.geom, list(tasks.values()))
        bt.run()
        torch.cuda.synchronize()
        bt.free()
        frac = {}
        tot64 = 0
        for it in mwl.items:
            frac[it["name"]] = {q: round(int(counts[(it["name"], q)].item()) / max(1, c), 4)
                                for q, c in it["count"].items()}
            tot64 += sum(int(counts[(it["name"], q)].item()) for q in it["count"])
        res[form]["fp64_lane_fraction"] = frac
        res[form]["fp64_lane_fraction_all"] = round(tot64 / mwl.queries_per_step, 4)
    B.set_mixed_policy(B.POLICY_MEASURED)
    res["_captured"] = capture(mwl.items, outs["mixed"])
    return res


def nav_bench(args, dev, stream, with_cpu):
    """Navigation on the nested geometry: 2^20 tracks from N(0, 5 mm) vertices
    with isotropic directions, ABBoxNavigator steps with pstep = kInfinity
    until every track has left the world (the TraceTrack loop).  The fused
    form runs the whole loop on the device (one launch); the plain form
    launches one FindNextBoundaryAndStep per step over the whole basket from
    the host, as the reference's GPU navigation benchmark does, and updates
    positions between launches."""
    import torch

    from vecgeom_amd import geometry as G
    from vecgeom_amd import nav as N
    from vecgeom_amd import workloads as W

    _, flat = W.nested_geometry()
    geom = G.DeviceGeometry(flat)
    n = args.nav_tracks
    p0, d0 = W.track_starts(n)
    t = lambda a: tuple(torch.from_numpy(np.ascontiguousarray(a[i])).to(dev) for i in range(3))
    p, d = t(p0), t(d0)
    out = N.TraceResult(n, dev)
    N.trace_tracks(geom, p, d, navigator=N.ABBOX, out=out)
    torch.cuda.synchronize()
    total_steps = int(out.nsteps.to(torch.int64).sum().item())
    exact = int(out.exact.to(torch.int64).sum().item())
    ms = []
    for _ in range(args.nav_steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        N.trace_tracks(geom, p, d, navigator=N.ABBOX, out=out)
        e1.record(stream)
        ms.append((e0, e1))
    torch.cuda.synchronize()
    ms = [a.elapsed_time(b) for a, b in ms]
    fused_ms = float(np.mean(ms))
    res = {"workload": f"nested geometry (64 tube shells x 32 boxes x 8 Trds), {n} tracks from N(0,5mm) vertices, "
                       "isotropic, ABBoxNavigator, pstep=inf, until exit",
           "tracks": n, "steps": total_steps, "mean_steps_per_track": round(total_steps / n, 3),
           "exact_walk_steps": exact, "kernel": "TraceKernel (fused TraceTrack loop)",
           "avg_launch_ms": round(fused_ms, 3), "value": round(total_steps / (fused_ms * 1e-3) / 1e9, 4),
           "unit": "G steps/s", "dtype": "f64"}
    # The mixed-precision navigator (FP32 filters in the daughter walk and
    # the locate scans), same tracks; identical results required.
    ref_ns, ref_hash = out.nsteps.clone(), out.hash.clone()
    ref_fin = [a.clone() for a in out.final]
    mout = N.TraceResult(n, dev)
    N.trace_tracks(geom, p, d, navigator=N.ABBOX_MIXED, out=mout)
    torch.cuda.synchronize()
    ev = []
    for _ in range(args.nav_steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        N.trace_tracks(geom, p, d, navigator=N.ABBOX_MIXED, out=mout)
        e1.record(stream)
        ev.append((e0, e1))
    torch.cuda.synchronize()
    mixed_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    same = bool(torch.equal(mout.nsteps, ref_ns) and torch.equal(mout.hash, ref_hash)
                and all(torch.equal(a, b) for a, b in zip(mout.final, ref_fin)))
    res["mixed"] = {"kernel": "TraceKernel<mixed> (FP32 daughter and locate filters)", "avg_launch_ms": round(mixed_ms, 3),
                    "value": round(total_steps / (mixed_ms * 1e-3) / 1e9, 4), "unit": "G steps/s",
                    "identical_to_fp64": same, "speedup_vs_fused_fp64": round(fused_ms / mixed_ms, 3)}
    # Plain form: one launch per step from the host.
    cur = N.locate_points(geom, p)
    new = N.NavStates(geom, n, device=dev)
    step = torch.empty(n, dtype=torch.float64, device=dev)
    q = [a.clone() for a in p]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    launches = 0
    for _ in range(100000):
        act = cur.level > 0
        if not bool(act.any()):
            break
        N.find_next_boundary_and_step(geom, tuple(q), d, cur, new=new, step=step)
        s = step + 1e-6
        for i in range(3):
            q[i] = torch.where(act, q[i] + d[i] * s, q[i])
        cur, new = new, cur
        launches += 1
    torch.cuda.synchronize()
    pel = time.perf_counter() - t0
    res["plain_per_step_launch"] = {"value": round(total_steps / pel / 1e9, 4), "unit": "G steps/s",
                                    "launches": launches, "seconds": round(pel, 4),
                                    "speedup_of_fused": round((total_steps / (fused_ms * 1e-3)) / (total_steps / pel), 3)}
    res["mixed"]["speedup_vs_plain_fp64"] = round((total_steps / (mixed_ms * 1e-3)) / (total_steps / pel), 3)
    if with_cpu:
        res["cpu_baseline"] = nav_cpu_baseline(flat, p0, d0, args.cpu_seconds / 2)
    geom.free()
    return res


def serve_bench(args, dev):
    """Serving on the nested geometry: baskets with sizes log-uniform in
    16..65536 tracks (vertex/direction distribution of the navigation run)
    arrive as a seeded Poisson process, are admitted by footprint into a
    32 GiB track arena, take `serve_steps` navigation steps each (one request
    per step, the next submitted when the previous completes) and close.
    Latency is per request, submission to completion.  The batched forms
    merge all pending requests into one launch on each of two streams, with
    the mixed-precision or the FP64 navigator; the per-basket form launches
    each request alone, one at a time, with the FP64 navigator.  Rates are
    0.5 and 0.9 of the per-basket form's saturation rate and 0.5 and 0.9 of
    the batched form's (saturation: all baskets arriving at once)."""
    from vecgeom_amd import geometry as G
    from vecgeom_amd import serve as S
    from vecgeom_amd import workloads as W

    _, flat = W.nested_geometry()
    geom = G.DeviceGeometry(flat)
    rng = np.random.Generator(np.random.Philox(2024))
    nb, steps = args.serve_baskets, args.serve_steps
    sizes = np.exp(rng.uniform(np.log(16), np.log(65536), nb)).astype(np.uint64)
    p, d = W.track_starts(int(sizes.sum()), seed=4040)
    from vecgeom_amd import nav as N

    servers = {"fp64": S.Server(geom, 32 << 30), "mixed": S.Server(geom, 32 << 30, navigator=N.ABBOX_MIXED)}
    # form: (server, max requests per launch, launch streams)
    forms = {"batched_mixed": ("mixed", 4096, 2), "batched": ("fp64", 4096, 2), "per_basket": ("fp64", 1, 1)}

    def run(form, arrival):
        srv = servers[forms[form][0]]
        srv.set_batching(*forms[form][1:])
        s0 = srv.stats()
        lat, wait, wall = srv.replay(arrival, sizes, steps, p, d)
        s1 = srv.stats()
        ts = s1["track_steps"] - s0["track_steps"]
        req = s1["requests"] - s0["requests"]
        la = lat.ravel() * 1e3
        return {"p50_ms": round(float(np.percentile(la, 50)), 4), "p99_ms": round(float(np.percentile(la, 99)), 4),
                "mean_ms": round(float(la.mean()), 4), "admit_wait_p99_ms": round(float(np.percentile(wait, 99) * 1e3), 4),
                "steps_per_s": round(ts / wall / 1e9, 5), "wall_s": round(wall, 4),
                "requests": req, "launches": s1["launches"] - s0["launches"]}

    zero = np.zeros(nb)
    run("batched", zero)  # warm-ups
    run("batched_mixed", zero)
    sat = {f: run(f, zero) for f in forms}
    base_rate = nb / sat["per_basket"]["wall_s"]
    batched_rate = nb / sat["batched"]["wall_s"]
    rates = {}
    for rate in [f * base_rate for f in (0.5, 0.9)] + [f * batched_rate for f in (0.5, 0.9)]:
        gaps = rng.exponential(1.0 / rate, nb)
        arrival = np.cumsum(gaps) - gaps[0]
        rates[f"{rate:.0f}/s"] = {f: run(f, arrival) for f in forms}
    res = {"workload": f"nested geometry, {nb} baskets, sizes log-uniform 16..65536 tracks "
                       f"({int(sizes.sum())} tracks), {steps} steps per basket, Poisson arrivals, 32 GiB arena",
           "unit_steps": "G steps/s", "saturation": sat,
           "saturation_basket_rate_per_basket_form": round(base_rate, 1),
           "saturation_basket_rate_batched_form": round(batched_rate, 1),
           "fixed_rates": rates, "dtype": "f64",
           "forms": "batched_mixed: requests merged into shared launches, mixed-precision ABBox navigator; batched: "
                    "the same with the FP64 navigator; per_basket: one launch per request, one at a time, FP64"}
    for srv in servers.values():
        srv.destroy()
    geom.free()
    return res


def nav_cpu_baseline(flat, p0, d0, target_s):
    """The reference's own ABBoxNavigator (built from its sources with its
    Release flags) on one core, on a bounded prefix of the same tracks: one
    warm-up pass, then the median of 5.  The navigator keeps a process-wide
    hit list, so it runs on one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    ref = po.reference(fast=True)
    kind, lib, cores = "reference", ref, 1
    if ref is None:
        kind, lib, cores = "port", po.oracle(), 1
    nav = po.Nav(lib, flat)
    m = 256
    want = target_s / 6
    while True:
        t0 = time.perf_counter()
        ns, _, _, _ = nav.trace(p0[:, :m], d0[:, :m], kind=0, nthreads=cores)
        dt = time.perf_counter() - t0
        if dt >= 0.6 * want or m >= p0.shape[1]:
            break
        m = min(p0.shape[1], max(m + 1, int(m * want / max(dt, 1e-3))))
    runs = []
    for _ in range(5):
        t0 = time.perf_counter()
        ns, _, _, _ = nav.trace(p0[:, :m], d0[:, :m], kind=0, nthreads=cores)
        runs.append(time.perf_counter() - t0)
    dt = float(np.median(runs))
    steps = int(ns.sum())
    return {"value": round(steps / dt / 1e9, 6), "unit": "G steps/s", "cores": cores, "kind": kind,
            "median_s": round(dt, 4),
            "sample": f"first {m} tracks of the same workload ({steps} steps per pass), "
                      + ("reference ABBoxNavigator, Release flags (-O3 -ffast-math)" if kind == "reference"
                         else "oracle C++ restatement")
                      + "; one warm-up pass, median of 5 passes on one thread"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU sharding is "
                                   "described in DESIGN.md, not built)"}))
        sys.exit(2)
    import torch

    from vecgeom_amd import queries as Q

    Q.init(0)
    dev = torch.device("cuda:0")
    wl = Workload(args.points, dev)
    launches = wl.launches()
    stream = torch.cuda.current_stream()

    def step():
        for _, fn in launches:
            fn()

    if args.copy_calibration > 0:
        # A plain device-to-device copy of known size, so that HBM counters
        # (FETCH_SIZE / WRITE_SIZE) can be scaled against it.
        nbytes = args.copy_calibration << 20
        src = torch.empty(nbytes // 8, dtype=torch.float64, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    log(f"workload ready: {len(wl.shapes)} shapes x {args.points} points")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    # Census (untimed): HIP events around every launch give the per-kernel
    # breakdown and pick the dominant kernel.  The median over five steps:
    # the first launch of the first step also holds the host's gap after the
    # synchronisation that starts the run.
    _, cen = time_steps(launches, 5, stream)
    keys = [key for key, _ in launches]
    count = {(it["name"], q): c for it in wl.items for q, c in it["count"].items()}
    rows = []
    for k, key in enumerate(keys):
        ms = float(np.median(cen[k]))
        cnt = count[key]
        rows.append({"shape": key[0], "query": key[1], "n": cnt, "avg_us": 1e3 * ms,
                     "gqps": cnt / (ms * 1e-3) / 1e9 if ms > 0 else None,
                     "gbs": cnt * BYTES[key[1]] / (ms * 1e-3) / 1e9 if ms > 0 else None})
    kdom = max(range(len(rows)), key=lambda k: rows[k]["avg_us"])

    # Timed region: K steps; HIP events on the launch stream (torch's current
    # stream, which the C-ABI calls use) around the dominant kernel only.
    elapsed, ev = time_steps(launches, args.steps, stream, watch=[kdom])
    ms_per_step = 1e3 * elapsed / args.steps
    value = wl.queries_per_step * args.steps / elapsed / 1e9
    dom = dict(rows[kdom])
    dom_ms = float(np.mean(ev[kdom]))
    dom_bytes = dom["n"] * BYTES[dom["query"]]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": profiled_traffic(dom["shape"], dom["query"], dom["n"]),
            "kernel": kernel_name(dom['shape'], dom['query']), "avg_launch_us": round(1e3 * dom_ms, 3),
            "bytes_per_query": BYTES[dom["query"]], "queries_per_launch": dom["n"], "bytes_per_launch": dom_bytes}

    log(f"timed shape steps done: {value:.3f} Gqueries/s")
    plain = None
    if not args.no_plain:
        # The plain FP64 GPU form: same queries, one thread per query with
        # parameters chased from global memory (vg_set_kernel_variant(1)).
        from vecgeom_amd import capi
        capi.call("vg_set_kernel_variant", 1)
        for _ in range(max(1, args.warmup)):
            step()
        pel, _ = time_steps(launches, args.steps, stream, watch=[])
        capi.call("vg_set_kernel_variant", 0)
        pval = wl.queries_per_step * args.steps / pel / 1e9
        plain = {"value": round(pval, 3), "unit": "Gqueries/s", "ms_per_step": round(1e3 * pel / args.steps, 4),
                 "speedup": round(value / pval, 3)}

    # Serving runs before the batched forms: the server overlaps launches on
    # two streams, and after the batched forms have run in the same process
    # (side streams, a captured graph) its saturation rate drops from 6250 to
    # 3600 baskets/s, most likely because its streams then share some of the
    # process's 4 hardware queues.  A serving process runs nothing else.
    serve = None
    if not args.no_serve:
        log("serving")
        serve = serve_bench(args, dev)

    batched = fp32 = None
    if not args.no_batch:
        log("batched forms")
        batched, fp32 = batch_forms(args, wl, stream)
    mixed = mwl = None
    if not args.no_mixed:
        log("mixed-precision workload")
        mwl = MixedWorkload(wl, dev)
        log("mixed-precision forms")
        mixed = mixed_forms(args, mwl, stream)
    forms = {"per_launch": value}
    if batched:
        forms["batched"] = batched["value"]
        forms["batched_graph"] = batched["graph_value"]
    best_form = max(forms, key=forms.get)
    value_per_launch = value
    value = forms[best_form]
    roof_per_launch = roof
    if best_form.startswith("batched"):
        # The roofline of the form that produced value: its dominant class
        # kernel, event-timed inside the batched timed region.
        roof = batched["roofline"]
    ms_per_step = 1e3 * wl.queries_per_step / (value * 1e9)

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (shapes)")
        cpu = cpu_baseline(wl, args.cpu_seconds)
        # The CPU leg also checks sampled device outputs against the oracle
        # on the inputs the device saw.
        log("sample checks against the oracle")
        for form, promote in ((batched, False), (fp32, True)):
            if form is not None:
                form["check_vs_oracle"] = sample_check(wl.items, wl.shapes, wl.xfs, wl.n, form["_captured"], promote)
        if mixed is not None:
            mixed["check_vs_oracle"] = sample_check(mwl.items, wl.shapes, wl.xfs, wl.n, mixed["_captured"], False)
    for form in (batched, fp32, mixed):
        if form is not None:
            form.pop("_captured", None)

    nav = None
    if not args.no_nav:
        log("navigation")
        nav = nav_bench(args, dev, stream, not args.no_cpu)

    boolean = None
    if not args.no_boolean:
        log("Boolean volumes")
        boolean = boolean_bench(args, dev, stream, not args.no_cpu)

    normals = None
    if not args.no_normal:
        log("surface normals")
        normals = normal_bench(args, dev, stream, not args.no_cpu)

    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "Gqueries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "shape zoo FP64, random placements, 2^%d points per shape, six queries"
                               % int(np.log2(args.points)), "form": best_form,
                   "points_per_shape": args.points, "shapes": [nm for nm, _ in wl.shapes],
                   "shapes_without_kernels": wl.missing, "queries_per_step": wl.queries_per_step},
        "roofline": roof,
        "roofline_per_launch_form": roof_per_launch,
        "forms_gqps": {k: round(v, 3) for k, v in forms.items()},
        "per_launch_value": round(value_per_launch, 3),
        "batched": batched,
        "fp32_storage": fp32,
        "mixed": mixed,
        "plain_fp64_gpu": plain,
        "cpu_baseline": cpu,
        "nav": nav,
        "serve": serve,
        "boolean": boolean,
        "normals": normals,
        "per_kernel": [{k: (round(v, 3) if isinstance(v, float) else v) for k, v in r.items()} for r in rows],
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
