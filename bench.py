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


PROFILE = os.path.join(ROOT, "profiles", "shape_zoo_fp64_kernels.md")
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
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline sample duration")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-FP64-form denominator")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched and FP32-storage forms")
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
            self.items.append(it)
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
            out += [B.Task(pid, B.CONTAINS, p, out=o["contains"]), B.Task(pid, B.INSIDE, p, out=o["inside"]),
                    B.Task(pid, B.DISTANCE_TO_IN, pout, dout, it["smout"], out=o["dist_in"]),
                    B.Task(pid, B.SAFETY_TO_IN, pout, out=o["safe_in"]),
                    B.Task(pid, B.DISTANCE_TO_OUT, pin, din, it["smin"], out=o["dist_out"]),
                    B.Task(pid, B.SAFETY_TO_OUT, pin, out=o["safe_out"])]
        return out

    def launches(self):
        """(key, callable) for every launch of one step."""
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


def cpu_baseline(wl, target_s):
    """The reference's CPU basket path on the host: its SOA3D basket
    virtuals, built from the reference's sources with its Release flags
    (oracle/_ref/libvgref_fast.so), over contiguous ranges on every host core;
    the oracle restatement stands in where that build is absent.  The same
    queries on a bounded, evenly strided sample of the same inputs."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    ref = po.reference(fast=True)
    o = ref if ref is not None and ref.has("basket") else po.oracle()
    kind = "reference" if o is ref else "port"
    nthreads = max(1, min(16, os.cpu_count() or 1))
    # Calibrate: one shape, all queries, small sample.
    frac = 1.0 / 64
    total_q = 0
    reps = 1
    t0 = time.perf_counter()
    while True:
        total_q = 0
        t0 = time.perf_counter()
        for it, (name, u), xf in [z for _ in range(reps) for z in zip(wl.items, wl.shapes, wl.xfs)]:
            h = o.placed(u.type, u.params, xf.array())
            run = h.basket if kind == "reference" else h.query
            hp = it["host"]
            step = int(round(1 / frac))
            sel_all = slice(0, wl.n, step)
            iout = it["outside_idx"][::step]
            iin = it["inside_idx"][::step]
            P, D, LP, LD = hp["p"], hp["d"], hp["lp"], hp["ld"]
            run(po.Q_CONTAINS, P[:, sel_all], nthreads=nthreads)
            run(po.Q_INSIDE, P[:, sel_all], nthreads=nthreads)
            run(po.Q_DIST_IN, P[:, iout], D[:, iout], np.full(len(iout), 1e30), nthreads=nthreads)
            run(po.Q_SAFE_IN, P[:, iout], nthreads=nthreads)
            run(po.Q_DIST_OUT, LP[:, iin], LD[:, iin], np.full(len(iin), 1e30), nthreads=nthreads)
            run(po.Q_SAFE_OUT, LP[:, iin], nthreads=nthreads)
            total_q += 2 * len(range(0, wl.n, step)) + 2 * len(iout) + 2 * len(iin)
        dt = time.perf_counter() - t0
        if dt >= 0.5 * target_s:
            break
        if frac >= 1.0:
            reps = max(reps + 1, int(reps * target_s / max(dt, 1e-3)))
        else:
            frac = min(1.0, frac * max(2.0, target_s / max(dt, 1e-3)))
    return {
        "value": total_q / dt / 1e9,
        "unit": "Gqueries/s",
        "cores": nthreads,
        "kind": kind,
        "sample": f"every {int(round(1 / frac))}th query of the same workload x{reps} ({total_q} queries, {dt:.1f} s), "
                  + ("reference SOA3D basket virtuals, Scalar backend, Release flags (-O3 -ffast-math)"
                     if kind == "reference" else "oracle C++ restatement of the reference Scalar backend")
                  + ", std::thread over contiguous ranges",
    }


def batch_forms(args, wl, stream):
    """The step as one batch (one launch per shape weight class), autotuned;
    the same batch replayed from a captured hipGraph; and the FP32-storage
    form (float32 inputs promoted to double in registers)."""
    import torch

    from vecgeom_amd import batch as B

    res = {}
    bt = B.Batch(wl.geom, wl.tasks())
    best_ms = bt.autotune(3)
    it, blocks, items = bt.launch()
    el, _ = time_steps([(("batch", "all"), bt.run)], args.steps, stream, watch=[])
    res = {"value": round(wl.queries_per_step * args.steps / el / 1e9, 3), "ms_per_step": round(1e3 * el / args.steps, 4),
           "items_per_lane": it, "blocks": blocks, "autotune_ms": round(best_ms, 4),
           "achieved_gbs": round(wl.bytes_per_step / (el / args.steps) / 1e9, 1)}
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
    f32 = B.Batch(wl.geom, wl.tasks(fp32=True))
    f32.autotune(3)
    it32, blocks32, _ = f32.launch()
    el32, _ = time_steps([(("batch32", "all"), f32.run)], args.steps, stream, watch=[])
    fp32 = {"value": round(wl.queries_per_step * args.steps / el32 / 1e9, 3), "unit": "Gqueries/s",
            "ms_per_step": round(1e3 * el32 / args.steps, 4), "items_per_lane": it32, "blocks": blocks32,
            "achieved_gbs": round(wl.bytes_per_step_f32 / (el32 / args.steps) / 1e9, 1),
            "note": "float32 inputs promoted to double before any arithmetic: results equal the FP64 form on the "
                    "same float values (max error 0 vs the oracle on promoted inputs, classifications identical)"}
    f32.free()
    return res, fp32


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
    Latency is per request, submission to completion.  The batched form
    merges all pending requests into one launch on each of two streams; the
    per-basket form launches each request alone, one at a time.  Rates are
    0.25, 0.5 and 0.9 of the per-basket form's saturation rate and 0.5 and
    0.9 of the batched form's (saturation: all baskets arriving at once)."""
    from vecgeom_amd import geometry as G
    from vecgeom_amd import serve as S
    from vecgeom_amd import workloads as W

    _, flat = W.nested_geometry()
    geom = G.DeviceGeometry(flat)
    rng = np.random.Generator(np.random.Philox(2024))
    nb, steps = args.serve_baskets, args.serve_steps
    sizes = np.exp(rng.uniform(np.log(16), np.log(65536), nb)).astype(np.uint64)
    p, d = W.track_starts(int(sizes.sum()), seed=4040)
    srv = S.Server(geom, 32 << 30)
    forms = {"batched": (4096, 2), "per_basket": (1, 1)}

    def run(form, arrival):
        srv.set_batching(*forms[form])
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
    run("batched", zero)  # warm-up
    sat = {f: run(f, zero) for f in forms}
    base_rate = nb / sat["per_basket"]["wall_s"]
    batched_rate = nb / sat["batched"]["wall_s"]
    rates = {}
    for rate in [f * base_rate for f in (0.25, 0.5, 0.9)] + [f * batched_rate for f in (0.5, 0.9)]:
        gaps = rng.exponential(1.0 / rate, nb)
        arrival = np.cumsum(gaps) - gaps[0]
        rates[f"{rate:.0f}/s"] = {f: run(f, arrival) for f in forms}
    res = {"workload": f"nested geometry, {nb} baskets, sizes log-uniform 16..65536 tracks "
                       f"({int(sizes.sum())} tracks), {steps} steps per basket, Poisson arrivals, 32 GiB arena",
           "unit_steps": "G steps/s", "saturation": sat,
           "saturation_basket_rate_per_basket_form": round(base_rate, 1),
           "saturation_basket_rate_batched_form": round(batched_rate, 1),
           "fixed_rates": rates, "dtype": "f64"}
    srv.destroy()
    geom.free()
    return res


def nav_cpu_baseline(flat, p0, d0, target_s):
    """The reference's own ABBoxNavigator (built from its sources with its
    Release flags) on one core, on a bounded prefix of the same tracks; the
    navigator keeps a process-wide hit list, so it runs on one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po

    ref = po.reference(fast=True)
    kind, lib, cores = "reference", ref, 1
    if ref is None:
        kind, lib, cores = "port", po.oracle(), max(1, min(16, os.cpu_count() or 1))
    nav = po.Nav(lib, flat)
    m = 256
    while True:
        t0 = time.perf_counter()
        ns, _, _, _ = nav.trace(p0[:, :m], d0[:, :m], kind=0, nthreads=cores)
        dt = time.perf_counter() - t0
        if dt >= 0.5 * target_s or m >= p0.shape[1]:
            break
        m = min(p0.shape[1], int(m * max(2.0, target_s / max(dt, 1e-3))))
    steps = int(ns.sum())
    return {"value": round(steps / dt / 1e9, 6), "unit": "G steps/s", "cores": cores, "kind": kind,
            "sample": f"first {m} tracks of the same workload ({steps} steps, {dt:.1f} s), "
                      + ("reference ABBoxNavigator, Release flags (-O3 -ffast-math)" if kind == "reference"
                         else "oracle C++ restatement")}


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
    # breakdown and pick the dominant kernel.
    _, cen = time_steps(launches, 3, stream)
    keys = [key for key, _ in launches]
    count = {(it["name"], q): c for it in wl.items for q, c in it["count"].items()}
    rows = []
    for k, key in enumerate(keys):
        ms = float(np.mean(cen[k]))
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
    forms = {"per_launch": value}
    if batched:
        forms["batched"] = batched["value"]
        forms["batched_graph"] = batched["graph_value"]
    best_form = max(forms, key=forms.get)
    value_per_launch = value
    value = forms[best_form]
    ms_per_step = 1e3 * wl.queries_per_step / (value * 1e9)

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (shapes)")
        cpu = cpu_baseline(wl, args.cpu_seconds)

    nav = None
    if not args.no_nav:
        log("navigation")
        nav = nav_bench(args, dev, stream, not args.no_cpu)

    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "Gqueries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "shape zoo FP64, random placements, 2^%d points per shape, six queries"
                               % int(np.log2(args.points)), "form": best_form,
                   "points_per_shape": args.points, "shapes": [nm for nm, _ in wl.shapes],
                   "shapes_without_kernels": wl.missing, "queries_per_step": wl.queries_per_step},
        "roofline": roof,
        "forms_gqps": {k: round(v, 3) for k, v in forms.items()},
        "per_launch_value": round(value_per_launch, 3),
        "batched": batched,
        "fp32_storage": fp32,
        "plain_fp64_gpu": plain,
        "cpu_baseline": cpu,
        "nav": nav,
        "serve": serve,
        "per_kernel": [{k: (round(v, 3) if isinstance(v, float) else v) for k, v in r.items()} for r in rows],
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
