"""GPU parity of the device navigator (HIP kernels through the C-ABI).

Bar: LocatePoint paths, next-volume paths and boundary flags bit-exact
against the reference (golden vectors from the reference's own navigators)
and against the oracle on seeded tracks; steps, safeties, final points and
path lengths within max(1e-9 mm, 1e-12 relative) per step; the per-track hash
of the whole (volume, boundary) sequence of a TraceTrack loop identical.
"""
import ctypes

import numpy as np
import pytest
import torch

import pyoracle as po
from conftest import load_golden
from vecgeom_amd import capi
from vecgeom_amd import geometry as G
from vecgeom_amd import nav as N
from vecgeom_amd import workloads as W

pytestmark = pytest.mark.gpu

KINDS = ((N.ABBOX, "abbox"), (N.SIMPLE, "simple"))
# The device navigators and the oracle navigator each must equal: the
# mixed-precision ABBox navigator gives the ABBox navigator's results.
DEVICE_KINDS = KINDS + ((N.ABBOX_MIXED, "abbox"),)
ORACLE_KIND = {"abbox": N.ABBOX, "simple": N.SIMPLE}
ZOO_REF_NAMES = [n for n, _ in W.NAV_ZOO if n != "Parallelepiped"]


def dev(a, device):
    return tuple(torch.from_numpy(np.ascontiguousarray(a[i])).to(device) for i in range(3))


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def close(a, b, scale=1.0):
    a, b = np.asarray(a), np.asarray(b)
    tol = np.maximum(1e-9, 1e-12 * np.abs(b)) * scale
    both_inf = np.isinf(a) & np.isinf(b) & (np.sign(a) == np.sign(b))
    return both_inf | (np.abs(a - b) <= tol)


def states_host(s):
    return host(s.level), host(s.path), host(s.boundary)


def check_nav(geom, nav_ref, g, device):
    """Device navigator against reference/oracle outputs in dict g."""
    p, d = dev(g["x"], device), dev(g["d"], device)
    s = N.locate_points(geom, p)
    lvl, path, bnd = states_host(s)
    assert (lvl == g["level"]).all()
    assert (path == g["path"]).all()
    assert (bnd == 0).all()
    for kind, tag in DEVICE_KINDS:
        new, step = N.find_next_boundary_and_step(geom, p, d, s, navigator=kind)
        lo, pa, bo = states_host(new)
        assert (lo == g[f"step_{tag}_level"]).all(), tag
        assert (pa == g[f"step_{tag}_path"]).all(), tag
        assert (bo == g[f"step_{tag}_boundary"]).all(), tag
        assert close(host(step), g[f"step_{tag}_step"]).all(), tag
        t = N.trace_tracks(geom, p, d, navigator=kind)
        ns = host(t.nsteps)
        assert (ns == g[f"trace_{tag}_nsteps"]).all(), tag
        assert (t.hash_u64() == g[f"trace_{tag}_hash"]).all(), tag
        fin = np.stack([host(x) for x in t.final])
        assert close(fin, g[f"trace_{tag}_final"], 1 + ns).all(), tag
        assert close(host(t.travelled), g[f"trace_{tag}_travelled"], 1 + ns).all(), tag
    assert close(host(N.get_safety(geom, p, s)), g["safety"]).all()


def oracle_dict(nav, p, d):
    level, path = nav.locate(p)
    g = {"x": p, "d": d, "level": level, "path": path, "safety": nav.safety(p, level, path, nthreads=8)}
    for kind, tag in KINDS:
        lo, pa, bo, st = nav.step(p, d, level, path, kind=kind, nthreads=8)
        g.update({f"step_{tag}_level": lo, f"step_{tag}_path": pa, f"step_{tag}_boundary": bo, f"step_{tag}_step": st})
        ns, fp, tr, hs = nav.trace(p, d, kind=kind, nthreads=8)
        g.update({f"trace_{tag}_nsteps": ns, f"trace_{tag}_final": fp, f"trace_{tag}_travelled": tr,
                  f"trace_{tag}_hash": hs})
    return g


@pytest.fixture(scope="module")
def nested_dev(gpu):
    gm, flat = W.nested_geometry()
    return gm, flat, G.DeviceGeometry(flat)


def test_gpu_nav_nested_golden(gpu, nested_dev):
    _, _, geom = nested_dev
    check_nav(geom, None, load_golden("nav_nested.npz"), gpu)


def test_gpu_nav_zoo_golden(gpu):
    _, flat, _ = W.nav_zoo_geometry(names=ZOO_REF_NAMES)
    check_nav(G.DeviceGeometry(flat), None, load_golden("nav_zoo.npz"), gpu)


def test_gpu_nav_zoo_all_shapes_vs_oracle(gpu, oracle):
    # Includes Parallelepiped, which the reference cannot box (no Extent).
    _, flat, placed = W.nav_zoo_geometry()
    centres = [flat.placed_transform[12 * i:12 * i + 3] for _, i in placed]
    p, d = W.nav_zoo_tracks(6000, centres, seed=4242)
    check_nav(G.DeviceGeometry(flat), None, oracle_dict(po.Nav(oracle, flat), p, d), gpu)


def test_gpu_nav_nested_random_vs_oracle(gpu, oracle, nested_dev):
    _, flat, geom = nested_dev
    rng = np.random.Generator(np.random.Philox(77))
    n = 8192
    r = rng.uniform(0.0, 1000.0, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    p = np.stack([r * np.cos(ph), r * np.sin(ph), rng.uniform(-1950, 1950, n)])
    d = W.isotropic(n, rng)
    check_nav(geom, None, oracle_dict(po.Nav(oracle, flat), p, d), gpu)


def test_gpu_nav_limited_pstep_vs_oracle(gpu, oracle, nested_dev):
    _, flat, geom = nested_dev
    p, d = W.track_starts(4096, seed=515)
    p = p + d * 150.0  # start between the inner shells
    nav = po.Nav(oracle, flat)
    level, path = nav.locate(p)
    for kind, tag in DEVICE_KINDS:
        lo, pa, bo, st = nav.step(p, d, level, path, pstep=7.5, kind=ORACLE_KIND[tag])
        pd, dd = dev(p, gpu), dev(d, gpu)
        s = N.locate_points(geom, pd)
        ps = torch.full((p.shape[1],), 7.5, dtype=torch.float64, device=gpu)
        new, step = N.find_next_boundary_and_step(geom, pd, dd, s, pstep=ps, navigator=kind)
        glo, gpa, gbo = states_host(new)
        assert (glo == lo).all() and (gpa == pa).all() and (gbo == bo).all(), tag
        assert close(host(step), st).all(), tag
        assert (bo == 0).any() and (bo == 1).any()


def test_gpu_step_loop_matches_trace(gpu, nested_dev):
    """The basket API (NavStates + FindNextBoundaryAndStep per step) gives
    the same tracks as the fused trace kernel."""
    _, _, geom = nested_dev
    p0, d0 = W.track_starts(4096, seed=616)
    p, d = list(dev(p0, gpu)), dev(d0, gpu)
    cur = N.locate_points(geom, tuple(p))
    new = N.NavStates(geom, cur.capacity, device=gpu)
    steps = torch.zeros(p0.shape[1], dtype=torch.int32, device=gpu)
    for _ in range(2000):
        active = cur.level > 0
        if not bool(active.any()):
            break
        new, step = N.find_next_boundary_and_step(geom, tuple(p), d, cur, new=new)
        s = step + 1e-6
        for i in range(3):
            p[i] = torch.where(active, p[i] + d[i] * s, p[i])
        steps += active.int()
        cur, new = new, cur
    else:
        pytest.fail("tracks did not leave the world within 2000 steps")
    t = N.trace_tracks(geom, dev(p0, gpu), d)
    assert (host(t.nsteps) == host(steps)).all()
    for i in range(3):
        assert (host(t.final[i]) == host(p[i])).all()


def test_gpu_nav_full_size_config(gpu, oracle, nested_dev):
    """2^20 tracks from the vertex distribution, ABBox navigator, until they
    leave the world; a strided sample is compared with the oracle."""
    _, flat, geom = nested_dev
    n = 1 << 20
    p0, d0 = W.track_starts(n)
    t = N.trace_tracks(geom, dev(p0, gpu), dev(d0, gpu), max_steps=100000)
    ns = host(t.nsteps)
    assert (ns > 0).all() and ns.max() < 100000
    fin = np.stack([host(x) for x in t.final])
    # Every track ends just outside the world box (half-lengths 1000, 1000, 2000).
    out = (np.abs(fin[0]) > 1000) | (np.abs(fin[1]) > 1000) | (np.abs(fin[2]) > 2000)
    assert out.all()
    idx = np.arange(0, n, 64)
    ons, ofp, otr, ohs = po.Nav(oracle, flat).trace(p0[:, idx], d0[:, idx], kind=0, nthreads=16)
    assert (ns[idx] == ons).all()
    assert (t.hash_u64()[idx] == ohs).all()
    assert close(fin[:, idx], ofp, 1 + ons).all()
    exact = host(t.exact)
    print(f"steps {int(ns.sum())}, mean {ns.mean():.2f}, exact-path steps {int(exact.sum())}")
    # The mixed-precision navigator: every one of the 2^20 tracks identical
    # to the FP64 navigator's, bit for bit.
    m = N.trace_tracks(geom, dev(p0, gpu), dev(d0, gpu), max_steps=100000, navigator=N.ABBOX_MIXED)
    assert (host(m.nsteps) == ns).all()
    assert (m.hash_u64() == t.hash_u64()).all()
    for i in range(3):
        assert np.array_equal(host(m.final[i]), fin[i])
    assert np.array_equal(host(m.travelled), host(t.travelled))


def test_gpu_nav_states_pool_c_abi(gpu, nested_dev):
    _, _, geom = nested_dev
    L = capi.lib()
    s = N.CNavStates()
    capi.call("vg_nav_states_create", geom.handle, 100, 0, ctypes.byref(s))
    try:
        assert s.depth == geom.max_depth() and s.capacity == 100
        lvl = np.arange(64, dtype=np.int32) % 5
        path = np.arange(64 * s.depth, dtype=np.int32).reshape(s.depth, 64)
        bnd = (np.arange(64) % 2).astype(np.uint8)
        pp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        capi.call("vg_nav_states_from_host", ctypes.byref(s), 64, pp(lvl), pp(path), pp(bnd))
        l2, p2, b2 = np.zeros_like(lvl), np.zeros_like(path), np.zeros_like(bnd)
        capi.call("vg_nav_states_to_host", ctypes.byref(s), 64, pp(l2), pp(p2), pp(b2))
        assert (l2 == lvl).all() and (p2 == path).all() and (b2 == bnd).all()
        rc, _ = capi.status("vg_nav_states_to_host", ctypes.byref(s), 101, pp(l2), pp(p2), pp(b2))
        assert rc == -5
    finally:
        capi.call("vg_nav_states_free", ctypes.byref(s))
    # Argument errors are reported, not launched.
    x = torch.zeros(8, dtype=torch.float64, device=gpu)
    small = N.NavStates(geom, 4, device=gpu)
    rc, msg = capi.status("vg_locate_points", geom.handle, x.data_ptr(), x.data_ptr(), x.data_ptr(), 8,
                          ctypes.byref(small.c()), None)
    assert rc == -5 and "capacity" in msg
    shallow = small.c()
    shallow.depth = 1
    rc, _ = capi.status("vg_locate_points", geom.handle, x.data_ptr(), x.data_ptr(), x.data_ptr(), 4,
                        ctypes.byref(shallow), None)
    assert rc == -1
    c = small.c()
    st = torch.zeros(4, dtype=torch.float64, device=gpu)
    rc, msg = capi.status("vg_find_next_boundary_and_step", geom.handle, 0, *[x.data_ptr()] * 6, None, 4,
                          ctypes.byref(c), ctypes.byref(c), st.data_ptr(), None)
    assert rc == -1 and "alias" in msg
    rc, _ = capi.status("vg_trace_tracks", geom.handle, 7, *[x.data_ptr()] * 6, 0, 1.0, 10, *[None] * 8)
    assert rc == -1


def test_gpu_basket_navigator_vs_oracle(gpu, oracle, nested_dev):
    """SimpleNavigator's basket step (navigator kind 3) on baskets of tracks
    sharing a current volume, against the oracle (itself checked against the
    reference's basket interface); also a basket of world states whose
    points lie outside the world box, where it differs from the scalar
    navigator."""
    from test_oracle_nav import baskets_by_volume
    _, flat, geom = nested_dev
    rng = np.random.Generator(np.random.Philox(2991))
    n = 6000
    r = rng.uniform(0.0, 1000.0, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    p = np.stack([r * np.cos(ph), r * np.sin(ph), rng.uniform(-1900, 1900, n)])
    d = W.isotropic(n, rng)
    nav = po.Nav(oracle, flat)
    level, path = nav.locate(p)
    baskets = [(p[:, i].copy(), d[:, i].copy(), level[i].copy(), path[:, i].copy())
               for i in baskets_by_volume(level, path)]
    m = 64
    q = np.stack([rng.uniform(1001, 1100, m), rng.uniform(-900, 900, m), rng.uniform(-1900, 1900, m)])
    lw = np.ones(m, np.int32)
    pw = np.full((nav.depth, m), -1, np.int32)
    pw[0] = flat.world
    baskets.append((q, W.isotropic(m, rng), lw, pw))
    for q, v, lo, pa in baskets:
        s = N.NavStates.from_arrays(geom, lo, pa, device=gpu)
        for pstep in (np.inf, 7.5):
            olo, opa, obo, ost = nav.step(q, v, lo, pa, pstep=pstep, kind=3)
            ps = torch.full((q.shape[1],), pstep, dtype=torch.float64, device=gpu)
            new, step = N.find_next_boundary_and_step(geom, dev(q, gpu), dev(v, gpu), s, pstep=ps,
                                                      navigator=N.SIMPLE_BASKET)
            glo, gpa, gbo = states_host(new)
            assert (glo == olo).all() and (gpa == opa).all() and (gbo == obo).all()
            assert close(host(step), ost).all()
