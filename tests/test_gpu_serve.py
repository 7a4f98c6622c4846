"""GPU tests of the navigation server (vg_server_* through the C-ABI).

Each request is one TraceTrack step of every track of a basket, so after K
requests a basket's tracks must equal vg_trace_tracks with max_steps = K
bit for bit (position, step count, hash of the volume/boundary sequence),
whether the basket ran alone on the server or merged with other baskets of
uneven sizes submitted concurrently from several client threads.
"""
import threading

import numpy as np
import pytest
import torch

from vecgeom_amd import capi
from vecgeom_amd import geometry as G
from vecgeom_amd import nav as N
from vecgeom_amd import serve as S
from vecgeom_amd import workloads as W

pytestmark = pytest.mark.gpu

STEPS = 6


@pytest.fixture(scope="module")
def nested(gpu):
    gm, flat = W.nested_geometry()
    return G.DeviceGeometry(flat)


def traced(geom, p, d, k, device):
    """vg_trace_tracks with max_steps = k, on the host."""
    t = N.trace_tracks(geom, tuple(torch.from_numpy(np.ascontiguousarray(a)).to(device) for a in p),
                       tuple(torch.from_numpy(np.ascontiguousarray(a)).to(device) for a in d), max_steps=k)
    torch.cuda.synchronize()
    return {"x": t.final[0].cpu().numpy(), "y": t.final[1].cpu().numpy(), "z": t.final[2].cpu().numpy(),
            "nsteps": t.nsteps.cpu().numpy(), "hash": t.hash_u64()}


def assert_same(got, want, tag):
    for key in ("x", "y", "z", "nsteps", "hash"):
        assert np.array_equal(got[key], want[key]), f"{tag}: {key}"


def baskets(sizes, seed):
    out = []
    for i, n in enumerate(sizes):
        p, d = W.track_starts(n, seed=seed + i)
        out.append((p, d))
    return out


def test_server_basket_alone_equals_trace(gpu, nested):
    p, d = W.track_starts(3000, seed=41)
    srv = S.Server(nested, 64 << 20)
    b = srv.open(p, d)
    lats = [srv.wait(srv.submit(b)) for _ in range(STEPS)]
    assert all(x > 0 for x in lats)
    got = srv.read(b, p.shape[1])
    assert_same(got, traced(nested, p, d, STEPS, gpu), "alone")
    assert (got["level"] > 0).any()
    srv.close(b)
    st = srv.stats()
    assert st["launches"] == STEPS and st["requests"] == STEPS and st["track_steps"] == STEPS * 3000
    srv.destroy()


@pytest.mark.parametrize("navigator", [N.ABBOX, N.ABBOX_MIXED])
def test_server_mixed_concurrent_clients_equal_alone(gpu, nested, navigator):
    """Uneven baskets (1 .. 70000 tracks, and an empty one) stepped by four
    client threads at once: every basket equals its own trace (the FP64 ABBox
    navigator's, also for the mixed-precision server), and the server merged
    requests into fewer launches than requests."""
    sizes = [1, 17, 128, 129, 4099, 70000, 0, 333, 2500, 9000, 31, 64]
    data = baskets(sizes, seed=900)
    srv = S.Server(nested, 256 << 20, navigator=navigator)
    ids = [srv.open(p, d) for p, d in data]
    errors = []

    def client(mine):
        try:
            for _ in range(STEPS):
                reqs = [srv.submit(ids[i]) for i in mine]
                for r in reqs:
                    srv.wait(r)
        except Exception as e:  # reported by the main thread
            errors.append(e)

    threads = [threading.Thread(target=client, args=(list(range(k, len(sizes), 4)),)) for k in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(120)
    assert not errors, errors
    st = srv.stats()
    assert st["requests"] == STEPS * len(sizes)
    assert st["track_steps"] == STEPS * sum(sizes)
    assert st["launches"] < st["requests"], st
    for i, (p, d) in enumerate(data):
        n = sizes[i]
        if n == 0:
            continue
        assert_same(srv.read(ids[i], n), traced(nested, p, d, STEPS, gpu), f"basket {i} (n={n})")
    for b in ids:
        srv.close(b)
    srv.destroy()


def test_server_admission_by_footprint(gpu, nested):
    depth = nested.max_depth()
    srv = S.Server(nested, 1000 * S.slot_bytes(depth))
    assert srv.capacity() == (1000, 1000)
    (p1, d1), (p2, d2), (p3, d3) = baskets([600, 600, 400], seed=5)
    b1 = srv.open(p1, d1)
    assert srv.capacity() == (1000, 400)
    assert srv.try_open(p2, d2) is None  # does not fit
    b3 = srv.open(p3, d3)
    assert srv.capacity() == (1000, 0)
    r = srv.submit(b1)
    srv.wait(r)
    srv.close(b1)
    b2 = srv.try_open(p2, d2)  # fits once b1's slots are free
    assert b2 is not None
    for _ in range(STEPS):
        srv.wait(srv.submit(b2))
    assert_same(srv.read(b2, 600), traced(nested, p2, d2, STEPS, gpu), "re-used slots")
    srv.close(b2)
    srv.close(b3)
    assert srv.capacity() == (1000, 1000)
    srv.destroy()


def test_server_replay_admits_as_slots_free(gpu, nested):
    """The load generator with a budget of two baskets' worth: later baskets
    wait for admission, every basket takes all its steps."""
    sizes = np.array([500, 500, 500, 500, 20, 499], dtype=np.uint64)
    data = baskets([int(n) for n in sizes], seed=77)
    p = np.concatenate([x[0] for x in data], axis=1)
    d = np.concatenate([x[1] for x in data], axis=1)
    srv = S.Server(nested, 1000 * S.slot_bytes(nested.max_depth()))
    arrival = np.zeros(len(sizes))
    lat, wait, wall = srv.replay(arrival, sizes, 3, p, d)
    assert (lat > 0).all() and wall > 0
    assert (wait[2:] > 0).all()  # admitted only after earlier baskets closed
    st = srv.stats()
    assert st["requests"] == 3 * len(sizes) and st["track_steps"] == 3 * int(sizes.sum())
    assert srv.capacity() == (1000, 1000)
    srv.set_batching(1, 1)
    lat, _, _ = srv.replay(np.linspace(0, 1e-3, len(sizes)), sizes, 2, p, d)
    assert (lat > 0).all()
    srv.destroy()


def test_server_argument_errors(gpu, nested):
    srv = S.Server(nested, 1 << 20)
    rc, msg = capi.status("vg_server_submit", srv.handle, 12345, S.ctypes.byref(S.c_i64()))
    assert rc == capi.VG_ERR_INVALID and "basket" in msg
    rc, _ = capi.status("vg_server_close", srv.handle, 12345)
    assert rc == capi.VG_ERR_INVALID
    rc, _ = capi.status("vg_server_set_batching", srv.handle, 0, 1)
    assert rc == capi.VG_ERR_INVALID
    rc, _ = capi.status("vg_server_set_batching", srv.handle, 1, 3)
    assert rc == capi.VG_ERR_INVALID
    h = S.c_vp()
    rc, _ = capi.status("vg_server_create", nested.handle, 10, 0, 1e30, S.ctypes.byref(h))
    assert rc == capi.VG_ERR_CAPACITY and not h.value
    big = np.zeros((3, 20000))
    with pytest.raises(RuntimeError, match="larger than the whole arena"):
        srv.replay([0.0], [20000], 1, big, big)
    srv.destroy()
