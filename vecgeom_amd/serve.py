"""Navigation server with continuous batching (vg_server_* C-ABI).

Clients open baskets of tracks, which stay resident on the device with
their navigation states, and submit requests; one request is one step of
the TraceTrack loop (test/globalbenchmarks/TraceTrack.cpp:188-220) for
every track of its basket.  A native thread merges all pending requests
into one launch whenever one of two device streams is free.  Latencies are
measured by the server from submission to completion.
"""
import ctypes

import numpy as np

from . import capi
from .nav import ABBOX

c_vp = ctypes.c_void_p
c_u64 = ctypes.c_uint64
c_i64 = ctypes.c_int64
c_i32 = ctypes.c_int32
c_dbl = ctypes.c_double
P = ctypes.POINTER

capi.register({
    "vg_server_create": [c_vp, c_u64, c_i32, c_dbl, P(c_vp)],
    "vg_server_capacity": [c_vp, P(c_u64), P(c_u64)],
    "vg_server_open": [c_vp, c_u64] + [c_vp] * 6 + [P(c_i64)],
    "vg_server_submit": [c_vp, c_i64, P(c_i64)],
    "vg_server_wait": [c_vp, c_i64, P(c_dbl)],
    "vg_server_poll": [c_vp, c_i64, P(c_i32), P(c_dbl)],
    "vg_server_read": [c_vp, c_i64, c_u64] + [c_vp] * 8,
    "vg_server_basket_size": [c_vp, c_i64, P(c_u64)],
    "vg_server_close": [c_vp, c_i64],
    "vg_server_stats": [c_vp, P(c_u64), P(c_u64), P(c_u64)],
    "vg_server_destroy": [c_vp],
    "vg_server_set_batching": [c_vp, c_i32, c_i32],
    "vg_server_replay": [c_vp, c_u64, c_vp, c_vp, c_i32] + [c_vp] * 6 + [c_vp, c_vp, c_vp],
})


def slot_bytes(depth):
    """Arena bytes per resident track."""
    return 8 * 8 + 3 * 4 + 4 * depth


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Server:
    """vg_server over a geometry; hbm_budget bounds the track arena."""

    def __init__(self, geom, hbm_budget, navigator=ABBOX, pstep=float("inf")):
        self.geom = geom  # the geometry must outlive the server
        h = c_vp()
        capi.call("vg_server_create", geom.handle, int(hbm_budget), int(navigator), float(pstep), ctypes.byref(h))
        self.handle = h

    def capacity(self):
        tot, free = c_u64(), c_u64()
        capi.call("vg_server_capacity", self.handle, ctypes.byref(tot), ctypes.byref(free))
        return tot.value, free.value

    @staticmethod
    def _tracks(p, d):
        """Six contiguous float64 host arrays of one length (the C call reads
        n entries from each)."""
        if len(p) != 3 or len(d) != 3:
            raise ValueError("p and d must be (x, y, z) triples")
        arrs = [_f64(a).ravel() for a in list(p) + list(d)]
        if any(a.size != arrs[0].size for a in arrs):
            raise ValueError("positions and directions must all hold the same number of tracks")
        return arrs

    def open(self, p, d):
        """Admits a basket of tracks (host arrays, shape (3, n)); returns its id."""
        arrs = self._tracks(p, d)
        b = c_i64()
        capi.call("vg_server_open", self.handle, arrs[0].size, *[a.ctypes.data for a in arrs], ctypes.byref(b))
        return b.value

    def try_open(self, p, d):
        """open(), but returns None when the basket does not fit the budget."""
        arrs = self._tracks(p, d)
        p = arrs[:3]
        d = arrs[3:]
        b = c_i64()
        rc, msg = capi.status("vg_server_open", self.handle, p[0].size, *[a.ctypes.data for a in p + d],
                              ctypes.byref(b))
        if rc == capi.VG_ERR_CAPACITY:
            return None
        if rc != capi.VG_OK:
            raise RuntimeError(f"vg_server_open failed ({rc}): {msg}")
        return b.value

    def submit(self, basket):
        r = c_i64()
        capi.call("vg_server_submit", self.handle, int(basket), ctypes.byref(r))
        return r.value

    def wait(self, request):
        """Blocks until the request completed; returns its latency in seconds."""
        lat = c_dbl()
        capi.call("vg_server_wait", self.handle, int(request), ctypes.byref(lat))
        return lat.value

    def poll(self, request):
        """Returns the latency if the request completed, else None."""
        done, lat = c_i32(), c_dbl()
        capi.call("vg_server_poll", self.handle, int(request), ctypes.byref(done), ctypes.byref(lat))
        return lat.value if done.value else None

    def basket_size(self, basket):
        n = c_u64()
        capi.call("vg_server_basket_size", self.handle, int(basket), ctypes.byref(n))
        return n.value

    def read(self, basket, n=None):
        """The basket's tracks: dict of x, y, z, level, boundary, nsteps,
        hash, last_step.  The arrays are sized by the basket itself; n, if
        given, must equal its track count."""
        size = self.basket_size(basket)
        if n is not None and int(n) != size:
            raise ValueError(f"basket {basket} holds {size} tracks, not {n}")
        n = size
        out = {"x": np.empty(n), "y": np.empty(n), "z": np.empty(n), "level": np.empty(n, np.int32),
               "boundary": np.empty(n, np.int32), "nsteps": np.empty(n, np.int32),
               "hash": np.empty(n, np.uint64), "last_step": np.empty(n)}
        capi.call("vg_server_read", self.handle, int(basket), n,
                  *[out[k].ctypes.data for k in ("x", "y", "z", "level", "boundary", "nsteps", "hash", "last_step")])
        return out

    def close(self, basket):
        capi.call("vg_server_close", self.handle, int(basket))

    def stats(self):
        a, b, c = c_u64(), c_u64(), c_u64()
        capi.call("vg_server_stats", self.handle, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return {"launches": a.value, "requests": b.value, "track_steps": c.value}

    def set_batching(self, max_requests_per_launch, streams):
        """(1, 1) is the per-basket form; the default is (4096, 2)."""
        capi.call("vg_server_set_batching", self.handle, int(max_requests_per_launch), int(streams))

    def replay(self, arrival_s, sizes, steps_per_basket, p, d):
        """Runs the load generator: baskets of sizes[i] tracks (taken in
        order from the concatenated p, d of shape (3, sum(sizes))) arrive at
        arrival_s[i], wait for admission, take steps_per_basket steps and
        close.  Returns (latency (nbaskets, steps), admit_wait, wall_s)."""
        arrival_s = _f64(arrival_s)
        sizes = np.ascontiguousarray(sizes, dtype=np.uint64)
        p = [_f64(a) for a in p]
        d = [_f64(a) for a in d]
        if any(a.size != int(sizes.sum()) for a in p + d):
            raise ValueError("p and d must hold sum(sizes) tracks")
        nb = arrival_s.size
        lat = np.zeros((nb, steps_per_basket))
        wait = np.zeros(nb)
        wall = c_dbl()
        capi.call("vg_server_replay", self.handle, nb, arrival_s.ctypes.data, sizes.ctypes.data,
                  int(steps_per_basket), *[a.ctypes.data for a in p + d], lat.ctypes.data, wait.ctypes.data,
                  ctypes.byref(wall))
        return lat, wait, wall.value

    def destroy(self):
        if self.handle:
            h, self.handle = self.handle, None
            capi.call("vg_server_destroy", h)

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
