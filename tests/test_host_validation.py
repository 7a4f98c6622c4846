"""Argument checks of the Python layer, on the CPU: every array handed to a
kernel must be a contiguous CUDA tensor of the expected dtype and length on
the points' device, so a host tensor or a short output raises before any
call reaches the library (no GPU is touched here)."""
import numpy as np
import pytest
import torch

from vecgeom_amd import batch as B
from vecgeom_amd import queries as Q
from vecgeom_amd import serve as S


class _Geom:
    handle = None  # never reached: the checks raise first


def _soa(n, dtype=torch.float64):
    return tuple(torch.zeros(n, dtype=dtype) for _ in range(3))


CALLS = [
    ("contains", lambda p, d: Q.contains(_Geom(), 0, p)),
    ("inside", lambda p, d: Q.inside(_Geom(), 0, p)),
    ("distance_to_in", lambda p, d: Q.distance_to_in(_Geom(), 0, p, d)),
    ("distance_to_out", lambda p, d: Q.distance_to_out(_Geom(), 0, p, d)),
    ("placed_distance_to_out", lambda p, d: Q.placed_distance_to_out(_Geom(), 0, p, d)),
    ("safety_to_in", lambda p, d: Q.safety_to_in(_Geom(), 0, p)),
    ("safety_to_out", lambda p, d: Q.safety_to_out(_Geom(), 0, p)),
    ("normal", lambda p, d: Q.normal(_Geom(), 0, p)),
    ("distance_to_out_normal", lambda p, d: Q.distance_to_out_normal(_Geom(), 0, p, d)),
]


@pytest.mark.parametrize("name,fn", CALLS)
def test_host_tensors_rejected(name, fn):
    with pytest.raises(ValueError):
        fn(_soa(8), _soa(8))


def test_mismatched_triples_rejected():
    p = _soa(8)
    with pytest.raises(ValueError):
        Q.contains(_Geom(), 0, (p[0], p[1]))
    with pytest.raises(ValueError):
        Q.contains(_Geom(), 0, (p[0], p[1], torch.zeros(7, dtype=torch.float64)))
    with pytest.raises(ValueError):
        Q.contains(_Geom(), 0, (p[0], p[1], torch.zeros(8, dtype=torch.float32)))


def test_array_checker_rules():
    dev = torch.device("cpu")
    # The rule applied to out, step_max, inout and next arrays.
    with pytest.raises(ValueError):
        Q._check_arr(torch.zeros(8, dtype=torch.float64), "out", torch.float64, 8, dev)  # not on a GPU
    with pytest.raises(ValueError):
        Q._check_arr(np.zeros(8), "out", torch.float64, 8, dev)  # not a tensor
    with pytest.raises(ValueError):
        Q._out(torch.zeros(8, dtype=torch.int32), torch.float64, 8, dev)


def test_task_rejects_host_and_short_arrays():
    with pytest.raises(ValueError):
        B.Task(0, B.CONTAINS, _soa(8))
    with pytest.raises(ValueError):
        B.Task(0, B.DISTANCE_TO_IN, _soa(8))  # no directions
    with pytest.raises(ValueError):
        B.Task(0, 9, _soa(8))  # unknown query


def test_server_track_arrays_checked():
    p = np.zeros((3, 10))
    d = np.zeros((3, 10))
    arrs = S.Server._tracks(p, d)
    assert len(arrs) == 6 and all(a.size == 10 for a in arrs)
    with pytest.raises(ValueError):
        S.Server._tracks(p, np.zeros((3, 9)))
    with pytest.raises(ValueError):
        S.Server._tracks(p[:2], d)
