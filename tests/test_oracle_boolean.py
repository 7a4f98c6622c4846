"""Boolean volumes in the oracle (oracle/vgo_boolean.cpp): against the live
reference build (its UnplacedBooleanVolume and SpecializedBooleanVolume
compiled from volumes/kernel/Boolean*Implementation.h) on random points,
bit for bit, and against closed forms that pin the operations themselves."""
import numpy as np
import pytest

import pyoracle as po
from vecgeom_amd import geometry as G
from vecgeom_amd import workloads as W

QUERIES = [("contains", po.Q_CONTAINS), ("inside", po.Q_INSIDE), ("dist_in", po.Q_DIST_IN),
           ("dist_out", po.Q_DIST_OUT), ("safe_in", po.Q_SAFE_IN), ("safe_out", po.Q_SAFE_OUT),
           ("placed_dist_out", po.Q_PLACED_DIST_OUT)]
I = G.Transformation3D()


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype.kind == "f":
        return (a == b) | (np.isnan(a) & np.isnan(b))
    return a == b


def run(h, q, inp):
    sm = np.full(inp["p"].shape[1], 1e30)
    if q in (po.Q_DIST_OUT, po.Q_SAFE_OUT):
        return h.query(q, inp["lp"], inp["ld"] if q == po.Q_DIST_OUT else None, sm if q == po.Q_DIST_OUT else None)
    if q in (po.Q_DIST_IN, po.Q_PLACED_DIST_OUT):
        return h.query(q, inp["p"], inp["d"], sm)
    return h.query(q, inp["p"])


@pytest.mark.parametrize("name", [n for n, _ in W.boolean_shapes()])
def test_boolean_oracle_against_live_reference(oracle, reference, name):
    u = dict(W.boolean_shapes())[name]
    rng = np.random.Generator(np.random.Philox(404))
    for xf in (I, W.random_placement(rng)):
        ho = oracle.placed(u.type, u.params, xf.array())
        hr = reference.placed(u.type, u.params, xf.array())
        inp = W.shape_inputs(((-60, -60, -60), (60, 60, 60)), xf, 20000, seed=505)
        for qn, q in QUERIES:
            a, b = run(ho, q, inp), run(hr, q, inp)
            ok = same(a, b)
            assert ok.all(), f"{name} {qn}: {int((~ok).sum())} lanes differ, first {np.flatnonzero(~ok)[:5]}"


def two_boxes(op, right_xf):
    return G.UnplacedBooleanVolume(op, G.UnplacedBox(10, 10, 10), I, G.UnplacedBox(5, 5, 5), right_xf)


def q1(oracle, u, q, p, d=None):
    h = oracle.placed(u.type, u.params, I.array())
    p = np.array(p, dtype=np.float64).reshape(3, 1)
    if d is None:
        return h.query(q, p)[0]
    return h.query(q, p, np.array(d, dtype=np.float64).reshape(3, 1), np.full(1, 1e30))[0]


def test_boolean_closed_forms(oracle):
    # Union of a box (half 10) and a disjoint box (half 5) centred at x = 20,
    # spanning [15, 25]: from x = -50 along +x it is entered at 40, from
    # x = 50 along -x at 25; from the origin along +x the walk leaves at the
    # first box's face, x = 10, since the second box does not hold 10 + 1e-9.
    u = two_boxes(G.kUnion, G.Transformation3D(20, 0, 0))
    assert q1(oracle, u, po.Q_CONTAINS, [0, 0, 0]) == 1 and q1(oracle, u, po.Q_CONTAINS, [20, 0, 0]) == 1
    assert q1(oracle, u, po.Q_CONTAINS, [12.5, 0, 0]) == 0
    assert q1(oracle, u, po.Q_DIST_IN, [-50, 0, 0], [1, 0, 0]) == 40.0
    assert q1(oracle, u, po.Q_DIST_IN, [50, 0, 0], [-1, 0, 0]) == 25.0
    assert abs(q1(oracle, u, po.Q_DIST_OUT, [0, 0, 0], [1, 0, 0]) - 10.0) < 1e-12
    # Overlapping union (second box at x = 12 spans [7, 17]): the walk from the
    # origin along +x crosses into the second box and leaves the union at 17.
    u = two_boxes(G.kUnion, G.Transformation3D(12, 0, 0))
    assert abs(q1(oracle, u, po.Q_DIST_OUT, [0, 0, 0], [1, 0, 0]) - 17.0) < 1e-8
    assert q1(oracle, u, po.Q_SAFE_IN, [30, 0, 0]) == 13.0
    # Intersection of the two (overlap x in [7, 10]): entered from x = -50
    # along +x at 57, from x = 50 along -x at 40; the point (8.5, 0, 0) is
    # inside with SafetyToOut 1.5 (the nearer x faces).
    u = two_boxes(G.kIntersection, G.Transformation3D(12, 0, 0))
    assert q1(oracle, u, po.Q_CONTAINS, [8.5, 0, 0]) == 1 and q1(oracle, u, po.Q_CONTAINS, [5, 0, 0]) == 0
    assert abs(q1(oracle, u, po.Q_DIST_IN, [-50, 0, 0], [1, 0, 0]) - 57.0) < 1e-8
    assert abs(q1(oracle, u, po.Q_DIST_IN, [50, 0, 0], [-1, 0, 0]) - 40.0) < 1e-8
    assert q1(oracle, u, po.Q_SAFE_OUT, [8.5, 0, 0]) == 1.5
    assert q1(oracle, u, po.Q_DIST_OUT, [8.5, 0, 0], [-1, 0, 0]) == 1.5
    # Subtraction of a centred box (half 5) from the box (half 10): a shell.
    # From the origin (in the hole) along +x the shell starts at 5; from
    # x = 7 along -x the hole is reached after 2, along +x the outside after 3.
    u = two_boxes(G.kSubtraction, I)
    assert q1(oracle, u, po.Q_CONTAINS, [0, 0, 0]) == 0 and q1(oracle, u, po.Q_CONTAINS, [7, 0, 0]) == 1
    assert abs(q1(oracle, u, po.Q_DIST_IN, [0, 0, 0], [1, 0, 0]) - 5.0) < 1e-7
    assert q1(oracle, u, po.Q_DIST_OUT, [7, 0, 0], [-1, 0, 0]) == 2.0
    assert q1(oracle, u, po.Q_DIST_OUT, [7, 0, 0], [1, 0, 0]) == 3.0
    assert q1(oracle, u, po.Q_SAFE_OUT, [7, 0, 0]) == 2.0
    assert q1(oracle, u, po.Q_INSIDE, [7, 0, 0]) == 0 and q1(oracle, u, po.Q_INSIDE, [0, 0, 0]) == 2


def test_boolean_operands_must_be_primitive():
    u = two_boxes(G.kUnion, I)
    with pytest.raises(ValueError):
        G.UnplacedBooleanVolume(G.kUnion, u, I, G.UnplacedBox(1, 1, 1), I)
    with pytest.raises(ValueError):
        G.UnplacedBooleanVolume(7, G.UnplacedBox(1, 1, 1), I, G.UnplacedBox(1, 1, 1), I)
