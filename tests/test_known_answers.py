"""Known answers of the reference's unit tests (TestBox.cpp, TestTube.cpp, TestPolycone.cpp,
TestPolyhedra.cpp, TestOrb.cpp, TestTrd.cpp, TestTrap.cpp,
TestCons.cpp, TestSphere.cpp, TestVecGeomPolycone.cpp, TestBoxlikeTrapezoid.cpp),
transcribed as data in tests/golden/known_answers.json, checked on the
oracle here and on the GPU in test_gpu_shapes.py."""
import json
import os

import numpy as np
import pytest

import pyoracle as po
from conftest import GOLDEN

QMAP = {"inside": po.Q_INSIDE, "dist_in": po.Q_DIST_IN, "dist_out": po.Q_DIST_OUT, "safe_in": po.Q_SAFE_IN,
        "safe_out": po.Q_SAFE_OUT, "contains": po.Q_CONTAINS}
IDENT = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1], np.float64)


def rows():
    with open(os.path.join(GOLDEN, "known_answers.json")) as f:
        return json.load(f)["rows"]


def approx_equal(x, y):
    """test/unit_tests/ApproxEqual.h:17-29"""
    if x == y:
        return True
    if x * y == 0.0:
        return abs(x - y) < 1e-6
    return abs(x - y) / (abs(x) + abs(y)) < 1e-6


def check(row, got):
    e = row["expect"]
    if row["cmp"] == "exact":
        return got == e
    if row["cmp"] == "inf":
        return got >= 1e30 or np.isinf(got)
    if row["cmp"] == "le0":
        return got <= 0.0
    if row["cmp"] == "abs":
        return abs(float(got) - float(e)) < row["tol"]
    return approx_equal(float(got), float(e))


def evaluate(lib, row):
    h = lib.placed(row["type"], row["params"], IDENT)
    p = np.array(row["p"], np.float64).reshape(3, 1)
    d = None if row["d"] is None else np.array(row["d"], np.float64).reshape(3, 1)
    return h.query(QMAP[row["query"]], p, d, None)[0]


@pytest.mark.parametrize("i", range(len(rows())))
def test_known_answer_oracle(oracle, i):
    row = rows()[i]
    got = evaluate(oracle, row)
    assert check(row, got), f"{row['src']}: {row['query']} p={row['p']} d={row['d']} got {got} want {row['expect']}"
