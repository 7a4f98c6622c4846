import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def golden_cases(prefix="shape_"):
    return sorted(f for f in os.listdir(GOLDEN) if f.startswith(prefix) and f.endswith(".npz"))


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def same(a, b):
    """Bit-level agreement for outputs: equal values, inf == inf, NaN == NaN."""
    a = np.asarray(a)
    b = np.asarray(b)
    if a.dtype.kind == "f":
        return (a == b) | (np.isnan(a) & np.isnan(b))
    return a == b


def close(a, b):
    """The FP64 distance bar: |a - b| <= max(1e-9 mm, 1e-12 |b|), or equal
    (inf == inf, NaN == NaN)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    eq = (a == b) | (np.isnan(a) & np.isnan(b))
    with np.errstate(invalid="ignore"):
        return eq | (np.abs(a - b) <= np.maximum(1e-9, 1e-12 * np.abs(b)))


# Distances and safeties that pass through per-point library transcendentals
# (the Sphere theta safeties: asin, sin; the Sphere phi exit: atan2), where
# device and host math libraries may differ in the last ulp, and the Torus
# distances: held to the FP64 distance bar.  Classifications (Contains,
# Inside) and every other query must agree bit for bit.
TOLERANT = {"Torus": {"dist_in", "dist_out", "placed_dist_out"},
            "Sphere": {"dist_out", "placed_dist_out", "safe_in", "safe_out"}}


def agree(shape, query, got, want):
    if query in TOLERANT.get(shape, ()):
        return close(got, want)
    return same(got, want)


def local_rays(xf, p, d, local):
    """The local-frame rays of a query's inputs: (p, d) themselves for the
    local queries, their Transform under the placement xf otherwise."""
    if local:
        return p, d
    from vecgeom_amd import workloads
    return workloads.master_to_local(xf, p), workloads.master_to_local(xf, d, direction=True)


def assert_agree(shape, query, got, want, what="", torus=None):
    """agree() on every lane.  The device's Torus distances come from an
    exact real-root solver where the reference takes a closed form whose
    acceptance test drops or misplaces roots on near-tangent rays: a Torus
    lane beyond the bar passes only if the device's value is within the
    surface tolerance of the exact first crossing (pyoracle.torus_distance_ok),
    for which torus = (params, lp, ld) gives the torus and the local rays."""
    import pyoracle
    ok = agree(shape, query, got, want)
    if shape == "Torus" and query in TOLERANT["Torus"] and not ok.all():
        assert torus is not None, f"{what} Torus {query}: lanes beyond the bar and no rays to judge them"
        params, lp, ld = torus
        bad = np.flatnonzero(~ok)
        ok = ok.copy()
        ok[bad] = pyoracle.torus_distance_ok(params, np.asarray(lp)[:, bad], np.asarray(ld)[:, bad],
                                             np.asarray(got)[bad], query != "dist_in")
    bad = int((~ok).sum())
    assert bad == 0, f"{what} {shape} {query}: {bad} of {ok.size} lanes differ, first {np.flatnonzero(~ok)[:5]}"


@pytest.fixture(scope="session")
def oracle():
    import pyoracle
    return pyoracle.oracle()


@pytest.fixture(scope="session")
def reference():
    import pyoracle
    ref = pyoracle.reference()
    if ref is None:
        pytest.skip("reference build (oracle/_ref) not present")
    return ref


@pytest.fixture(scope="session")
def gpu():
    """Initialises the HIP library on cuda:0; fails loudly if it cannot."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected on a machine without a GPU")
    from vecgeom_amd import queries
    queries.init(0)
    return torch.device("cuda:0")
