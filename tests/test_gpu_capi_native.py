"""The C-ABI called from C++ (tests/native/capi_gpu.cpp, linked against
vecgeom_amd/lib/libvecgeom_amd.so, no Python in between): basket and scalar
shape queries, Contains(point, local&), UnplacedContains, Normal and the
USolids DistanceToOut with the exit normal, the
no-allocation batch and the navigator on a small geometry with closed-form
answers."""
import os
import subprocess

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "tests", "native", "capi_gpu")


def test_cpp_caller_of_the_c_abi(gpu):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is missing: build it with __graft_entry__.build()")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all" in r.stdout and "checks passed" in r.stdout
