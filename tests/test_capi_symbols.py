"""The C-ABI library loads and exports every entry point declared in
include/*.h (no compute calls: runs without a GPU)."""
import ctypes
import glob
import os
import re

from conftest import ROOT

LIB = os.path.join(ROOT, "vecgeom_amd", "lib", "libvecgeom_amd.so")


def declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names.update(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(vg_\w+)\s*\(", src, flags=re.M))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared()
    assert "vg_init" in names and "vg_contains" in names
    assert len(names) >= 20


def test_library_exports_all_declared_symbols():
    assert os.path.exists(LIB), "libvecgeom_amd.so not built"
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_python_binding_covers_header():
    from vecgeom_amd import capi
    import vecgeom_amd.batch  # noqa: F401  (registers task and batch signatures)
    import vecgeom_amd.nav  # noqa: F401  (registers navigation signatures)
    import vecgeom_amd.serve  # noqa: F401  (registers server signatures)
    missing = [n for n in declared() if n not in capi._SIGS]
    assert not missing, f"no ctypes signature for: {missing}"
