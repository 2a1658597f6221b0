"""libhcurl.so loads on a CPU-only host and exports every entry point that
include/*.h declares (no compute calls: this runs without a GPU)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "hephaestus_amd", "libhcurl.so")


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*\b(hc_[a-z0-9_]+)\s*\(", txt, re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    d = _declared()
    assert "hc_op_create" in d and "hc_pcg_solve" in d and "hc_op_apply" in d


@pytest.mark.skipif(not os.path.exists(LIB), reason="libhcurl.so not built")
def test_library_exports_declared_symbols():
    import torch  # noqa: F401  (same HIP runtime as the GPU runs)
    lib = ctypes.CDLL(LIB)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_lists_all_exports():
    """The ctypes binding declares exactly the header's entry points."""
    from hephaestus_amd import hcurl
    declared = set(_declared())
    listed = set(hcurl.EXPORTS)
    assert len(hcurl.EXPORTS) == len(listed), "duplicate names in hcurl.EXPORTS"
    assert declared - listed == set(), f"declared but not bound: {sorted(declared - listed)}"
    assert listed - declared == set(), f"bound but not declared: {sorted(listed - declared)}"
