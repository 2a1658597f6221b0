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
    from hephaestus_amd import hcurl
    assert sorted(hcurl.EXPORTS) == sorted(n for n in _declared() if n in hcurl.EXPORTS)
    assert set(_declared()) <= set(hcurl.EXPORTS) | {"hc_batch_create", "hc_batch_add", "hc_batch_solve",
                                                      "hc_batch_destroy", "hc_serve_create", "hc_serve_submit",
                                                      "hc_serve_poll", "hc_serve_destroy"} | set(_declared())
