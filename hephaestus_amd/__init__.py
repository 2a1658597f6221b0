"""MI355X-native H(curl) curl-curl + sigma-mass operator path of hephaestus.

The compute lives in libhcurl.so (hand-written HIP for gfx950, C ABI in
include/hc_hcurl.h); this package holds the build script and a thin ctypes
binding used by the tests and the benchmark.  The C++ host layer that mirrors
hephaestus' Kernel / EquationSystem API sits in hephaestus_amd/host/.
"""
from .hcurl import (FACE_ALL, PREC_FP32, PREC_FP64, SOLVE_AUX, SOLVE_FP32_IR, SOLVE_FP64, SOLVE_MG, Batch, HcError,  # noqa: F401
                    Operator, Server, load, tile_list)
