"""OpenZL codec hot path on MI355X (gfx950).

The codecs run in libozl_hip.so (HIP kernels behind a C ABI declared in
include/openzl_amd.h). This package only loads that library and offers thin
Python conveniences for tests and benchmarks; it holds no codec arithmetic.
"""
from ._lib import lib, OzlError  # noqa: F401
