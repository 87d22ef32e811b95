"""splatrs_amd: the MI355X (gfx950) differentiable splat training path of splatrs.

The compute lives in libsplat.so (hand-written HIP kernels behind the C ABI of
include/splat.h).  This package holds the host-side plumbing around it: record layouts
(`records`), cameras and synthetic scenes (`scenes`) and a ctypes mirror of the
reference API (`abi`).
"""
from . import records, scenes  # noqa: F401
