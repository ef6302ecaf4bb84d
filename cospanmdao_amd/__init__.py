"""cospanmdao_amd: an MI355X-native (gfx950) execution backend for CoSpan's fused-run path.

The product is native: libcospan_hip.so (the C boundary of include/cospan_hip.h: runtime, device scan, column build,
initializers) and libcospan_workloads.so (the GEMVER and NS-residual models built through the header-only C++ API under
cospanmdao_amd/include/naga/cospan). This package only loads them (ctypes) for tests, the benchmark and Python callers;
it never computes anything itself, and it raises if the libraries are missing.
"""
from ._lib import hip, workloads, LIB_DIR, check, CsError  # noqa: F401

__all__ = ["hip", "workloads", "LIB_DIR", "check", "CsError"]
