"""rest_tensors_amd — MI355X-native RI hot path of rest_tensors.

The compute lives in librest_tensors_amd.so (hand-written gfx950 HIP kernels
behind a C ABI, see include/restmatr.h and include/rest_tensors_amd.h).  This
package is the Python plumbing used by the tests and the benchmark: a ctypes
binding, device buffers and numpy-facing helpers.  There is no CPU fallback.
"""
from ._lib import LIB_PATH, RtError, check, lib  # noqa: F401

__all__ = ["lib", "check", "RtError", "LIB_PATH"]
