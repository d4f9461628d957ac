"""Loads libawkward_hip.so, the C-ABI library of hand-written gfx950 kernels.

There is no fallback: if the library has not been built, or no HIP device is
visible, every kernel call raises.  Build it with ``python -m awkward_amd.build``.
"""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# AWKWARD_HIP_LIBRARY names another build of the same library (A/B timing of
# kernel variants); the in-tree build is the default.
LIBPATH = os.environ.get("AWKWARD_HIP_LIBRARY") or os.path.join(HERE, "lib", "libawkward_hip.so")

_lib = None


class LibraryMissing(RuntimeError):
    pass


def _warn_if_stale() -> None:
    import glob
    import sys

    built = os.path.getmtime(LIBPATH)
    sources = glob.glob(os.path.join(HERE, "csrc", "*"))
    if any(os.path.getmtime(s) > built + 1 for s in sources):
        print(f"warning: {LIBPATH} is older than its sources; rebuild with "
              "`python -m awkward_amd.build`", file=sys.stderr)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise LibraryMissing(
                f"{LIBPATH} is not built; run `python -m awkward_amd.build` "
                "(the hip backend has no CPU fallback)"
            )
        if "AWKWARD_HIP_LIBRARY" not in os.environ:
            _warn_if_stale()
        handle = ctypes.CDLL(LIBPATH)
        handle.awkward_hip_set_stream.argtypes = [ctypes.c_void_p]
        handle.awkward_hip_set_stream.restype = None
        handle.awkward_hip_set_deferred_errors.argtypes = [ctypes.c_int]
        handle.awkward_hip_set_deferred_errors.restype = None
        handle.awkward_hip_version.restype = ctypes.c_char_p
        handle.awkward_hip_reserve_scratch.argtypes = [ctypes.c_int64]
        handle.awkward_hip_reserve_scratch.restype = ctypes.c_int
        _lib = handle
    return _lib
