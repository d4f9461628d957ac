"""CPU oracle for the hip backend -- TEST INFRASTRUCTURE ONLY.

``liboracle.so`` (built from ``oracle.cpp`` by ``oracle/Makefile``) restates
the awkward-cpp CPU kernels of the jagged-array path on host memory, with the
same symbol names, argument order and ``struct Error`` convention.  Only the
test suite, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
use it, and only as the checker or the timed CPU baseline; the product
package ``awkward_amd`` never imports it.

Parity pinning (what fixes each function of the oracle):

* every family with vectors in the reference's ``kernel-test-data.json``
  (``tests/golden/kernel_vectors.json``) is run against all of them, for
  every specialization whose types can hold the vector's values;
* reductions additionally against numpy closed forms (``np.add.reduceat``
  style sums in exact integer arithmetic, first-occurrence ``argmax`` on
  NaN-free data, identities for empty lists) and ``math.fsum`` bounds;
* ``combinations`` against ``itertools.combinations`` /
  ``combinations_with_replacement`` and ``math.comb`` for the lengths;
* ``next_range`` against Python's ``range(*slice.indices(n))``;
* ``argsort`` (no reference vectors: its spec sets ``automatic-tests:
  false``) against ``numpy.lexsort`` with a NaN-first key and the literal
  expectations of ``tests/test_0945_argsort_sort_nan_array.py:12-25``.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
#: ORACLE_LIB selects another build of the same kernels (the sanitized one)
LIBPATH = os.environ.get("ORACLE_LIB") or os.path.join(HERE, "liboracle.so")


class ERROR(ctypes.Structure):
    _fields_ = [
        ("str", ctypes.c_char_p),
        ("filename", ctypes.c_char_p),
        ("id", ctypes.c_int64),
        ("attempt", ctypes.c_int64),
    ]


_SCALARS = {
    "bool": ctypes.c_bool,
    "int8_t": ctypes.c_int8,
    "uint8_t": ctypes.c_uint8,
    "int16_t": ctypes.c_int16,
    "uint16_t": ctypes.c_uint16,
    "int32_t": ctypes.c_int32,
    "uint32_t": ctypes.c_uint32,
    "int64_t": ctypes.c_int64,
    "uint64_t": ctypes.c_uint64,
    "float": ctypes.c_float,
    "double": ctypes.c_double,
}


def build() -> None:
    """Compile liboracle.so from oracle.cpp (g++ only, seconds)."""
    subprocess.run(["make", "-s", "-B", "-C", HERE], check=True)


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            build()
        _lib = ctypes.CDLL(LIBPATH)
    return _lib


def argtype(spec_type: str):
    """ctypes type of one argument (pointers travel as raw addresses)."""
    if "List[" in spec_type:
        return ctypes.c_void_p
    return _SCALARS[spec_type]


def kernel(name: str):
    """The oracle function for specialization ``name`` with argtypes set."""
    from awkward_amd._kernel_signatures import SPECIALIZATIONS

    _, args, _ = SPECIALIZATIONS[name]
    f = getattr(lib(), name)
    f.argtypes = [argtype(t) for _, t, _ in args]
    f.restype = ERROR
    return f
