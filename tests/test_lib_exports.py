"""The C-ABI library loads without a GPU and exports every declared symbol."""

import ctypes
import os
import re

import pytest

from awkward_amd._kernel_signatures import SPECIALIZATIONS
from awkward_amd._lib import LIBPATH

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header="awkward_hip.h"):
    with open(os.path.join(ROOT, "include", header)) as f:
        text = f.read()
    return re.findall(r"^\S.*?\b(awkward_\w+)\(", text, flags=re.M)


def test_header_declares_every_specialization():
    names = set(declared())
    missing = [n for n in SPECIALIZATIONS if n not in names]
    assert not missing, missing
    # 163 on the jagged path, 38 for prod / any / all (and max / min of booleans), 22 option-type, 4 non-local
    assert len(SPECIALIZATIONS) == 227


@pytest.mark.skipif(not os.path.exists(LIBPATH), reason="libawkward_hip.so not built")
@pytest.mark.parametrize("header", ["awkward_hip.h", "awkward_hip_ext.h"])
def test_library_exports_every_declared_symbol(header):
    lib = ctypes.CDLL(LIBPATH)
    names = declared(header)
    assert names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIBPATH), reason="libawkward_hip.so not built")
def test_every_export_is_declared():
    """Nothing the library exports under awkward_ is missing from the headers."""
    import subprocess

    out = subprocess.run(["nm", "-D", "--defined-only", LIBPATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.split() and line.split()[-1].startswith("awkward_")}
    headers = set(declared("awkward_hip.h")) | set(declared("awkward_hip_ext.h"))
    assert exported - headers == set()


def test_product_does_not_import_oracle():
    # the hip package never loads the oracle (test infrastructure only)
    pkg = os.path.join(ROOT, "awkward_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".hip", ".h")):
                with open(os.path.join(dirpath, fn)) as f:
                    text = f.read()
                assert "import oracle" not in text and "from oracle" not in text, fn
                assert "liboracle" not in text, fn


def test_run_time_pair_kernels_compile():
    """The pair kernels the library compiles at run time (hipRTC) build for
    gfx950 from the embedded sources, without a device."""
    import ctypes

    from awkward_amd._lib import LIBPATH

    lib = ctypes.CDLL(LIBPATH)
    lib.awkward_hip_jit_selftest.argtypes = [ctypes.c_char_p]
    lib.awkward_hip_jit_selftest.restype = ctypes.c_int
    assert lib.awkward_hip_jit_selftest(b"gfx950") == 0
