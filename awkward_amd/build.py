"""Builds libawkward_hip.so (gfx950) from awkward_amd/csrc/*.hip with hipcc.

    python -m awkward_amd.build            # or awkward_amd.build.build()
    python -m awkward_amd.build sanitize   # host code under ASan + UBSan

Objects go to awkward_amd/_build/, the library to awkward_amd/lib/.  Every
source compiles in parallel; the whole build takes about a minute.  The
sanitized build instruments the host code only (each sanitizer flag after
-Xarch_host: device code is compiled as usual) into
awkward_amd/lib/san/libawkward_hip.so; load it with AWKWARD_HIP_LIBRARY and
clang's ASan runtime (SAN_RUNTIME) preloaded.
"""

from __future__ import annotations

import concurrent.futures
import glob
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJDIR = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIBPATH = os.path.join(LIBDIR, "libawkward_hip.so")
ARCH = os.environ.get("AWKWARD_HIP_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    # packed-FP32 instructions are kept out of every kernel
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


#: host-only sanitizers for the sanitized build
SAN_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined", "-Xarch_host",
             "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer"]
SAN_OBJDIR = os.path.join(HERE, "_build_san")
SAN_LIBPATH = os.path.join(LIBDIR, "san", "libawkward_hip.so")


def san_runtime() -> str:
    """The ASan runtime of hipcc's host compiler (the process must preload it)."""
    found = sorted(glob.glob(os.path.join(os.path.dirname(os.path.dirname(HIPCC)), "llvm", "lib", "clang", "*",
                                          "lib", "linux", "libclang_rt.asan-x86_64.so")))
    if not found:
        raise FileNotFoundError("clang's ASan runtime is not installed")
    return found[-1]


def _embed(objdir: str) -> None:
    """jit_embed.h: the headers of the kernels compiled at run time (jit.hip),
    as string constants -- expr.h without its <cmath> include (the run-time
    compiler has the device math built in), pair_jit.h and tuples.h."""
    texts = {}
    with open(os.path.join(CSRC, "expr.h")) as f:
        texts["kJitExprH"] = f.read().replace("#include <cmath>\n", "")
    with open(os.path.join(CSRC, "pair_jit.h")) as f:
        texts["kJitPairH"] = f.read()
    with open(os.path.join(CSRC, "tuples.h")) as f:
        texts["kJitTuplesH"] = f.read()
    delim = "AKHIP_EMBED"
    out = "// generated by awkward_amd/build.py from csrc/expr.h, pair_jit.h and tuples.h\n#pragma once\n"
    for name, text in texts.items():
        if ")" + delim in text:
            raise ValueError(f"{name} holds the raw-string delimiter")
        out += f'static const char {name}[] = R"{delim}({text}){delim}";\n'
    path = os.path.join(objdir, "jit_embed.h")
    if not os.path.exists(path) or open(path).read() != out:
        with open(path, "w") as f:
            f.write(out)


def _compile(src: str, objdir: str = OBJDIR, extra=()) -> str:
    obj = os.path.join(objdir, os.path.basename(src).replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, *extra, f"-I{objdir}", "-c", src, "-o", obj]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{proc.stderr}")
    return obj


def build(verbose: bool = True, sanitize: bool = False) -> str:
    objdir, libpath = (SAN_OBJDIR, SAN_LIBPATH) if sanitize else (OBJDIR, LIBPATH)
    extra = SAN_FLAGS if sanitize else ()
    os.makedirs(objdir, exist_ok=True)
    os.makedirs(os.path.dirname(libpath), exist_ok=True)
    sources = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    _embed(objdir)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(sources))) as pool:
        objects = list(pool.map(lambda src: _compile(src, objdir, extra), sources))
    tmp = libpath + ".tmp"
    # the sanitized library takes its runtime from the preloaded ASan runtime
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objects, "-lhiprtc", "-o", tmp]
    if sanitize:
        cmd[3:3] = ["-Wl,--allow-shlib-undefined", "-Wl,--unresolved-symbols=ignore-all"]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{proc.stderr}")
    os.replace(tmp, libpath)
    if not sanitize:
        jit_selftest(libpath)
    if verbose:
        print(f"built {libpath} from {len(sources)} sources for {ARCH}")
    return libpath


def jit_selftest(libpath: str = LIBPATH) -> None:
    """Compiles the run-time pair kernel of representative programs for ARCH
    through the built library's hipRTC path (no device needed), so that a
    template that no longer compiles fails the build rather than sending
    every call to the interpreter."""
    import ctypes

    lib = ctypes.CDLL(libpath)
    lib.awkward_hip_jit_selftest.argtypes = [ctypes.c_char_p]
    lib.awkward_hip_jit_selftest.restype = ctypes.c_int
    failed = lib.awkward_hip_jit_selftest(ARCH.encode())
    if failed:
        raise RuntimeError(f"{failed} run-time pair kernel(s) failed to compile for {ARCH} (log above)")


if __name__ == "__main__":
    import sys

    build(sanitize=len(sys.argv) > 1 and sys.argv[1] == "sanitize")
