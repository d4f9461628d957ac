"""The oracle's golden-vector and property suites re-run against a build of
oracle.cpp under AddressSanitizer and UndefinedBehaviorSanitizer (any
undefined behaviour aborts), in a subprocess with the ASan runtime
preloaded: out-of-bounds reads or writes, use of freed memory, signed
overflow, misaligned or null accesses in the CPU kernels fail this test."""

import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _runtime():
    if os.environ.get("LD_PRELOAD"):
        pytest.skip("a preload is already set: the sanitized run needs the ASan runtime first, not instead")
    try:
        path = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True,
                              check=True).stdout.strip()
    except (OSError, subprocess.CalledProcessError):
        return None
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.mark.timeout(600)
def test_oracle_suites_under_asan_and_ubsan():
    runtime = _runtime()
    if runtime is None:
        pytest.skip("no ASan runtime for the host compiler")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    lib = os.path.join(ROOT, "oracle", "_san", "liboracle.so")
    env = dict(os.environ, LD_PRELOAD=runtime, ORACLE_LIB=lib,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu",
                        os.path.join(HERE, "test_oracle_golden.py"), os.path.join(HERE, "test_oracle_props.py")],
                       cwd=ROOT, env=env, capture_output=True, text=True)
    tail = (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0, tail
    assert "passed" in r.stdout and "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, tail


def test_the_sanitized_run_catches_a_planted_overflow(tmp_path):
    """Control: a one-line heap overflow built with the same flags and
    loaded the same way is reported, so a clean run above means something."""
    runtime = _runtime()
    if runtime is None:
        pytest.skip("no ASan runtime for the host compiler")
    src = tmp_path / "planted.cpp"
    src.write_text('#include <cstdlib>\nextern "C" int planted(int n) { int* p = (int*)std::malloc(n * sizeof(int));'
                   ' p[n] = 1; int v = p[0]; std::free(p); return v; }\n')
    lib = tmp_path / "libplanted.so"
    subprocess.run(["g++", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fPIC",
                    "-shared", "-o", str(lib), str(src)], check=True)
    env = dict(os.environ, LD_PRELOAD=runtime, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([sys.executable, "-c", f"import ctypes; ctypes.CDLL({str(lib)!r}).planted(4)"], env=env,
                       capture_output=True, text=True)
    assert r.returncode != 0 and "heap-buffer-overflow" in r.stderr
