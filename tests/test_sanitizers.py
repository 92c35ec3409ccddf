"""The host regrid and table build (threaded row-pipelined balance passes, shared node pool,
parallel table phases) built with g++ under AddressSanitizer + UndefinedBehaviorSanitizer and,
separately, under ThreadSanitizer, and run on a mesh large enough for the parallel passes: no
sanitizer report, the threaded result equals the one-thread result, and the cached neighbours
equal the recursive finders' (tests/sanitize/regrid_check.cpp).
CPU only; the host sources are compiled directly, without the HIP files."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "fdm-cell-based-amr-cartesian_amd", "csrc")
DRIVER = os.path.join(ROOT, "tests", "sanitize", "regrid_check.cpp")


def _build_and_run(tmp_path, name, flags, env_extra):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / name)
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-pthread", *flags, "-I", CSRC,
           DRIVER, os.path.join(CSRC, "host_tree.cpp"), os.path.join(CSRC, "host_tables.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=600)
    env = dict(os.environ, **env_extra)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=900, env=env)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0, r.stderr[-4000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
    assert "WARNING: ThreadSanitizer" not in r.stderr
    assert r.stdout.count("(= one thread)") == 4


def test_regrid_under_address_and_undefined_sanitizers(tmp_path):
    _build_and_run(tmp_path, "regrid_asan", ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"],
                   {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0", "UBSAN_OPTIONS": "print_stacktrace=1"})


def test_regrid_under_thread_sanitizer(tmp_path):
    _build_and_run(tmp_path, "regrid_tsan", ["-fsanitize=thread"], {"TSAN_OPTIONS": "halt_on_error=1"})
