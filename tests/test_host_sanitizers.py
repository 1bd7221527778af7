"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only, g++).

* tests/cpp/test_plan_fuzz.cpp with ttv_amd/csrc/plan.cpp: the planner's validation and folding
  against brute-force enumeration over random and malformed problems.
* tests/cpp/test_metadata.cpp: the drop-in metadata headers (include/tlib/detail) on the golden
  tables of the reference's test suite.

Any out-of-bounds access, use after free, leak, signed overflow or misaligned access in these
paths aborts the program (-fno-sanitize-recover=all) and fails the test."""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OUT = os.path.join(REPO, "ttv_amd", "lib", "san")
FLAGS = ["-std=c++17", "-O1", "-g", "-Wall", "-Wextra", "-fsanitize=address,undefined",
         "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]


def _build(name, sources):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    os.makedirs(OUT, exist_ok=True)
    exe = os.path.join(OUT, name)
    subprocess.run(["g++"] + FLAGS + ["-I" + os.path.join(REPO, "include")] + sources + ["-o", exe], check=True)
    return exe


def _run(exe, args=()):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    return subprocess.run([exe] + list(args), capture_output=True, text=True, env=env, timeout=300)


def test_planner_fuzz_under_sanitizers():
    exe = _build("plan_fuzz", [os.path.join(REPO, "tests", "cpp", "test_plan_fuzz.cpp"),
                               os.path.join(REPO, "ttv_amd", "csrc", "plan.cpp")])
    r = _run(exe, ["6000"])
    assert r.returncode == 0, r.stdout + r.stderr[-4000:]
    assert " 0 failures" in r.stdout


def test_metadata_headers_under_sanitizers():
    exe = _build("metadata", [os.path.join(REPO, "tests", "cpp", "test_metadata.cpp")])
    r = _run(exe)
    assert r.returncode == 0, r.stdout + r.stderr[-4000:]
    assert "0 failures" in r.stdout
