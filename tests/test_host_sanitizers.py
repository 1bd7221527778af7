"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

* tests/cpp/test_plan_fuzz.cpp with ttv_amd/csrc/plan.cpp (g++): the planner's validation and
  folding against brute-force enumeration over random and malformed problems.
* tests/cpp/test_metadata.cpp (g++): the drop-in metadata headers (include/tlib/detail) on the
  golden tables of the reference's test suite.
* tests/cpp/test_select_host.cpp against the whole library built by hipcc with the sanitizers on
  its host code only (device code unoptimised, never run): planning, features, the selector's
  host inference, tuning table and decision cache through the C-ABI, in every selection mode.

Any out-of-bounds access, use after free, leak, signed overflow or misaligned access in these
paths aborts the program (-fno-sanitize-recover=all) and fails the test."""
import concurrent.futures as cf
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


HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SAN_HOST = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-sanitize-recover=all"]


def _sanitized_library():
    """libttvx_san.so: every library source, sanitizers on the host side (rebuilt when a source or
    header is newer than the library)."""
    from ttv_amd import build as B
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    os.makedirs(os.path.join(OUT, "obj"), exist_ok=True)
    lib = os.path.join(OUT, "libttvx_san.so")
    srcs = B.SOURCES + [s for s in B.OPTIONAL if os.path.exists(os.path.join(B.CSRC, s))]
    deps = [os.path.join(B.CSRC, f) for f in os.listdir(B.CSRC)] + [os.path.join(REPO, "include", "ttvx.h")]
    if os.path.exists(lib) and os.path.getmtime(lib) > max(os.path.getmtime(d) for d in deps):
        return lib
    flags = ["-g", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-I" + os.path.join(REPO, "include"),
             "-Xarch_host", "-O1", "-Xarch_device", "-O0", "-Xarch_host", "-fno-omit-frame-pointer"] + SAN_HOST

    def comp(src):
        obj = os.path.join(OUT, "obj", src + ".o")
        r = subprocess.run([HIPCC] + flags + ["-x", "hip", "-c", os.path.join(B.CSRC, src), "-o", obj],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]
        return obj

    with cf.ThreadPoolExecutor(max_workers=8) as ex:
        objs = list(ex.map(comp, srcs))
    subprocess.run([HIPCC, "-shared", "--offload-arch=gfx950"] + SAN_HOST + ["-o", lib] + objs, check=True)
    return lib


def test_planning_and_selection_under_sanitizers():
    lib = _sanitized_library()
    exe = os.path.join(OUT, "select_host")
    subprocess.run([HIPCC, "-O1", "-g", "-std=c++17"] + SAN_HOST +
                   ["-I" + os.path.join(REPO, "include"), os.path.join(REPO, "tests", "cpp", "test_select_host.cpp"),
                    "-L" + OUT, "-lttvx_san", "-Wl,-rpath," + OUT, "-o", exe], check=True, capture_output=True)
    assert os.path.exists(lib)
    r = _run(exe, ["1500"])
    assert r.returncode == 0, r.stdout + r.stderr[-4000:]
    assert " 0 failures" in r.stdout
