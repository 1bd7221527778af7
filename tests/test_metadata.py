"""Metadata layer of the drop-in C++ headers (include/tlib/detail) on the CPU.

* tests/cpp/test_metadata.cpp holds the golden tables of the reference's own test suite
  (gtest_tlib_{shape,layout,strides,workload}.cpp) and their invariants.
* The dump mode of the same program is compared with tests/golden/metadata_ref.json, written
  by the compiled reference for exhaustive small inputs (p <= 5 layouts, extents in {1,2,3}).
"""
import json
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "ttv_amd", "lib", "test_metadata")


@pytest.fixture(scope="module")
def meta_bin():
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    src = os.path.join(REPO, "tests", "cpp", "test_metadata.cpp")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-I" + os.path.join(REPO, "include"), src, "-o", BIN],
                   check=True)
    return BIN


def test_reference_golden_tables(meta_bin):
    r = subprocess.run([meta_bin], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout


def _dump(meta_bin, lines):
    r = subprocess.run([meta_bin, "dump"], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True)
    return r.stdout.split("\n")[:len(lines)]


def test_against_compiled_reference(meta_bin):
    with open(os.path.join(REPO, "tests", "golden", "metadata_ref.json")) as f:
        ref = json.load(f)
    queries, expect = [], []
    for na, q, out in ref["output_shape"]:
        queries.append("oshape %d %s %d" % (len(na), " ".join(map(str, na)), q))
        expect.append(" ".join(map(str, out)))
    for pia, q, out in ref["output_layout"]:
        queries.append("olayout %d %s %d" % (len(pia), " ".join(map(str, pia)), q))
        expect.append(" ".join(map(str, out)))
    for na, pia, out in ref["strides"]:
        queries.append("strides %d %s %s" % (len(na), " ".join(map(str, na)), " ".join(map(str, pia))))
        expect.append(" ".join(map(str, out)))
    for p, k, out in ref["k_order"]:
        queries.append("korder %d %d" % (p, k))
        expect.append(" ".join(map(str, out)))
    for na, m in ref["shape_class"]:
        queries.append("sclass %d %s" % (len(na), " ".join(map(str, na))))
        expect.append(str(m))
    for pia, q, r in ref["inverse_mode"]:
        queries.append("invmode %d %s %d" % (len(pia), " ".join(map(str, pia)), q))
        expect.append(str(r))
    got = _dump(meta_bin, queries)
    bad = [(qq, e, g.strip()) for qq, e, g in zip(queries, expect, got) if g.strip() != e.strip()]
    assert not bad, bad[:10]
    assert len(queries) > 4000
