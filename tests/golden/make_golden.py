"""Writes the golden fixtures of tests/golden/ from the reference compiled from its own headers.

Run in a container that has the reference sources (make -C oracle ref first):
    python tests/golden/make_golden.py

Outputs (committed; data only):
  ttv_ref_fixtures.npz  -- seeded random A (U[-1,1]) and b for p = 2..4, every layout, every q,
                           FP32 and FP64, with C from the reference's default policy
                           par_loop/subtensor/all (policy id 4 of oracle/ref_shim.cpp) and from
                           seq/slice/none (id 8); C in the dense output layout.
  metadata_ref.json     -- output shape / output layout / dense strides / k-order layouts /
                           shape classes computed by the reference for exhaustive small inputs.
"""
import itertools
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402

SHAPES = {2: [3, 5], 3: [3, 4, 5], 4: [2, 3, 4, 5]}


def fixtures():
    R = O.ref()
    if R is None:
        raise SystemExit("oracle/_ref/libttvref.so missing: run `make -C oracle ref`")
    out = {}
    rng = np.random.default_rng(20240611)
    for p, na in SHAPES.items():
        for dt, tag in ((np.float64, "f64"), (np.float32, "f32")):
            for q in range(1, p + 1):
                out[f"{tag}_p{p}_q{q}_b"] = rng.uniform(-1, 1, na[q - 1]).astype(dt)
            for pia in itertools.permutations(range(1, p + 1)):
                pia = list(pia)
                wa = O.strides(na, pia)
                key = f"{tag}_p{p}_pi{''.join(map(str, pia))}"
                a = rng.uniform(-1, 1, O.span(na, wa)).astype(dt)
                out[key + "_A"] = a
                for q in range(1, p + 1):
                    b = out[f"{tag}_p{p}_q{q}_b"]
                    out[f"{key}_q{q}_C"] = O.ref_ttv(q, a, na, wa, pia, b, policy=4)
                    out[f"{key}_q{q}_Cseq"] = O.ref_ttv(q, a, na, wa, pia, b, policy=8)
    np.savez_compressed(os.path.join(HERE, "ttv_ref_fixtures.npz"), **out)
    print("wrote", len(out), "arrays")


def metadata():
    import ctypes
    R = O.ref()

    def arr(v):
        a = (ctypes.c_size_t * max(1, len(v)))(*v)
        return a

    rec = {"output_shape": [], "output_layout": [], "strides": [], "k_order": [], "shape_class": [],
           "inverse_mode": []}
    for p in range(1, 6):
        for pia in itertools.permutations(range(1, p + 1)):
            pia = list(pia)
            for q in range(1, p + 1):
                if p >= 2:
                    out = arr([0] * (p - 1))
                    if R.ref_output_layout(arr(pia), p, q, out) == 0:
                        rec["output_layout"].append([pia, q, list(out[:p - 1])])
                rec["inverse_mode"].append([pia, q, R.ref_inverse_mode(arr(pia), p, q)])
    for p in range(1, 5):
        for na in itertools.product([1, 2, 3], repeat=p):
            na = list(na)
            rec["shape_class"].append([na, R.ref_shape_class(arr(na), p)])
            for q in range(1, p + 1):
                if p >= 2:
                    out = arr([0] * (p - 1))
                    if R.ref_output_shape(arr(na), p, q, out) == 0:
                        rec["output_shape"].append([na, q, list(out[:p - 1])])
            for pia in itertools.permutations(range(1, p + 1)):
                pia = list(pia)
                out = arr([0] * p)
                if R.ref_strides(arr(na), arr(pia), p, out) == 0:
                    rec["strides"].append([na, pia, list(out[:p])])
    for p in range(1, 11):
        for k in range(0, p + 2):
            out = arr([0] * p)
            R.ref_k_order_layout(p, k, out)
            rec["k_order"].append([p, k, list(out[:p])])
    with open(os.path.join(HERE, "metadata_ref.json"), "w") as f:
        json.dump(rec, f, separators=(",", ":"))
    print("wrote metadata:", {k: len(v) for k, v in rec.items()})


if __name__ == "__main__":
    fixtures()
    metadata()
