"""What the learned selector gains over the fixed rule, on the CPU.

ttv_amd/data/ holds the timing tables the exhaustive autotuner measured on an MI355X (every valid
variant of 20,000 training and 2,000 held-out shapes, HBM-cold, median of 3) and the network
trained on the GPU from the training table.  Here the library's own host-side selection
(ttvx_plan_info under TTVX_SELECT_LEARNED, no GPU needed) picks a variant for every held-out
shape, and its regret t(pick)/t(fastest) - 1 is compared with the fixed rule's.
The network itself is checked against the FP64 restatement in oracle/ (oracle.mlp_forward).
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DATA = os.path.join(REPO, "ttv_amd", "data")
TEST_TABLE = os.path.join(DATA, "selector_timings_test_gfx950.npz")
WEIGHTS = os.path.join(DATA, "selector_gfx950.npz")

pytestmark = pytest.mark.skipif(not (os.path.exists(TEST_TABLE) and os.path.exists(WEIGHTS)),
                                reason="selector weights / timing tables not generated yet")


def _regret(choice, times):
    best = np.nanmin(times, axis=1)
    return float(np.mean(times[np.arange(len(choice)), choice] / best - 1.0))


def test_learned_selector_beats_fixed_rule_on_held_out_shapes(ttvx):
    from ttv_amd import selector as S
    z = np.load(TEST_TABLE)
    times = z["times"]
    shapes = S.shapes_from_arrays(z)
    params = S.load_weights(WEIGHTS)
    ttvx.set_selector_weights(params)
    ttvx.set_selection(ttvx.SELECT_LEARNED)
    try:
        learned = np.array([ttvx.plan_info(s["q"], s["na"], s["wa"], s["pia"], S.DT_CODE[s["dtype"]])["variant"]
                            for s in shapes])
        ttvx.set_selection(ttvx.SELECT_HEURISTIC)
        fixed = np.array([ttvx.plan_info(s["q"], s["na"], s["wa"], s["pia"], S.DT_CODE[s["dtype"]])["variant"]
                          for s in shapes])
    finally:
        ttvx.restore_default_selection()
    assert np.all(np.isfinite(times[np.arange(len(learned)), learned])), "learned pick must be a timed variant"
    r_learned, r_fixed = _regret(learned, times), _regret(fixed, times)
    print("held-out regret: learned %.4f, fixed rule %.4f" % (r_learned, r_fixed))
    assert r_learned <= 0.05
    assert r_learned < r_fixed


def test_host_inference_matches_fp64_restatement(ttvx):
    from ttv_amd import selector as S
    z = np.load(TEST_TABLE)
    params = S.load_weights(WEIGHTS)
    feats = z["feats"][:300]
    y64, _ = O.mlp_forward(params, feats)
    valid = np.isfinite(z["times"][:300])
    pick64 = np.argmin(np.where(valid, y64, np.inf), axis=1)
    shapes = S.shapes_from_arrays(z)[:300]
    ttvx.set_selector_weights(params)
    ttvx.set_selection(ttvx.SELECT_LEARNED)
    try:
        host = np.array([ttvx.plan_info(s["q"], s["na"], s["wa"], s["pia"], S.DT_CODE[s["dtype"]])["variant"]
                         for s in shapes])
        feats_now = np.stack([ttvx.features(s["q"], s["na"], s["wa"], s["pia"], S.DT_CODE[s["dtype"]]) for s in shapes])
    finally:
        ttvx.restore_default_selection()
    assert np.array_equal(feats_now, feats)
    assert np.mean(host == pick64) >= 0.99


def test_library_loads_the_shipped_weights_by_default():
    """A fresh process selects with the MLP and the shipped weights (data/selector_gfx950.bin,
    found next to the library) without any call: same picks as with the weights set explicitly
    (the held-out shapes fold to no row of the tuning table)."""
    import subprocess
    import sys
    code = r'''
import sys, numpy as np
sys.path.insert(0, %r)
import ttv_amd
from ttv_amd import selector as S
z = np.load(%r)
shapes = S.shapes_from_arrays(z)[:300]
pick = lambda: [ttv_amd.plan_info(s["q"], s["na"], s["wa"], s["pia"], S.DT_CODE[s["dtype"]])["variant"] for s in shapes]
default = pick()
ttv_amd.set_selector_weights(S.load_weights()); ttv_amd.set_selection(ttv_amd.SELECT_LEARNED)
learned = pick()
ttv_amd.set_selection(ttv_amd.SELECT_HEURISTIC)
fixed = pick()
print(int(default == learned), int(default != fixed))
''' % (REPO, TEST_TABLE)
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-2000:]
    same, differs = out.stdout.split()[-2:]
    assert same == "1", "default picks differ from the shipped weights' picks"
    assert differs == "1", "default picks equal the fixed rule's: the weights were not loaded"


def test_tuning_table_rows_are_the_measured_fastest():
    """The shipped tuning table holds, for the training shapes, the fastest variant the exhaustive
    autotuner measured (pooled over shapes that fold to the same problem), and the default
    selection returns it for those shapes."""
    import ttv_amd
    from ttv_amd import selector as S
    rows = S.load_table()
    assert rows is not None and rows.shape[1] == 5
    table = {tuple(int(x) for x in r[:4]): int(r[4]) for r in rows}
    z = np.load(os.path.join(DATA, "selector_timings_train_gfx950.npz"))
    shapes = S.shapes_from_arrays(z)
    times = z["times"]
    pick = np.random.default_rng(0).choice(len(shapes), 200, replace=False)
    ttv_amd.restore_default_selection()
    slow = []
    for n in pick:
        s = shapes[n]
        code = S.DT_CODE[s["dtype"]]
        info = ttv_amd.plan_info(s["q"], s["na"], s["wa"], s["pia"], code)
        key = (code, info["O"], info["K"], info["I"])
        assert key in table, key
        assert info["variant"] == table[key]
        t = times[n]
        slow.append(t[table[key]] / np.nanmin(t))
    assert np.mean(slow) <= 1.01 and max(slow) <= 1.25
