"""The C-ABI library on the CPU (no compute calls): it loads, exports every symbol that
include/*.h declares, rejects invalid arguments with the reference's messages in the reference's
order (include/tlib/ttv.h:64-89, tensor_times_vector.h:147-168), and folds every layout to the
right [O][K][I] view."""
import ctypes
import itertools

import numpy as np
import pytest

from oracle import oracle as O


def test_exports_every_declared_symbol(ttvx):
    from ttv_amd import _lib
    names = _lib.declared_symbols()
    assert len(names) >= 20
    L = ttvx.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert L.ttvx_version() >= 10000
    assert L.ttvx_num_variants() == 15
    assert L.ttvx_selector_num_params() == 6783


def _call(ttvx, q, p, a, na, wa, pia, b, nb, c, nc, wc, pic):
    from ttv_amd._lib import sizes
    L = ttvx.lib()
    arr = lambda v: None if v is None else sizes(v)  # noqa: E731
    ptr = lambda x: ctypes.c_void_p(x)  # noqa: E731
    rc = L.ttvx_ttv_f64(5, 1, 2, q, p, ptr(a), arr(na), arr(wa), arr(pia), ptr(b), arr(nb), ptr(c), arr(nc),
                        arr(wc), arr(pic))
    return rc, L.ttvx_last_error().decode()


def _ref_msg(q, p, a, na, wa, pia, b, nb, c, nc, wc, pic):
    """Message of the compiled reference for the same call (None if it would not throw)."""
    R = O.ref()
    arr = lambda v: None if v is None else O._arr(v)  # noqa: E731
    ptr = lambda x: ctypes.c_void_p(x)  # noqa: E731
    rc = R.ref_ttv_f64(4, q, p, ptr(a), arr(na), arr(wa), arr(pia), ptr(b), arr(nb), ptr(c), arr(nc), arr(wc),
                       arr(pic))
    return R.ref_last_error().decode() if rc else None


def _invalid_cases():
    good = dict(q=2, p=3, a=8, na=[2, 3, 4], wa=[1, 2, 6], pia=[1, 2, 3], b=8, nb=[3], c=8, nc=[2, 4], wc=[1, 2],
                pic=[1, 2])
    cases = [
        dict(p=0), dict(q=0), dict(q=4), dict(a=0), dict(b=0), dict(c=0), dict(na=None), dict(nb=None),
        dict(nc=None), dict(wa=None), dict(wc=None), dict(pia=None), dict(pic=None), dict(nb=[4]),
        dict(na=[2, 3, 0], nb=[3]), dict(nc=[2, 0]), dict(pia=[1, 1, 3]), dict(pia=[0, 2, 3]), dict(pic=[2, 2]),
        dict(wa=[6, 2, 1]), dict(wc=[2, 1]),
        dict(pia=[2, 1, 3], wa=[3, 1, 6], pic=[2, 1], wc=[4, 1]),            # case 6, inconsistent C layout ok
        dict(pia=[3, 2, 1], wa=[12, 4, 1], q=2, nb=[3], pic=[1, 2]),         # case 8, beginning/end mismatch
        dict(pia=[1, 2, 3, 4], p=4, na=[2, 3, 4, 5], wa=[1, 2, 6, 24], nb=[3], nc=[2, 4, 5], wc=[1, 2, 8],
             pic=[1, 3, 2]),
    ]
    for over in cases:
        d = dict(good)
        d.update(over)
        yield d


def test_validation_messages(ttvx):
    bufs = [np.zeros(256) for _ in range(3)]
    seen = 0
    for d in _invalid_cases():
        # real buffers (the reference computes when its validation passes); 0 stays a null pointer
        for key, buf in zip("abc", bufs):
            d[key] = buf.ctypes.data if d[key] else 0
        rc, msg = _call(ttvx, **d)
        ref = _ref_msg(**d) if O.ref_available() and d["a"] and d["b"] and d["c"] else "skip"
        if rc in (0, 3):
            # passed validation (no GPU here): the reference must accept it too
            assert ref in (None, "skip"), (d, ref)
            continue
        assert msg.startswith("Error in tlib::"), msg
        seen += 1
        if ref != "skip":
            assert ref == msg, (d, ref, msg)
    assert seen >= 20


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")
def test_layout_consistency_matches_reference(ttvx):
    """Every (pia, pic) pair of order 3 and 4 is accepted or rejected exactly as the reference
    does (it checks only in case 8, with two distinct messages)."""
    for p in (3, 4):
        na = [2, 3, 4, 5][:p]
        for pia in itertools.permutations(range(1, p + 1)):
            pia = list(pia)
            wa = O.strides(na, pia)
            for q in range(1, p + 1):
                nc = O.output_shape(na, q)
                for pic in itertools.permutations(range(1, p)):
                    pic = list(pic)
                    wc = O.strides(nc, pic)
                    a = np.zeros(int(np.prod(na)))
                    b = np.zeros(na[q - 1])
                    c = np.zeros(int(np.prod(nc)))
                    ref = _ref_msg(q, p, a.ctypes.data, na, wa, pia, b.ctypes.data, [na[q - 1]], c.ctypes.data, nc,
                                   wc, pic)
                    rc, msg = _call(ttvx, q, p, a.ctypes.data, na, wa, pia, b.ctypes.data, [na[q - 1]],
                                    c.ctypes.data, nc, wc, pic)
                    if ref is None:
                        # accepted by the reference: ours fails only for lack of a GPU here
                        assert rc in (0, 3), (pia, q, pic, msg)
                    else:
                        assert msg == ref, (pia, q, pic, msg, ref)


def test_fold_every_layout(ttvx):
    """(O, K, I) of the folded view equals the brute-force products for p <= 5."""
    rng = np.random.default_rng(0)
    for p in range(2, 6):
        for pia in itertools.permutations(range(1, p + 1)):
            pia = list(pia)
            na = list(rng.integers(1, 5, p))
            wa = O.strides(na, pia)
            for q in range(1, p + 1):
                info = ttvx.plan_info(q, na, wa, pia)
                qpos = pia.index(q)
                I = int(np.prod([na[m - 1] for m in pia[:qpos]])) if qpos else 1
                Ov = int(np.prod([na[m - 1] for m in pia[qpos + 1:]])) if qpos < p - 1 else 1
                assert info["three_d"], (pia, q, na)
                assert (info["O"], info["K"], info["I"]) == (Ov, na[q - 1], I), (pia, q, na, info)
                cnr = 6 if pia[0] == q else 7 if pia[-1] == q else 8
                if p == 2:
                    cnr = {(1, 1): 2, (2, 1): 3, (1, 2): 4, (2, 2): 5}[(q, pia[0])]
                assert info["ref_case"] == cnr


def test_fold_general_strides(ttvx):
    """A padded view keeps two inner dimensions and goes to the strided kernel."""
    info = ttvx.plan_info(3, [4, 5, 6], [1, 8, 40], [1, 2, 3])
    assert not info["three_d"] and info["variant"] == 11
    info = ttvx.plan_info(2, [4, 5, 6], [1, 8, 40], [1, 2, 3])
    assert info["three_d"] and (info["O"], info["K"], info["I"]) == (6, 5, 4)


def test_variants_and_features(ttvx):
    names = ttvx.variant_names()
    assert len(names) == 15 and names[11] == "strided" and names[12] == "col_wide"
    assert names[13:] == ["col_u1_v8", "col_u2_v8"]
    na = [128] * 4
    for pia in ([1, 2, 3, 4], [4, 3, 2, 1]):
        wa = O.strides(na, pia)
        for q in range(1, 5):
            valid = [v for v in range(len(names)) if ttvx.variant_valid(q, na, wa, pia, v)]
            assert 11 in valid
            info = ttvx.plan_info(q, na, wa, pia)
            assert info["variant"] in valid
            f = ttvx.features(q, na, wa, pia)
            assert f.shape == (24,) and f[0] == 4 and abs(f[5] - 28) < 1e-6
    with pytest.raises(ttvx.TTVError):
        ttvx.set_selection(7, 0)
