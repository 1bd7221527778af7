"""Parity of the HIP path (through the C-ABI) with the oracle and the reference's answers.

Bars: bit-exact on the known-answer grid (integers < 2^24, exact in FP32 and FP64); relative L2
error <= 1e-12 (FP64) / 1e-5 (FP32) against the FP64 oracle on random data, as the north-star
parity targets state.
"""
import itertools
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = os.path.join(REPO, "tests", "golden", "ttv_ref_fixtures.npz")
TOL = {np.float32: 1e-5, np.float64: 1e-12}


def _dev_view(torch, device, flat, na, wa):
    return torch.from_numpy(np.ascontiguousarray(flat)).to(device).as_strided(na, wa)


def _flat(C):
    span = 1 + sum((n - 1) * w for n, w in zip(C.shape, C.stride()))
    return C.as_strided([span], [1]).cpu().numpy()


def _rel(got, ref):
    ref = np.asarray(ref, np.float64)
    n = np.linalg.norm(ref)
    return np.linalg.norm(np.asarray(got, np.float64) - ref) / (n if n > 0 else 1.0)


def test_cpp_dropin_interfaces():
    """tests/cpp/test_api.cpp: every policy tag x ranks 2-4 x {2,4,8}^p x layouts x q (known
    answer, exact), the three example programs, and the error messages."""
    exe = os.path.join(REPO, "ttv_amd", "lib", "test_api")
    assert os.path.exists(exe), "build() compiles tests/cpp/test_api.cpp"
    print("running", exe, flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert " 0 failures" in r.stdout


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_known_answer_grid_every_variant(ttvx, device, dt):
    """Known-answer grid (gtest_tlib_ttv.cpp:146-189) through the stream-ordered entry point,
    with the automatic choice and with every valid forced variant."""
    import torch
    code = ttvx.F32 if dt == np.float32 else ttvx.F64
    try:
        for rank in (2, 3, 4):
            shapes = O.generate_shapes([2] * rank, [3] * rank)
            if rank == 4:
                shapes = shapes[::4]
            for na in shapes:
                for pia in O.generate_permutations(rank):
                    wa = O.strides(na, pia)
                    for q in range(1, rank + 1):
                        a = O.known_answer_input(q, na, wa, pia).astype(dt)
                        A = _dev_view(torch, device, a, na, wa)
                        b = torch.ones(na[q - 1], dtype=A.dtype, device=device)
                        exp = O.known_answer_expected(na[q - 1], int(np.prod(O.output_shape(na, q))))
                        variants = [-1] + [v for v in range(ttvx.lib().ttvx_num_variants()) if ttvx.variant_valid(q, na, wa, pia, v, code)]
                        for v in variants:
                            if v < 0:
                                ttvx.set_selection(ttvx.SELECT_AUTO)
                            else:
                                ttvx.set_selection(ttvx.SELECT_FORCED, v)
                            C = ttvx.ttv(q, A, b)
                            got = _flat(C)
                            assert np.array_equal(got.astype(np.float64), exp), (na, pia, q, v)
    finally:
        ttvx.restore_default_selection()


def test_reference_fixtures(ttvx, device):
    """Seeded random A, b from tests/golden; C written by the compiled reference."""
    import torch
    fx = np.load(FIX)
    n = 0
    for key in fx.files:
        if not key.endswith("_A"):
            continue
        tag, ps, pis = key.split("_")[:3]
        p = int(ps[1:])
        pia = [int(ch) for ch in pis[2:]]
        na = {2: [3, 5], 3: [3, 4, 5], 4: [2, 3, 4, 5]}[p]
        wa = O.strides(na, pia)
        a = fx[key]
        dt = np.float64 if tag == "f64" else np.float32
        A = _dev_view(torch, device, a, na, wa)
        for q in range(1, p + 1):
            b = fx[f"{tag}_p{p}_q{q}_b"]
            cref = fx[f"{tag}_p{p}_pi{''.join(map(str, pia))}_q{q}_C"]
            C = ttvx.ttv(q, A, torch.from_numpy(b).to(device))
            assert _rel(_flat(C), cref) <= TOL[dt], (key, q)
            # the synchronous host-pointer entry point on the same inputs
            Ah = np.lib.stride_tricks.as_strided(a, shape=na, strides=[w * a.itemsize for w in wa])
            Ch = ttvx.ttv(q, Ah, b)
            flat = np.lib.stride_tricks.as_strided(Ch, shape=[cref.size], strides=[Ch.itemsize])
            assert _rel(flat, cref) <= TOL[dt], (key, q, "host")
            n += 1
    assert n == 2 * (2 * 2 + 6 * 3 + 24 * 4)


def test_random_shapes_every_variant(ttvx, device):
    """Random orders 2..7, extents 1..40 (odd sizes, unit extents), random layouts, every q and
    every valid variant, FP32 and FP64, against the FP64 oracle."""
    import torch
    rng = np.random.default_rng(11)
    try:
        for trial in range(60):
            p = int(rng.integers(2, 8))
            na = [int(x) for x in rng.integers(1, 41 if p <= 3 else (13 if p <= 5 else 6), p)]
            pia = [int(x) for x in rng.permutation(p) + 1]
            wa = O.strides(na, pia)
            dt = np.float32 if trial % 2 else np.float64
            code = ttvx.F32 if dt == np.float32 else ttvx.F64
            a = rng.uniform(-1, 1, O.span(na, wa)).astype(dt)
            A = _dev_view(torch, device, a, na, wa)
            for q in range(1, p + 1):
                b = rng.uniform(-1, 1, na[q - 1]).astype(dt)
                ref = O.naive(q, a, na, wa, b, pia=pia)
                for v in [v for v in range(ttvx.lib().ttvx_num_variants()) if ttvx.variant_valid(q, na, wa, pia, v, code)]:
                    ttvx.set_selection(ttvx.SELECT_FORCED, v)
                    C = ttvx.ttv(q, A, torch.from_numpy(b).to(device))
                    assert _rel(_flat(C), ref) <= TOL[dt], (na, pia, q, v, dt)
    finally:
        ttvx.restore_default_selection()


def test_unaligned_and_strided_views(ttvx, device):
    """Base pointers off the 16-byte grid, padded (non-dense) A, and a strided C."""
    import torch
    rng = np.random.default_rng(12)
    na, pia = [37, 12, 9], [1, 2, 3]
    wa = O.strides(na, pia)
    base = torch.from_numpy(rng.uniform(-1, 1, O.span(na, wa) + 3).astype(np.float32)).to(device)
    for off in (0, 1, 2, 3):
        A = base[off:].as_strided(na, wa)
        a = base[off:off + O.span(na, wa)].cpu().numpy()
        for q in (1, 2, 3):
            b = rng.uniform(-1, 1, na[q - 1]).astype(np.float32)
            C = ttvx.ttv(q, A, torch.from_numpy(b).to(device))
            assert _rel(_flat(C), O.naive(q, a, na, wa, b, pia=pia)) <= 1e-5, (off, q)
    # padded views: inner extents with gaps, folded and non-folded
    big = torch.from_numpy(rng.uniform(-1, 1, (9, 16, 40)).astype(np.float32)).to(device)
    for sl in ((slice(0, 7), slice(0, 12), slice(0, 33)), (slice(1, 9), slice(2, 16, 3), slice(0, 40, 2))):
        V = big[sl[0], sl[1], sl[2]]
        na, wa = list(V.shape), list(V.stride())
        pia = ttvx.layout_from_strides(na, wa)
        flat = big.flatten()[V.storage_offset():].cpu().numpy()[:O.span(na, wa)]
        for q in (1, 2, 3):
            b = rng.uniform(-1, 1, na[q - 1]).astype(np.float32)
            C = ttvx.ttv(q, V, torch.from_numpy(b).to(device))
            ref = O.naive(q, flat, na, wa, b, pia=pia)
            assert _rel(_flat(C), ref) <= 1e-5, (sl, q)
    # strided output (every other element)
    na, pia = [20, 30], [1, 2]
    a = rng.uniform(-1, 1, 600).astype(np.float32)
    A = _dev_view(torch, device, a, na, [1, 20])
    b = rng.uniform(-1, 1, 30).astype(np.float32)
    out_buf = torch.zeros(40, dtype=torch.float32, device=device)
    Cs = out_buf.as_strided([20], [2])
    ttvx.ttv(2, A, torch.from_numpy(b).to(device), out=Cs)
    got = out_buf.cpu().numpy()
    ref = O.naive(2, a, na, [1, 20], b, pia=pia)
    assert _rel(got[0::2], ref) <= 1e-5 and np.all(got[1::2] == 0)


def test_long_contractions_split_and_fp64_final(ttvx, device):
    """Few outputs with long fibers (split-n_q + two-pass reduction), and n_q > 2^20 (FP64 final
    reduction), q innermost and outermost; the split row and column kernels also forced."""
    import torch
    rng = np.random.default_rng(13)
    for na, pia, q in (([4, 300000], [2, 1], 2), ([3000, 7], [1, 2], 1), ([7, 3000], [1, 2], 2),
                       ([2, (1 << 20) + 17], [2, 1], 2), ([(1 << 20) + 5, 3], [1, 2], 1),
                       ([3, (1 << 20) + 9], [1, 2], 2)):
        wa = O.strides(na, pia)
        a = rng.uniform(-1, 1, O.span(na, wa)).astype(np.float32)
        b = rng.uniform(-1, 1, na[q - 1]).astype(np.float32)
        A = _dev_view(torch, device, a, na, wa)
        ref = O.naive(q, a, na, wa, b, pia=pia)
        for flags in (0, ttvx.FLAG_KAHAN, ttvx.FLAG_FP64_FINAL):
            C = ttvx.ttv(q, A, torch.from_numpy(b).to(device), flags=flags)
            assert _rel(_flat(C), ref) <= 1e-5, (na, pia, q, flags)
        # the split row (4) and column (9) kernels forced: a few outputs take hundreds of splits,
        # summed one block per output
        for v in (4, 9):
            if not ttvx.variant_valid(q, na, wa, pia, v, ttvx.F32):
                continue
            ttvx.set_selection(ttvx.SELECT_FORCED, v)
            try:
                for flags in (0, ttvx.FLAG_FP64_FINAL):
                    C = ttvx.ttv(q, A, torch.from_numpy(b).to(device), flags=flags)
                    torch.cuda.synchronize()
                    assert _rel(_flat(C), ref) <= 1e-5, (na, pia, q, flags, v)
            finally:
                ttvx.set_selection(ttvx.SELECT_AUTO)


def test_config1_fp64_256cubed(ttvx, device):
    """Order-3 FP64 (256,256,256), first-order, q = 2, U[-1,1] from a seed, vs the naive FP64
    triple loop: relative error <= 1e-12."""
    import torch
    rng = np.random.default_rng(2024)
    na, pia, q = [256, 256, 256], [1, 2, 3], 2
    wa = O.strides(na, pia)
    a = rng.uniform(-1, 1, 256 ** 3)
    b = rng.uniform(-1, 1, 256)
    C = ttvx.ttv(q, _dev_view(torch, device, a, na, wa), torch.from_numpy(b).to(device))
    assert _rel(_flat(C), O.naive(q, a, na, wa, b, pia=pia)) <= 1e-12


@pytest.mark.parametrize("shape", [[128] * 4, [16] * 7])
def test_config2_fp32_full_size(ttvx, device, shape):
    """128^4 and 16^7 FP32 (2^28 elements), every q, first- and last-order, against the FP64
    oracle and the restated reference CPU path (OpenBLAS): relative L2 <= 1e-5."""
    import torch
    rng = np.random.default_rng(7)
    a = rng.uniform(-1, 1, 1 << 28).astype(np.float32)
    Ad = torch.from_numpy(a).to(device)
    O.load_blas()
    p = len(shape)
    for pia in (list(range(1, p + 1)), list(range(p, 0, -1))):
        wa = O.strides(shape, pia)
        A = Ad.as_strided(shape, wa)
        for q in range(1, p + 1):
            b = rng.uniform(-1, 1, shape[q - 1]).astype(np.float32)
            print("config2", shape, pia, q, flush=True)
            got = _flat(ttvx.ttv(q, A, torch.from_numpy(b).to(device)))
            ref = O.naive(q, a, shape, wa, b, pia=pia)
            assert _rel(got, ref) <= 1e-5, (pia, q)
            if q in (1, p) or p == 4:
                cpu, _ = O.log_cpu(q, a, shape, wa, pia, b)
                assert _rel(got, cpu) <= 1e-5, (pia, q, "cpu path")


def test_argument_errors_on_device(ttvx, device):
    import torch
    A = torch.zeros(4, 3, 2, device=device)
    with pytest.raises(ttvx.TTVError, match="contraction dimension of A and B are not equal"):
        ttvx.ttv(2, A, torch.zeros(5, device=device))
    with pytest.raises(ttvx.TTVError, match="contraction mode should be greater zero"):
        ttvx.ttv(4, A, torch.zeros(2, device=device))
    with pytest.raises(ttvx.TTVError, match="input tensor order should be greater zero"):
        ttvx.ttv(1, torch.zeros((), device=device), torch.zeros(1, device=device))
    with pytest.raises(ttvx.TTVError, match="shape vector of C is not valid"):
        ttvx.ttv(1, torch.zeros(3, device=device), torch.zeros(3, device=device))
    # host memory is refused by the stream-ordered entry point before any kernel runs
    import ctypes
    from ttv_amd._lib import sizes
    host = np.zeros(16, np.float32)
    dev_buf = torch.zeros(16, device=device)
    rc = ttvx.lib().ttvx_ttv_f32_async(1, 2, ctypes.c_void_p(host.ctypes.data), sizes([4, 4]), sizes([1, 4]),
                                       sizes([1, 2]), ctypes.c_void_p(dev_buf.data_ptr()), sizes([4]),
                                       ctypes.c_void_p(dev_buf.data_ptr()), sizes([4]), sizes([1]), sizes([1]), 0, None)
    assert rc == 1 and "not device memory" in ttvx.lib().ttvx_last_error().decode()
