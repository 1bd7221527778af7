"""Reduced-precision storage of A on the GPU.

* The BF16 and E4M3 quantisers are bit-exact against the oracle's restatements
  (oracle.bf16_rne, oracle.quantize_fp8), codes and scales alike.
* The gfx950 E4M3 decoder (v_cvt_pk_f32_fp8 in the product kernel) reproduces the format's
  definition for all 256 codes (read back through a product with n_q = 1).
* BF16/FP8 products match the FP64 oracle on the dequantised values within FP32 accumulation
  error (relative L2 <= 1e-5), with and without Kahan; the error against the original FP32
  values is the storage error and is reported, not asserted beyond loose sanity bounds.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


def _flat(C):
    span = 1 + sum((n - 1) * w for n, w in zip(C.shape, C.stride()))
    return C.as_strided([span], [1]).cpu().numpy()


def test_bf16_quantiser_bit_exact(ttvx, device):
    import torch
    rng = np.random.default_rng(0)
    x = (rng.standard_normal(1 << 20) * rng.choice([1e-3, 1, 1e3], 1 << 20)).astype(np.float32)
    x[:6] = [0.0, -0.0, 1.00390625, 1.01171875, np.inf, -np.inf]
    q = ttvx.quantize_bf16(torch.from_numpy(x).to(device))
    got = q.cpu().numpy().view(np.uint16)
    assert np.array_equal(got, O.bf16_rne(x))


@pytest.mark.parametrize("na,pia,q,fps", [([64, 33, 20], [1, 2, 3], 2, 8), ([64, 33, 20], [3, 2, 1], 1, 1),
                                          ([128, 16, 16], [2, 3, 1], 3, 64), ([4096, 4, 8], [1, 2, 3], 1, 1)])
def test_fp8_quantiser_bit_exact(ttvx, device, na, pia, q, fps):
    import torch
    rng = np.random.default_rng(1)
    wa = O.strides(na, pia)
    a = rng.standard_normal(O.span(na, wa)).astype(np.float32)
    a[rng.uniform(size=a.size) < 0.01] *= 100.0
    A = torch.from_numpy(a).to(device).as_strided(na, wa)
    codes, scales, f = ttvx.quantize_fp8(q, A, fps)
    assert f == fps
    oc, osc = O.quantize_fp8(q, a, na, wa, pia, fps)
    assert np.array_equal(scales.cpu().numpy(), osc)
    assert np.array_equal(codes.as_strided([a.size], [1]).cpu().numpy(), oc)


def test_e4m3_hardware_decode_all_codes(ttvx, device):
    """A = all 256 codes as a 256 x 1 tensor, q = 2 (n_q = 1), b = 1, scale = 1: C = decode(code)."""
    import torch
    codes = torch.arange(256, dtype=torch.int32).to(torch.uint8).to(device).as_strided([256, 1], [1, 256])
    scales = torch.ones(256, dtype=torch.float32, device=device)
    b = torch.ones(1, dtype=torch.float32, device=device)
    C = ttvx.ttv_fp8(2, codes, scales, 1, b)
    got = C.cpu().numpy()
    ref = O.e4m3_decode(np.arange(256, dtype=np.uint8))
    fin = np.isfinite(ref)
    assert np.array_equal(got[fin], ref[fin])
    assert np.all(np.isnan(got[~fin]))


@pytest.mark.parametrize("fmt", ["bf16", "fp8"])
def test_quantised_products_vs_oracle(ttvx, device, fmt):
    import torch
    rng = np.random.default_rng(2)
    for na in ([64, 48, 40], [32, 16, 24, 8], [4096, 4, 8, 4]):
        p = len(na)
        for pia in (list(range(1, p + 1)), list(range(p, 0, -1)), [int(x) for x in rng.permutation(p) + 1]):
            wa = O.strides(na, pia)
            a = rng.standard_normal(O.span(na, wa)).astype(np.float32)
            a[rng.uniform(size=a.size) < 0.01] *= 100.0
            Ad = torch.from_numpy(a).to(device).as_strided(na, wa)
            for q in range(1, p + 1):
                b = rng.uniform(-1, 1, na[q - 1]).astype(np.float32)
                bd = torch.from_numpy(b).to(device)
                ref = O.naive(q, a, na, wa, b, pia=pia)
                if fmt == "bf16":
                    Aq = ttvx.quantize_bf16(Ad)
                    deq = O.bf16_to_f32(O.bf16_rne(a))
                    ref_q = O.naive(q, deq, na, wa, b, pia=pia)
                    runs = [ttvx.ttv_bf16(q, Aq, bd, flags=f) for f in (0, ttvx.FLAG_KAHAN)]
                    storage_bound = 1e-2
                else:
                    codes, scales, fps = ttvx.quantize_fp8(q, Ad)
                    table = O.e4m3_decode(np.arange(256, dtype=np.uint8))
                    deq = table[codes.as_strided([a.size], [1]).cpu().numpy()]
                    ref_q = O.naive(q, deq, na, wa, b, pia=pia)
                    ref_q *= scales.cpu().numpy().astype(np.float64)[np.arange(ref_q.size) // fps]
                    runs = [ttvx.ttv_fp8(q, codes, scales, fps, bd, flags=f) for f in (0, ttvx.FLAG_KAHAN)]
                    storage_bound = 6e-2
                for C in runs:
                    got = _flat(C)
                    assert _rel(got, ref_q) <= 1e-5, (fmt, na, pia, q)
                    assert _rel(got, ref) <= storage_bound, (fmt, na, pia, q)


def test_order8_precision_config(ttvx, device):
    """Order-8 FP32 tensor (4096,4,8,4,8,4,2,64) = 2^31 elements (8 GiB), layout drawn from a
    seed, q in {1,4,8}; A ~ N(0,1) with 1% of entries x100; BF16 and FP8-e4m3 storage with FP32
    accumulation (and Kahan).  Relative L2 and max errors against the FP64 oracle on the original
    values are printed; BF16 must stay within 1e-2, FP8 within its storage error (<= 6e-2)."""
    import torch
    na = [4096, 4, 8, 4, 8, 4, 2, 64]
    rng = np.random.default_rng(31)
    pia = [int(x) for x in rng.permutation(8) + 1]
    wa = O.strides(na, pia)
    n = int(np.prod(na))
    g = torch.Generator(device=device)
    g.manual_seed(31)
    A = torch.randn(n, generator=g, device=device)
    A.mul_(torch.where(torch.rand(n, generator=g, device=device) < 0.01, 100.0, 1.0))
    a = A.cpu().numpy()
    Av = A.as_strided(na, wa)
    Abf = ttvx.quantize_bf16(A).as_strided(na, wa)
    for q in (1, 4, 8):
        b = rng.uniform(-1, 1, na[q - 1]).astype(np.float32)
        bd = torch.from_numpy(b).to(device)
        print("order-8 config: pia", pia, "q", q, flush=True)
        ref = O.naive(q, a, na, wa, b, pia=pia)
        rn, rmax = np.linalg.norm(ref), np.max(np.abs(ref))
        res = {}
        res["fp32"] = _flat(ttvx.ttv(q, Av, bd))
        res["bf16"] = _flat(ttvx.ttv_bf16(q, Abf, bd))
        res["bf16_kahan"] = _flat(ttvx.ttv_bf16(q, Abf, bd, flags=ttvx.FLAG_KAHAN))
        codes, scales, fps = ttvx.quantize_fp8(q, Av)
        res["fp8"] = _flat(ttvx.ttv_fp8(q, codes, scales, fps, bd))
        res["fp8_kahan"] = _flat(ttvx.ttv_fp8(q, codes, scales, fps, bd, flags=ttvx.FLAG_KAHAN))
        del codes, scales
        for k, got in res.items():
            d = got.astype(np.float64) - ref
            print("  q=%d %-10s rel L2 %.3e  max abs %.3e (max |C| %.3e)" % (q, k, np.linalg.norm(d) / rn,
                                                                            np.max(np.abs(d)), rmax), flush=True)
        assert np.linalg.norm(res["fp32"] - ref) <= 1e-5 * rn
        assert np.linalg.norm(res["bf16"] - ref) <= 1e-2 * rn
        assert np.linalg.norm(res["fp8"] - ref) <= 6e-2 * rn
        del res, ref


def test_kahan_long_contraction(ttvx, device):
    """n_q = 2^22 with a large common offset: compensated summation stays at the FP64 answer
    while plain FP32 accumulation drifts."""
    import torch
    n = 1 << 22
    rng = np.random.default_rng(3)
    a = (1.0 + rng.uniform(-1e-3, 1e-3, n)).astype(np.float32)
    b = np.ones(n, np.float32)
    A = torch.from_numpy(a).to(device).as_strided([1, n], [n, 1])
    ref = O.naive(2, a, [1, n], [n, 1], b, pia=[1, 2])
    bd = torch.from_numpy(b).to(device)
    plain = _flat(ttvx.ttv(2, A, bd))
    kahan = _flat(ttvx.ttv(2, A, bd, flags=ttvx.FLAG_KAHAN))
    e_plain, e_kahan = _rel(plain, ref), _rel(kahan, ref)
    assert e_kahan <= 1e-7, (e_plain, e_kahan)
    assert e_kahan <= e_plain


@pytest.mark.parametrize("rows", ["narrow", "wide", "short"])
@pytest.mark.parametrize("fmt", ["f32", "bf16", "fp8"])
def test_output_alignment_padding_and_scale_groups(ttvx, device, fmt, rows):
    """Outputs written by vector stores: C at every 4-byte offset off the 16-byte grid, C with
    padded rows, every kernel variant that can run, and FP8 scale groups smaller and larger than
    a lane's vector (fibers_per_scale 1, 3, 16, 1000): each output equals the FP64 oracle on the
    values the kernel reads.  "wide": 300 vectors per output row, so whole waves lie along i
    (the wave-transposed stores of BF16/FP8) and the last tile of 256 lanes is partly past the row.
    "short": 16 columns and 16 k steps, so lane groups along k meet by shuffles and a lane may
    end with whole lane vectors (the vector stores of that path)."""
    import torch
    rng = np.random.default_rng(44)
    vec = {"f32": 4, "bf16": 8, "fp8": 16}[fmt]
    if rows == "narrow":
        na, pia, q = [64, 8, 6], [1, 2, 3], 2      # I = 64 columns (16-byte vectors), K = 8, O = 6
    elif rows == "short":
        na, pia, q = [16, 16, 40], [1, 2, 3], 2
    else:
        na, pia, q = [300 * vec, 4, 5], [1, 2, 3], 2
    ni, no = na[0], na[2]
    wa = O.strides(na, pia)
    a = rng.uniform(-1, 1, O.span(na, wa)).astype(np.float32)
    b = rng.uniform(-1, 1, na[q - 1]).astype(np.float32)
    Ad = torch.from_numpy(a).to(device).as_strided(na, wa)
    bd = torch.from_numpy(b).to(device)
    nc = [ni, no]
    if fmt == "f32":
        ref = O.naive(q, a, na, wa, b, pia=pia)
        groups = [None]
    elif fmt == "bf16":
        Aq = ttvx.quantize_bf16(Ad)
        ref = O.naive(q, O.bf16_to_f32(O.bf16_rne(a)), na, wa, b, pia=pia)
        groups = [None]
    else:
        groups = [1, 3, 16, 1000]
    table = O.e4m3_decode(np.arange(256, dtype=np.uint8))
    for fps in groups:
        if fmt == "fp8":
            codes, scales, fps = ttvx.quantize_fp8(q, Ad, fibers_per_scale=fps)
            deq = table[codes.as_strided([a.size], [1]).cpu().numpy()]
            ref = O.naive(q, deq, na, wa, b, pia=pia)
            ref = ref * scales.cpu().numpy().astype(np.float64)[np.arange(ref.size) // fps]
        for off, ld in ((0, ni), (1, ni), (2, ni), (3, ni), (0, ni + 2), (1, ni + 3)):
            variants = [-1] + [v for v in range(13)
                               if ttvx.variant_valid(q, na, wa, pia, v, {"f32": ttvx.F32, "bf16": ttvx.BF16,
                                                                         "fp8": ttvx.FP8E4M3}[fmt])]
            for v in variants:
                buf = torch.full((off + ld * no + 8,), -7.0, dtype=torch.float32, device=device)
                C = buf[off:].as_strided(nc, [1, ld])
                if v < 0:
                    ttvx.set_selection(ttvx.SELECT_AUTO)
                else:
                    ttvx.set_selection(ttvx.SELECT_FORCED, v)
                try:
                    if fmt == "f32":
                        ttvx.ttv(q, Ad, bd, out=C)
                    elif fmt == "bf16":
                        ttvx.ttv_bf16(q, Aq, bd, out=C)
                    else:
                        ttvx.ttv_fp8(q, codes, scales, fps, bd, out=C)
                finally:
                    ttvx.set_selection(ttvx.SELECT_AUTO)
                torch.cuda.synchronize()
                got = buf.cpu().numpy()
                mat = np.stack([got[off + o * ld: off + o * ld + ni] for o in range(no)])   # [o][i]
                assert _rel(mat.reshape(-1), ref) <= 1e-5, (fmt, fps, off, ld, v)
                # nothing written outside the outputs (padding and the bytes before C keep -7)
                mask = np.ones(got.size, bool)
                for o in range(no):
                    mask[off + o * ld: off + o * ld + ni] = False
                assert np.all(got[mask] == -7.0), (fmt, fps, off, ld, v)
