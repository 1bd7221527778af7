"""Rows whose length the 16-byte vector width does not divide, and tensors that start off the
16-byte grid, in all four storage types: the kernels keep full-width vector loads (element-aligned)
and clamp the last vector of a row to the row's end.  Every valid kernel variant is forced on
column problems (I = n*vmax + r, r = 1, 2, 3, vmax - 1; also I = 70*vmax + r, i-groups of a wave
or more, where BF16/FP8 lanes pass realigned dwords to their neighbours) and row problems
(K = n*vmax + r), with the tensor's base pointer 0, 1 and 3 elements past an aligned address; outputs must equal the FP64
oracle on the values the kernel reads (FP32 1e-5, FP64 1e-12), and nothing outside C is written.
The clamped vectors re-read elements that belong to the neighbouring lane; a non-finite value
there must not reach another output (checked with an infinity planted next to a row end)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

VMAX = {"f32": 4, "f64": 2, "bf16": 8, "fp8": 16}
TOL = {"f32": 1e-5, "f64": 1e-12, "bf16": 1e-5, "fp8": 1e-5}


def _rel(got, ref):
    ref = np.asarray(ref, np.float64)
    n = np.linalg.norm(ref)
    return np.linalg.norm(np.asarray(got, np.float64) - ref) / (n if n > 0 else 1.0)


def _shapes(fmt):
    v = VMAX[fmt]
    rs = sorted({1, 2, 3, v - 1} - {0}) if v > 1 else [1]
    out = []
    for r in rs:
        n = 3 * v + r
        out.append(([n, 7, 5], [1, 2, 3], 2))      # columns: I = n
        out.append(([n, 6, 4], [1, 2, 3], 1))      # rows: K = n contiguous
        out.append(([5, 9, n], [3, 2, 1], 1))      # columns in last-order layout: I = n (mode 3)
        out.append(([3, n * 9 + 1, 2], [2, 1, 3], 2))  # long rows: K = 9n + 1
        out.append(([70 * v + r, 37, 3], [1, 2, 3], 2))  # slab rows wider than a wave of vectors
    return out


def _variants(ttvx, fmt, q, na, wa, pia):
    code = {"f32": ttvx.F32, "f64": ttvx.F64, "bf16": ttvx.BF16, "fp8": ttvx.FP8E4M3}[fmt]
    return [-1] + [v for v in range(ttvx.lib().ttvx_num_variants()) if ttvx.variant_valid(q, na, wa, pia, v, code)]


@pytest.mark.parametrize("fmt", ["f32", "f64", "bf16", "fp8"])
def test_rows_off_the_vector_grid(ttvx, device, fmt):
    import torch
    rng = np.random.default_rng(91)
    table = O.e4m3_decode(np.arange(256, dtype=np.uint8))
    for na, pia, q in _shapes(fmt):
        wa = O.strides(na, pia)
        span = O.span(na, wa)
        b = rng.uniform(-1, 1, na[q - 1]).astype(np.float64 if fmt == "f64" else np.float32)
        bd = torch.from_numpy(b).to(device)
        for off in (0, 1, 3):
            a = rng.uniform(-1, 1, span).astype(np.float64 if fmt == "f64" else np.float32)
            if fmt in ("f32", "f64"):
                buf = torch.zeros(span + off, dtype=torch.float64 if fmt == "f64" else torch.float32, device=device)
                buf[off:] = torch.from_numpy(a).to(device)
                A = buf[off:].as_strided(na, wa)
                ref = O.naive(q, a, na, wa, b, pia=pia)
            elif fmt == "bf16":
                Af = torch.from_numpy(a).to(device)
                bits = ttvx.quantize_bf16(Af.as_strided([span], [1])).view(torch.int16)
                buf = torch.zeros(span + off, dtype=torch.int16, device=device)
                buf[off:] = bits
                A = buf.view(torch.uint16)[off:].as_strided(na, wa)
                ref = O.naive(q, O.bf16_to_f32(O.bf16_rne(a)), na, wa, b, pia=pia)
            else:
                Af = torch.from_numpy(a).to(device).as_strided(na, wa)
                codes, scales, fps = ttvx.quantize_fp8(q, Af)
                buf = torch.zeros(span + off, dtype=torch.uint8, device=device)
                buf[off:] = codes.as_strided([span], [1])
                A = buf[off:].as_strided(na, wa)
                deq = table[codes.as_strided([span], [1]).cpu().numpy()]
                ref = O.naive(q, deq, na, wa, b, pia=pia)
                ref = ref * scales.cpu().numpy().astype(np.float64)[np.arange(ref.size) // fps]
            nc = [n for m, n in enumerate(na) if m != q - 1]
            for v in _variants(ttvx, fmt, q, na, wa, pia):
                ttvx.set_selection(ttvx.SELECT_AUTO if v < 0 else ttvx.SELECT_FORCED, max(v, 0))
                try:
                    out = torch.full([int(np.prod(nc)) + 4], -7.0, device=device,
                                     dtype=torch.float64 if fmt == "f64" else torch.float32)
                    pic = ttvx.generate_output_layout(pia, q)
                    C = out[2:].as_strided(nc, ttvx.generate_strides(nc, pic))
                    if fmt in ("f32", "f64"):
                        ttvx.ttv(q, A, bd, out=C)
                    elif fmt == "bf16":
                        ttvx.ttv_bf16(q, A, bd, out=C)
                    else:
                        ttvx.ttv_fp8(q, A, scales, fps, bd, out=C)
                    torch.cuda.synchronize()
                finally:
                    ttvx.set_selection(ttvx.SELECT_AUTO)
                got = out.cpu().numpy()
                assert _rel(got[2:-2], ref) <= TOL[fmt], (fmt, na, pia, q, off, v)
                assert np.all(got[:2] == -7.0) and np.all(got[-2:] == -7.0), (fmt, na, pia, q, off, v)


def test_clamped_vectors_keep_outputs_apart(ttvx, device):
    """FP32, vector width 4.  Rows of 7 elements: an infinity at the start of row 1 makes output 1
    alone non-finite (a row's clamped last vector stays inside the row).  Columns, I = 7: the last
    vector of a row is loaded from column 3 and re-reads column 3 of the lane before; an infinity at
    column 3 of k-row 1 must make output column 3 alone non-finite, not the columns 4..6 that the
    clamped vector produces."""
    import torch
    for v_name, q, na, pia in (("rows", 1, [7, 40], [1, 2]), ("columns", 2, [7, 40, 3], [1, 2, 3])):
        wa = O.strides(na, pia)
        span = O.span(na, wa)
        a = np.ones(span, np.float32)
        a[7 if v_name == "rows" else 7 + 3] = np.inf     # row 1 element 0 / k-row 1 column 3
        A = torch.from_numpy(a).to(device).as_strided(na, wa)
        b = torch.ones(na[q - 1], device=device)
        for v in _variants(ttvx, "f32", q, na, wa, pia):
            ttvx.set_selection(ttvx.SELECT_AUTO if v < 0 else ttvx.SELECT_FORCED, max(v, 0))
            try:
                C = ttvx.ttv(q, A, b)
                torch.cuda.synchronize()
            finally:
                ttvx.set_selection(ttvx.SELECT_AUTO)
            got = C.as_strided([C.numel()], [1]).cpu().numpy()
            bad = ~np.isfinite(got)
            expect = np.zeros(got.size, bool)
            expect[1 if v_name == "rows" else 3] = True
            assert np.array_equal(bad, expect), (v_name, v, np.nonzero(bad))
