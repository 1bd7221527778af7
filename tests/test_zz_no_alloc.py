"""ttvx_ttv_f32_async_noalloc (the entry point that allocates no device memory) against
ttvx_ttv_f32_async on problems whose kernels write split partial sums: two calls in a row, the
second smaller, each bit-identical to the allocating entry point with the same kernel; and a
problem whose partials exceed the static scratch array is refused on the host."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(ttvx, fn, q, A, b, C, flags):
    import torch
    from ttv_amd._lib import sizes
    from ttv_amd.meta import generate_output_layout, layout_from_strides
    na, wa = list(A.shape), list(A.stride())
    pia = layout_from_strides(na, wa)
    nc, wc = list(C.shape), list(C.stride())
    pic = generate_output_layout(pia, q)
    return fn(q, len(na), ctypes.c_void_p(A.data_ptr()), sizes(na), sizes(wa), sizes(pia),
              ctypes.c_void_p(b.data_ptr()), sizes([b.shape[0]]), ctypes.c_void_p(C.data_ptr()), sizes(nc), sizes(wc),
              sizes(pic), flags, ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))


# (shape, q, variant name forced or None, flags): each writes partial sums
CASES = [
    ([8, 1 << 18], 2, "col_split", 0),                    # split n_q over 64-ish blocks, FP32 partials
    ([64, 128, 32], 2, None, 2),                           # FP64 final reduction (TTVX_FLAG_FP64_FINAL)
    ([1 << 17, 16], 1, "dot_split", 0),                    # rows split along k
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_noalloc_matches_allocating_entry_point(ttvx, device, case):
    """Two calls in a row on one problem kind, the second on half the leading extent."""
    import torch
    L = ttvx.lib()
    names = ttvx.variant_names()
    g = torch.Generator(device=device)
    g.manual_seed(1234 + case)
    shape0, q, vname, flags = CASES[case]
    try:
        for shrink in (1, 2):
            shape = list(shape0)
            shape[0] = shape[0] // shrink
            # first-order layout (column-major), the reference's default tensor layout
            A = torch.empty(list(reversed(shape)), device=device).uniform_(-1, 1, generator=g)
            A = A.permute(*reversed(range(len(shape))))
            b = torch.empty(shape[q - 1], device=device).uniform_(-1, 1, generator=g)
            if vname is None:
                ttvx.restore_default_selection()
            else:
                ttvx.set_selection(ttvx.SELECT_FORCED, names.index(vname))
            C1 = ttvx.ttv(q, A, b, flags=flags)
            C2 = torch.full_like(C1, float("nan"))
            print("no-alloc case", case, shape, "q", q, vname, "flags", flags, flush=True)
            ttvx.check(_run(ttvx, L.ttvx_ttv_f32_async_noalloc, q, A, b, C2, flags))
            torch.cuda.synchronize()
            assert torch.equal(C1, C2), "case %d %s: no-alloc result differs from ttvx_ttv_f32_async" % (case, shape)
    finally:
        ttvx.restore_default_selection()


def test_noalloc_refuses_partials_beyond_scratch(ttvx, device):
    """FP64 partials of 2200 x 1100 outputs (19.4 MB) exceed the 16 MiB array: refused before any
    kernel runs, with a message naming the limit."""
    import torch
    L = ttvx.lib()
    A = torch.zeros((2200, 2, 1100), device=device).permute(2, 1, 0).contiguous().permute(2, 1, 0)
    b = torch.zeros(2, device=device)
    C = torch.zeros((2200, 1100), device=device).t().contiguous().t()
    rc = _run(ttvx, L.ttvx_ttv_f32_async_noalloc, 2, A, b, C, 2)
    assert rc == 4, "expected TTVX_ERR_UNSUPPORTED, got %d" % rc
    assert "scratch" in L.ttvx_last_error().decode()
