"""MI355X mode-q tensor-times-vector product, Python front end.

Python counterpart of the reference's binding ``ttvpy`` (ttvpy/src/wrapped_ttv.cpp:18-207):

* ``ttv(q, A, b)``            -- C = A x_q b            (wrapped_ttv.cpp:18-78)
* ``ttvs(q, A, bs, order)``   -- all modes but q contracted, forward/backward/optimal order
                                 (wrapped_ttv.cpp:83-198)

Differences from ttvpy: any float32/float64 strided layout is accepted (the layout tuple is
derived from the strides instead of being forced to last-order, wrapped_ttv.cpp:44-45), and
torch tensors on the GPU are computed in place on the current stream.  numpy arrays and CPU
tensors go through the library's synchronous entry points (host buffers are staged).

Everything runs in libttvx.so (HIP kernels for gfx950).  There is no CPU fallback: without the
built library or without a GPU the calls raise.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import (BF16, F32, F64, FLAG_FP64_FINAL, FLAG_KAHAN, FP8E4M3, SELECT_AUTO, SELECT_FORCED,
                   SELECT_HEURISTIC, SELECT_LEARNED, TTVError, check, lib, sizes)
from .meta import (generate_k_order_layout, generate_output_layout, generate_output_shape, generate_strides,
                   layout_from_strides)

__all__ = [
    "ttv", "ttvs", "quantize_bf16", "quantize_fp8", "ttv_bf16", "ttv_fp8", "plan_info", "variant_names",
    "set_selection", "variant_valid", "features", "TTVError", "F32", "F64", "BF16", "FP8E4M3", "FLAG_KAHAN",
    "FLAG_FP64_FINAL", "SELECT_HEURISTIC", "SELECT_LEARNED", "SELECT_FORCED", "SELECT_AUTO",
    "restore_default_selection", "set_selector_table", "generate_output_shape",
    "generate_output_layout", "generate_strides", "generate_k_order_layout", "layout_from_strides",
]


def _is_torch(x):
    try:
        import torch
        return isinstance(x, torch.Tensor)
    except ImportError:
        return False


def _describe(x):
    """(shape, element strides, data pointer, dtype name, on_device) of a torch tensor or ndarray."""
    if _is_torch(x):
        return list(x.shape), list(x.stride()), x.data_ptr(), str(x.dtype).replace("torch.", ""), x.is_cuda
    x = np.asarray(x)
    if any(s % x.itemsize for s in x.strides) or any(s < 0 for s in x.strides):
        raise TTVError(_lib.ERR_ARGUMENT, "ttv_amd: numpy strides must be non-negative multiples of the item size")
    return list(x.shape), [s // x.itemsize for s in x.strides], x.ctypes.data, x.dtype.name, False


def _stream_ptr():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _vector(b, name="b"):
    """b as a 1-D vector with unit stride: a contiguous copy when it is strided."""
    if _is_torch(b):
        if b.dim() != 1:
            raise TTVError(_lib.ERR_ARGUMENT, "ttv_amd: %s must be a 1-D vector" % name)
        return b if b.stride(0) == 1 else b.contiguous()
    arr = np.asarray(b)
    if arr.ndim != 1:
        raise TTVError(_lib.ERR_ARGUMENT, "ttv_amd: %s must be a 1-D vector" % name)
    return np.ascontiguousarray(arr)


def _checked_out(out, nc, dtype_name, need_device):
    """Strides of a caller-supplied output after checking its element type, shape and placement."""
    shape, strides, _, dt, dev = _describe(out)
    if dt != dtype_name:
        raise TTVError(_lib.ERR_ARGUMENT, "ttv_amd: out must be %s, got %s" % (dtype_name, dt))
    if list(shape) != list(nc):
        raise TTVError(_lib.ERR_ARGUMENT, "ttv_amd: out must have shape %s (A's shape without mode q), got %s"
                       % (list(nc), list(shape)))
    if need_device and not dev:
        raise TTVError(_lib.ERR_ARGUMENT, "ttv_amd: out must be a GPU tensor")
    return strides


def ttv(q, A, b, *, flags=0, out=None):
    """C = A x_q b (q one-based).  C has A's shape without mode q and the reference's output
    layout (layout.h:143-207: A's layout with q removed) with dense strides.  b is a 1-D vector of
    A's element type (a strided b is copied); a given `out` must have C's shape and A's type."""
    na, wa, pa, dt, dev = _describe(A)
    b = _vector(b)
    nb, _, pb, dtb, devb = _describe(b)
    p = len(na)
    if dt not in ("float32", "float64") or dtb != dt:
        raise TTVError(_lib.ERR_UNSUPPORTED, "ttv_amd.ttv: A and b must both be float32 or both float64")
    pia = layout_from_strides(na, wa)
    if not (1 <= q <= p) or any(n == 0 for n in na) or p < 2:
        # let the library report the argument error with the reference's message
        one = ctypes.c_void_p(16)
        m = max(p - 1, 1)
        check(lib().ttvx_ttv_f32(5, 1, 2, q, p, one, sizes(na), sizes(wa), sizes(pia), one, sizes(nb or [0]),
                                 one, sizes([0] * m), sizes([1] * m), sizes(range(1, m + 1))))
        raise TTVError(_lib.ERR_ARGUMENT, "ttv_amd.ttv: invalid arguments")
    nc = generate_output_shape(na, q)
    pic = generate_output_layout(pia, q)
    wc = generate_strides(nc, pic)
    if out is None:
        if _is_torch(A):
            import torch
            out = torch.empty_strided(nc, wc, dtype=A.dtype, device=A.device)
        else:
            npdt = np.float32 if dt == "float32" else np.float64
            span = 1 + sum((n - 1) * w for n, w in zip(nc, wc))
            buf = np.empty(span, dtype=npdt)
            out = np.lib.stride_tricks.as_strided(buf, shape=nc, strides=[w * buf.itemsize for w in wc])
    else:
        _checked_out(out, nc, dt, False)
    _, wc, pc, _, devc = _describe(out)
    args = (q, p, ctypes.c_void_p(pa), sizes(na), sizes(wa), sizes(pia), ctypes.c_void_p(pb), sizes(nb),
            ctypes.c_void_p(pc), sizes(nc), sizes(wc), sizes(pic))
    if dev and devb and devc:
        fn = lib().ttvx_ttv_f32_async if dt == "float32" else lib().ttvx_ttv_f64_async
        check(fn(*args, flags, _stream_ptr()))
    else:
        fn = lib().ttvx_ttv_f32 if dt == "float32" else lib().ttvx_ttv_f64
        check(fn(5, 1, 2, *args))     # par_loop / subtensor / all: the reference's operator* policy
    return out


def ttvs(q, A, bs, order="optimal"):
    """Contract every mode except q with the vectors bs (one per mode, in mode order without q).

    Same contraction orders as wrapped_ttv.cpp:83-198:
      backward -- modes p..1 from the slowest (:141-150)
      forward  -- modes 1..p from the fastest (:152-160)
      optimal  -- longest vector first, remaining modes renumbered after each product (:161-194)
    and one of this library's own:
      fused    -- for A a dense GPU tensor (in any layout) of order >= 5 with modes stored both
                  before and after q: one product with the Kronecker product of the vectors of the
                  modes stored before q (a row product) and one with that of the modes stored after
                  q (a column product), the larger block first: two product launches and one
                  intermediate of A / max(I, O) elements, whatever p; other inputs take the
                  optimal chain
    """
    if order not in ("optimal", "backward", "forward", "fused"):
        raise ValueError("Error calling ttvpy::ttvs: multiplication order should be either 'optimal', 'backward' or 'forward'.")
    na = list(A.shape)
    p = len(na)
    if p == 0:
        raise ValueError("Error calling ttvpy::ttvs: input tensor order should be greater than zero.")
    if len(bs) != p - 1:
        raise ValueError("Error calling ttvpy::ttvs: number of input vectors is not equal to the tensor order - 1.")
    if q == 0 or q > p:
        raise ValueError("Error calling ttvpy::ttvs: contraction mode should be greater than zero or less than or equal to p.")
    if any(len(b.shape) != 1 for b in bs):
        raise ValueError("Error calling ttvpy::ttvs: some of the input vectors is not a vector.")
    nb = [b.shape[0] for b in bs]
    if nb[:q - 1] != na[:q - 1] or nb[q - 1:] != na[q:]:
        raise ValueError("Error calling ttvpy::ttvs: vector dimension is not compatible with the dimension of a tensor mode.")
    if order == "fused":
        return _ttvs_fused(q, A, bs)
    if order == "backward":
        r0 = p - 1 if q == p else p
        r1 = r0 - 1
        r2 = r1 if q == p else q - 1
        c = ttv(r0, A, bs[p - 2])
        while r1 > q:
            c = ttv(r1, c, bs[r1 - 2])
            r1 -= 1
        while r2 > 0:
            c = ttv(r2, c, bs[r2 - 1])
            r2 -= 1
        return c
    if order == "forward":
        r0 = 2 if q == 1 else 1
        r1 = 2
        r2 = 2 if q == 1 else q
        c = ttv(r0, A, bs[0])
        while r1 < q:
            c = ttv(1, c, bs[r1 - 1])
            r1 += 1
        while r2 < p:
            c = ttv(2, c, bs[r2 - 1])
            r2 += 1
        return c
    pairs = [[bs[r - 1], r] for r in range(1, q)] + [[bs[r - 2], r] for r in range(q + 1, p + 1)]
    pairs.sort(key=lambda pr: pr[0].shape[0])
    c = A
    for idx in range(len(pairs) - 1, -1, -1):
        vec, r = pairs[idx]
        c = ttv(r, c, vec)
        for j in range(idx, -1, -1):
            if pairs[j][1] > r:
                pairs[j][1] -= 1
    return c


def _kron(vectors, like):
    """Kronecker product of 1-D GPU vectors (vector 0 varies fastest), in like's dtype and device
    (a single vector is returned as it is)."""
    import torch
    vs = [_vector(torch.as_tensor(v, device=like.device, dtype=like.dtype)) for v in vectors]
    if len(vs) == 1:
        return vs[0]
    lens = [int(v.shape[0]) for v in vs]
    out = torch.empty(int(np.prod(lens)), dtype=like.dtype, device=like.device)
    ptrs = (ctypes.c_void_p * len(vs))(*[v.data_ptr() for v in vs])
    fn = lib().ttvx_kron_f32 if like.dtype == torch.float32 else lib().ttvx_kron_f64
    check(fn(ptrs, sizes(lens), len(vs), ctypes.c_void_p(out.data_ptr()), _stream_ptr()))
    return out


def _ttvs_fused(q, A, bs):
    """ttvs as c[j] = sum_o wO[o] sum_i A[o, j, i] wI[i] (A dense in its layout; see ttvs).  With q
    stored first or last one of the blocks is empty and the other would contract N / n_q elements
    into n_q outputs in one product; such problems take the optimal chain, and so do orders below 5,
    whose chain has at most three products (measured on MI355X, FP32 1 GiB: 16^7 with q in the
    middle 1.03-1.11x faster fused; 128^4 q = 2, 3 0.85-0.95x, its merged contraction of 16384
    elements needing a split)."""
    na, wa, _, dt, dev = _describe(A)
    pia = layout_from_strides(na, wa)
    qpos = pia.index(q) if q in pia else -1
    if not (_is_torch(A) and dev and dt in ("float32", "float64") and list(wa) == generate_strides(na, pia)
            and 0 < qpos < len(pia) - 1 and len(na) >= 5):
        return ttvs(q, A, bs, order="optimal")
    vec = lambda m: bs[m - 1] if m < q else bs[m - 2]
    inner, outer = pia[:qpos], pia[qpos + 1:]
    nq = na[q - 1]
    n_in = int(np.prod([na[m - 1] for m in inner], dtype=np.int64))
    n_out = int(np.prod([na[m - 1] for m in outer], dtype=np.int64))
    flat = A.as_strided([A.numel()], [1])
    w_in, w_out = _kron([vec(m) for m in inner], A), _kron([vec(m) for m in outer], A)
    if n_out >= n_in:
        # the larger block first (the intermediate is A / O): a column product over o, then rows of i
        s = ttv(2, flat.as_strided([nq * n_in, n_out], [1, nq * n_in]), w_out)        # s[i + I*j]
        return ttv(2, s.as_strided([nq, n_in], [n_in, 1]), w_in)
    s = ttv(2, flat.as_strided([n_out * nq, n_in], [n_in, 1]), w_in)                  # s[j + nq*o]
    return ttv(2, s.as_strided([nq, n_out], [1, nq]), w_out)


# ---- reduced precision --------------------------------------------------------------------

def quantize_bf16(A):
    """bfloat16 bits (torch.uint16 tensor with A's shape and strides), round to nearest even."""
    import torch
    if not (A.is_cuda and A.dtype == torch.float32):
        raise TTVError(_lib.ERR_ARGUMENT, "ttv_amd.quantize_bf16: A must be a float32 GPU tensor")
    na, wa = list(A.shape), list(A.stride())
    span = 1 + sum((n - 1) * w for n, w in zip(na, wa))
    base = A.as_strided([span], [1])
    out = torch.empty(span, dtype=torch.uint16, device=A.device)
    check(lib().ttvx_quantize_bf16_async(span, ctypes.c_void_p(base.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                         _stream_ptr()))
    return out.as_strided(na, wa)


def default_fibers_per_scale(K):
    """Fibers per FP8 scale: groups of at least 4096 elements (scale overhead <= 0.1%)."""
    return max(1, -(-4096 // int(K)))


def quantize_fp8(q, A, fibers_per_scale=None):
    """(codes, scales, fibers_per_scale): e4m3 codes with A's strides and per-group scales."""
    import torch
    if not (A.is_cuda and A.dtype == torch.float32):
        raise TTVError(_lib.ERR_ARGUMENT, "ttv_amd.quantize_fp8: A must be a float32 GPU tensor")
    na, wa = list(A.shape), list(A.stride())
    pia = layout_from_strides(na, wa)
    fps = fibers_per_scale or default_fibers_per_scale(na[q - 1])
    ns = lib().ttvx_fp8_num_scales(q, len(na), sizes(na), sizes(wa), sizes(pia), fps)
    if ns == 0:
        check(_lib.ERR_ARGUMENT)
    span = 1 + sum((n - 1) * w for n, w in zip(na, wa))
    codes = torch.empty(span, dtype=torch.uint8, device=A.device)
    scales = torch.empty(ns, dtype=torch.float32, device=A.device)
    check(lib().ttvx_quantize_fp8_async(q, len(na), ctypes.c_void_p(A.data_ptr()), sizes(na), sizes(wa), sizes(pia),
                                        ctypes.c_void_p(codes.data_ptr()), ctypes.c_void_p(scales.data_ptr()), fps,
                                        _stream_ptr()))
    return codes.as_strided(na, wa), scales, fps


def _quantised_operands(q, Aq, storage, b, out, what):
    """Checked (na, wa, pia, b, C, nc, wc, pic) of a reduced-precision product: Aq a GPU tensor of
    the storage type, b a float32 GPU vector (copied when strided), C float32 of C's shape."""
    import torch
    if not (_is_torch(Aq) and Aq.is_cuda and Aq.dtype == storage):
        raise TTVError(_lib.ERR_ARGUMENT, "ttv_amd.%s: A must be a %s GPU tensor" % (what, str(storage).replace("torch.", "")))
    na, wa = list(Aq.shape), list(Aq.stride())
    if not (1 <= q <= len(na)):
        raise TTVError(_lib.ERR_ARGUMENT, "Error in tlib::tensor_times_vector: contraction mode should be greater zero or less than or equal to p.")
    pia = layout_from_strides(na, wa)
    if not _is_torch(b) or not b.is_cuda or b.dtype != torch.float32:
        raise TTVError(_lib.ERR_ARGUMENT, "ttv_amd.%s: b must be a float32 GPU vector" % what)
    b = _vector(b)
    nc = generate_output_shape(na, q)
    pic = generate_output_layout(pia, q)
    if out is None:
        wc = generate_strides(nc, pic)
        out = torch.empty_strided(nc, wc, dtype=torch.float32, device=Aq.device)
    else:
        if not _is_torch(out):
            raise TTVError(_lib.ERR_ARGUMENT, "ttv_amd.%s: out must be a GPU tensor" % what)
        wc = _checked_out(out, nc, "float32", True)
    return na, wa, pia, b, out, nc, wc, pic


def ttv_bf16(q, Aq, b, *, flags=0, out=None):
    """C = A x_q b for bfloat16 A (torch.uint16 bits), float32 b and C."""
    import torch
    na, wa, pia, b, c, nc, wc, pic = _quantised_operands(q, Aq, torch.uint16, b, out, "ttv_bf16")
    check(lib().ttvx_ttv_bf16_async(q, len(na), ctypes.c_void_p(Aq.data_ptr()), sizes(na), sizes(wa), sizes(pia),
                                    ctypes.c_void_p(b.data_ptr()), sizes([b.shape[0]]), ctypes.c_void_p(c.data_ptr()),
                                    sizes(nc), sizes(wc), sizes(pic), flags, _stream_ptr()))
    return c


def ttv_fp8(q, codes, scales, fibers_per_scale, b, *, flags=0, out=None):
    """C = A x_q b for e4m3 A (codes + scales from quantize_fp8), float32 b and C."""
    import torch
    na, wa, pia, b, c, nc, wc, pic = _quantised_operands(q, codes, torch.uint8, b, out, "ttv_fp8")
    if not (_is_torch(scales) and scales.is_cuda and scales.dtype == torch.float32 and scales.dim() == 1):
        raise TTVError(_lib.ERR_ARGUMENT, "ttv_amd.ttv_fp8: scales must be a float32 GPU vector")
    scales = _vector(scales, "scales")
    check(lib().ttvx_ttv_fp8_async(q, len(na), ctypes.c_void_p(codes.data_ptr()), sizes(na), sizes(wa), sizes(pia),
                                   ctypes.c_void_p(scales.data_ptr()), fibers_per_scale,
                                   ctypes.c_void_p(b.data_ptr()), sizes([b.shape[0]]), ctypes.c_void_p(c.data_ptr()),
                                   sizes(nc), sizes(wc), sizes(pic), flags, _stream_ptr()))
    return c


# ---- planning and selection ---------------------------------------------------------------

_DT = {"float32": F32, "float64": F64, "bfloat16": BF16, "fp8": FP8E4M3}


def plan_info(q, na, wa, pia, dtype=F32):
    """dict(O, K, I, three_d, ref_case, variant) of the folded problem."""
    out = (ctypes.c_int64 * 6)()
    check(lib().ttvx_plan_info(dtype, q, len(na), sizes(na), sizes(wa), sizes(pia), out))
    return dict(O=out[0], K=out[1], I=out[2], three_d=bool(out[3]), ref_case=out[4], variant=out[5])


def variant_names():
    L = lib()
    return [L.ttvx_variant_name(v).decode() for v in range(L.ttvx_num_variants())]


def set_selection(mode, forced_variant=0):
    check(lib().ttvx_set_selection(mode, forced_variant))


def variant_valid(q, na, wa, pia, v, dtype=F32):
    return bool(lib().ttvx_variant_valid(dtype, q, len(na), sizes(na), sizes(wa), sizes(pia), v))


def features(q, na, wa, pia, dtype=F32):
    out = (ctypes.c_float * _lib.NUM_FEATURES)()
    check(lib().ttvx_features(dtype, q, len(na), sizes(na), sizes(wa), sizes(pia), out))
    return np.array(out[:], dtype=np.float32)


def restore_default_selection():
    """Back to the library's default: the shipped tuning table, then the MLP selector with the
    shipped weights, then the fixed rule (for what is absent)."""
    from .selector import load_table, load_weights
    set_selector_weights(load_weights())
    set_selector_table(load_table())
    set_selection(SELECT_AUTO)


def set_selector_table(rows):
    """Tuning table rows (n x 5 int64: dtype, O, K, I, variant); None clears it."""
    L = lib()
    if rows is None:
        check(L.ttvx_selector_set_table(None, 0))
        return
    arr = np.ascontiguousarray(rows, dtype=np.int64).reshape(-1, 5)
    check(L.ttvx_selector_set_table(arr.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), arr.shape[0]))


def set_selector_weights(params):
    """Loads the selector MLP parameters (float32 vector, layout in ttvx.h); None unloads."""
    L = lib()
    if params is None:
        check(L.ttvx_selector_set_weights(None, 0))
        return
    arr = np.ascontiguousarray(params, dtype=np.float32)
    check(L.ttvx_selector_set_weights(arr.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), arr.size))
