"""CPU oracle of the mode-q tensor-times-vector product -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module, and
only as the checker (or, for bench.py, as the timed CPU baseline).  The product path
(libttvx.so and the ttv_amd package) never imports it.

What is here, and what pins it (tests/test_oracle.py):

* ``naive`` -- FP64 textbook loop (README.md:13-18, include/tlib/ttv.h:28-53) from
  ttv_oracle.c; pinned by the known-answer closed form (test/src/gtest_tlib_ttv.cpp:132),
  numpy.einsum, the ttvpy test vectors (ttvpy/tests/test.py:5-27), the example outputs
  (example/interface*.cpp) and the compiled reference in oracle/_ref.
* ``log_cpu`` -- the reference's default CPU path restated in C with OpenBLAS
  (tensor_times_vector.h:1207-1256, matrix_times_vector.h:356-373); pinned like ``naive``.
* metadata (``output_shape``, ``output_layout``, ``strides``, ``k_order_layout``) -- restated
  from include/tlib/detail/{shape,layout,strides}.h; pinned by the golden tables of
  test/src/gtest_tlib_{shape,layout,strides}.cpp and the compiled reference.
* ``known_answer`` -- gtest_tlib_ttv.cpp:27-71 and :132.
* ``bf16_rne`` / ``e4m3_encode`` / ``e4m3_decode`` / ``quantize_fp8`` -- bit-level restatements of
  the formats (round to nearest even; OCP E4M3, bias 7, max 448, NaN S.1111.111); pinned by
  exhaustive decode of all 256 codes against the format's definition, closed-form points
  (powers of two, halfway cases) and numpy's float32 arithmetic.
* ``mlp_*`` -- FP64 forward/backward of the selector network; pinned by central finite
  differences.
"""
import ctypes
import glob
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libttvoracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libttvref.so")

_sz = ctypes.c_size_t
_szp = ctypes.POINTER(ctypes.c_size_t)
_vp = ctypes.c_void_p

_lib = None
_ref = None
_blas_path = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_ttv_naive_f64.argtypes = [_sz, _sz, _vp, _szp, _szp, _vp, _vp, _szp]
        L.oracle_ttv_naive_f32in.argtypes = [_sz, _sz, _vp, _szp, _szp, _vp, _vp, _szp]
        for n in ("oracle_ttv_log_f32", "oracle_ttv_log_f64"):
            getattr(L, n).argtypes = [_sz, _sz, _vp, _szp, _szp, _szp, _vp, _vp, _szp, _szp, ctypes.c_int]
            getattr(L, n).restype = ctypes.c_int
        L.oracle_blas_load.argtypes = [ctypes.c_char_p]
        L.oracle_blas_load.restype = ctypes.c_int
        L.oracle_blas_loaded.restype = ctypes.c_int
        L.oracle_number_cores.restype = ctypes.c_uint
        L.oracle_ttv_init_known.argtypes = [_sz, _sz, _vp, _szp, _szp, _szp]
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    """The reference compiled from its own headers (oracle/_ref), or None when absent."""
    global _ref
    if _ref is None and ref_available():
        R = ctypes.CDLL(REF_PATH)
        for n in ("ref_ttv_f64", "ref_ttv_f32"):
            getattr(R, n).argtypes = [ctypes.c_int, _sz, _sz, _vp, _szp, _szp, _szp, _vp, _szp, _vp, _szp, _szp, _szp]
            getattr(R, n).restype = ctypes.c_int
        R.ref_last_error.restype = ctypes.c_char_p
        R.ref_output_shape.argtypes = [_szp, _sz, _sz, _szp]
        R.ref_output_layout.argtypes = [_szp, _sz, _sz, _szp]
        R.ref_strides.argtypes = [_szp, _szp, _sz, _szp]
        R.ref_k_order_layout.argtypes = [_sz, _sz, _szp]
        R.ref_shape_class.argtypes = [_szp, _sz]
        R.ref_is_valid_layout.argtypes = [_szp, _sz]
        R.ref_is_valid_strides.argtypes = [_szp, _sz, _szp]
        R.ref_inverse_mode.argtypes = [_szp, _sz, _sz]
        R.ref_number_cores.restype = ctypes.c_uint
        _ref = R
    return _ref


def _arr(v):
    a = (ctypes.c_size_t * max(1, len(v)))()
    for i, x in enumerate(v):
        a[i] = int(x)
    return a


def _ptr(x):
    return ctypes.c_void_p(x.ctypes.data)


# ---- metadata (include/tlib/detail/{shape,layout,strides}.h) ---------------------------------

def is_valid_shape(n):
    return len(n) > 0 and all(x != 0 for x in n)


def is_scalar(n):
    return is_valid_shape(n) and all(x == 1 for x in n)


def is_vector(n):
    if not is_valid_shape(n):
        return False
    if len(n) == 1:
        return n[0] > 1
    return (n[0] > 1 or n[1] > 1) and (n[0] == 1 or n[1] == 1) and all(x == 1 for x in n[2:])


def is_matrix(n):
    return is_valid_shape(n) and len(n) >= 2 and n[0] > 1 and n[1] > 1 and all(x == 1 for x in n[2:])


def is_tensor(n):
    return is_valid_shape(n) and len(n) >= 3 and any(x > 1 for x in n[2:])


def is_valid_layout(pi):
    p = len(pi)
    return p > 0 and sorted(pi) == list(range(1, p + 1))


def k_order_layout(p, k):
    """layout.h:57-84: positions 0..m-1 hold m..1, then m+1..p; k=0 means last-order."""
    if k == 0:
        k = p
    m = min(k, p)
    return list(range(m, 0, -1)) + list(range(m + 1, p + 1))


def output_shape(na, q):
    return [x for r, x in enumerate(na, 1) if r != q]


def output_layout(pia, q):
    return [m - 1 if m > q else m for m in pia if m != q]


def strides(n, pi):
    """strides.h:31-57 including the all-ones rule for vector/scalar shapes (:44-45)."""
    p = len(n)
    w = [1] * p
    if is_vector(n) or is_scalar(n):
        return w
    for r in range(1, p):
        w[pi[r] - 1] = w[pi[r - 1] - 1] * n[pi[r - 1] - 1]
    return w


def is_valid_strides(pi, w):
    return all(w[pi[r] - 1] <= w[pi[r + 1] - 1] for r in range(len(pi) - 1))


def inverse_mode(pi, mode):
    return pi.index(mode) + 1


def span(n, w):
    return 1 + sum((a - 1) * b for a, b in zip(n, w))


# ---- products ---------------------------------------------------------------------------------

def naive(q, a_flat, na, wa, b, wc=None, pia=None):
    """FP64 oracle on a flat buffer with strides wa; C in the dense output layout (or wc)."""
    p = len(na)
    if wc is None:
        pia = pia if pia is not None else list(range(1, p + 1))
        wc = strides(output_shape(na, q), output_layout(pia, q)) if p > 1 else [1]
    nc = output_shape(na, q)
    c = np.zeros(span(nc, wc) if p > 1 else 1, dtype=np.float64)
    a_flat = np.ascontiguousarray(a_flat)
    b = np.ascontiguousarray(b)
    if a_flat.dtype == np.float64:
        b64 = np.ascontiguousarray(b, dtype=np.float64)
        lib().oracle_ttv_naive_f64(q, p, _ptr(a_flat), _arr(na), _arr(wa), _ptr(b64), _ptr(c), _arr(wc))
    elif a_flat.dtype == np.float32:
        b32 = np.ascontiguousarray(b, dtype=np.float32)
        lib().oracle_ttv_naive_f32in(q, p, _ptr(a_flat), _arr(na), _arr(wa), _ptr(b32), _ptr(c), _arr(wc))
    else:
        raise TypeError("naive oracle takes float32 or float64 A")
    return c


def default_blas_path():
    try:
        import scipy
        d = os.path.join(os.path.dirname(os.path.dirname(scipy.__file__)), "scipy.libs")
        cands = sorted(glob.glob(os.path.join(d, "libscipy_openblas*.so")))
        if cands:
            return cands[0]
    except Exception:
        pass
    for c in ("/usr/lib/x86_64-linux-gnu/libopenblas.so.0", "/usr/lib/x86_64-linux-gnu/libopenblas.so"):
        if os.path.exists(c):
            return c
    return None


def load_blas(path=None):
    """Binds OpenBLAS for log_cpu; returns the path bound or None (non-BLAS loops then)."""
    global _blas_path
    if lib().oracle_blas_loaded():
        return _blas_path
    path = path or default_blas_path()
    if path and lib().oracle_blas_load(path.encode()):
        _blas_path = path
        return path
    return None


def number_cores():
    return int(lib().oracle_number_cores())


def log_cpu(q, a_flat, na, wa, pia, b, c_out=None, threads=0):
    """The reference's default CPU path (par_loop/subtensor/all; mtv(par_blas) for cases 2-7)."""
    p = len(na)
    nc = output_shape(na, q)
    pic = output_layout(pia, q)
    wc = strides(nc, pic)
    dt = a_flat.dtype
    if c_out is None:
        c_out = np.zeros(span(nc, wc), dtype=dt)
    b = np.ascontiguousarray(b, dtype=dt)
    fn = lib().oracle_ttv_log_f32 if dt == np.float32 else lib().oracle_ttv_log_f64
    case = fn(q, p, _ptr(a_flat), _arr(na), _arr(wa), _arr(pia), _ptr(b), _ptr(c_out), _arr(wc), _arr(pic), threads)
    return c_out, case


def ref_ttv(q, a_flat, na, wa, pia, b, policy=4):
    """The compiled reference (policy id per oracle/ref_shim.cpp; 4 = par_loop/subtensor/all)."""
    R = ref()
    if R is None:
        raise RuntimeError("oracle/_ref not built")
    p = len(na)
    nc = output_shape(na, q)
    pic = output_layout(pia, q)
    wc = strides(nc, pic)
    dt = a_flat.dtype
    c = np.zeros(span(nc, wc), dtype=dt)
    b = np.ascontiguousarray(b, dtype=dt)
    fn = R.ref_ttv_f32 if dt == np.float32 else R.ref_ttv_f64
    rc = fn(policy, q, p, _ptr(a_flat), _arr(na), _arr(wa), _arr(pia), _ptr(b), _arr([len(b)]), _ptr(c),
            _arr(nc), _arr(wc), _arr(pic))
    if rc != 0:
        raise RuntimeError(R.ref_last_error().decode())
    return c


# ---- known answers (test/src/gtest_tlib_ttv.cpp) -----------------------------------------------

def known_answer_input(q, na, wa, pia):
    """A filled by ttv_init (gtest_tlib_ttv.cpp:27-71); returns the flat FP64 buffer."""
    p = len(na)
    a = np.zeros(span(na, wa), dtype=np.float64)
    lib().oracle_ttv_init_known(q, p, _ptr(a), _arr(na), _arr(wa), _arr(pia))
    return a


def known_answer_expected(nq, count):
    """C[j] = (nq^2 (2(j+1) - 1) + nq) / 2 for b = 1 (gtest_tlib_ttv.cpp:132)."""
    i = np.arange(1, count + 1, dtype=np.float64)
    return (nq * nq * (2 * i - 1) + nq) / 2


def generate_shapes(start, steps):
    """gtest_aux.h:28-63: every extent doubles from start, steps values per mode, mode 1 fastest."""
    import itertools
    rank = len(start)
    per_mode = [[start[r] * (2 ** j) for j in range(steps[r])] for r in range(rank)]
    out = []
    for combo in itertools.product(*reversed(per_mode)):
        out.append(list(reversed(combo)))
    return out


def generate_permutations(rank):
    """gtest_aux.h:65-79: all rank! layouts in lexicographic order."""
    import itertools
    return [list(x) for x in itertools.permutations(range(1, rank + 1))]


# ---- reduced precision formats --------------------------------------------------------------

def bf16_rne(x):
    """float32 -> bfloat16 bits (uint16), round to nearest even; NaN -> quiet NaN."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan = ((u & 0x7F800000) == 0x7F800000) & ((u & 0x007FFFFF) != 0)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, ((u >> 16) | 0x40) & 0xFFFF, r)
    return r.astype(np.uint16)


def bf16_to_f32(h):
    return (np.asarray(h, dtype=np.uint32) << 16).view(np.float32)


def e4m3_decode(codes):
    """OCP E4M3 codes (uint8) -> float32.  Exponent bias 7; field 0 = subnormal m/8 * 2^-6."""
    c = np.asarray(codes, dtype=np.uint8).astype(np.int64)
    s = np.where(c & 0x80, -1.0, 1.0)
    e = (c >> 3) & 0xF
    m = c & 0x7
    mag = np.where(e == 0, m / 8.0 * 2.0 ** -6, (1.0 + m / 8.0) * 2.0 ** (e - 7))
    mag = np.where((c & 0x7F) == 0x7F, np.nan, mag)
    return (s * mag).astype(np.float32)


def e4m3_encode(x):
    """float32 -> OCP E4M3 code, round to nearest even, saturating to +-448 (NaN -> 0x7F)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    sign = ((u >> 24) & 0x80).astype(np.uint64)
    au = u & 0x7FFFFFFF
    ax = au.astype(np.uint32).view(np.float32)
    # normal range: keep 3 mantissa bits with RNE on the float32 bit pattern
    r = au + 0x7FFFF + ((au >> 20) & 1)
    e = (r >> 23).astype(np.int64) - 127 + 7
    m = ((r >> 20) & 7).astype(np.int64)
    normal = (sign | (np.clip(e, 0, 15).astype(np.uint64) << 3) | m.astype(np.uint64))
    sat = (e > 15) | ((e == 15) & (m == 7))
    normal = np.where(sat, sign | 0x7E, normal)
    sub = sign | np.rint(np.minimum(ax, np.float32(1.0)) * np.float32(512.0)).astype(np.uint64)
    out = np.where(ax < np.float32(0.015625), sub, normal)
    out = np.where(ax >= np.float32(448.0), sign | 0x7E, out)
    out = np.where(au > 0x7F800000, sign | 0x7F, out)
    return out.astype(np.uint8)


def quantize_fp8(q, a_flat, na, wa, pia, fps):
    """Per-group absmax e4m3 quantisation, the same arithmetic as ttvx_quantize_fp8_async:
    group g = fibers j in [g*fps, (g+1)*fps) (j = C offset of the fiber); scale = absmax/448
    (1 if absmax == 0); code = e4m3(x * (448/absmax)).  Returns (codes like a_flat, scales)."""
    a_flat = np.ascontiguousarray(a_flat, dtype=np.float32)
    fib, offs = fiber_offsets(q, na, wa, pia)      # [nfib, nq] element offsets
    vals = a_flat[offs]
    nfib = vals.shape[0]
    ng = (nfib + fps - 1) // fps
    pad = ng * fps - nfib
    absmax = np.abs(vals).max(axis=1)
    absmax = np.concatenate([absmax, np.zeros(pad, np.float32)]).reshape(ng, fps).max(axis=1).astype(np.float32)
    scale = np.where(absmax > 0, absmax / np.float32(448.0), np.float32(1.0)).astype(np.float32)
    inv = np.where(absmax > 0, np.float32(448.0) / np.where(absmax > 0, absmax, 1), np.float32(1.0)).astype(np.float32)
    g_of_fiber = np.arange(nfib) // fps
    codes_fib = e4m3_encode((vals * inv[g_of_fiber][:, None]).astype(np.float32))
    codes = np.zeros(a_flat.shape, dtype=np.uint8)
    codes[offs] = codes_fib
    return codes, scale


def fiber_offsets(q, na, wa, pia):
    """Element offsets of every mode-q fiber, fibers in C-offset order (dense output layout)."""
    p = len(na)
    qpos = list(pia).index(q)
    order = [pia[s] - 1 for s in range(p) if s != qpos]        # free modes, fastest first
    offs = np.zeros(1, dtype=np.int64)
    for md in order:
        offs = (offs[None, :] + (np.arange(na[md], dtype=np.int64) * wa[md])[:, None]).reshape(-1)
    kk = np.arange(na[q - 1], dtype=np.int64) * wa[q - 1]
    return len(offs), offs[:, None] + kk[None, :]


# ---- selector network (FP64 restatement of the layout in ttv_amd/csrc/mlp.hpp) ----------------
MLP_F, MLP_H, MLP_V = 24, 64, 15


def mlp_split(params):
    o = 0
    out = {}
    for name, shape in (("mean", (MLP_F,)), ("inv_std", (MLP_F,)), ("W1", (MLP_H, MLP_F)), ("b1", (MLP_H,)),
                        ("W2", (MLP_H, MLP_H)), ("b2", (MLP_H,)), ("W3", (MLP_V, MLP_H)), ("b3", (MLP_V,))):
        n = int(np.prod(shape))
        out[name] = np.asarray(params[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def mlp_forward(params, X):
    P = mlp_split(params)
    xs = (np.asarray(X, np.float64) - P["mean"]) * P["inv_std"]
    z1 = xs @ P["W1"].T + P["b1"]
    h1 = np.maximum(z1, 0)
    z2 = h1 @ P["W2"].T + P["b2"]
    h2 = np.maximum(z2, 0)
    y = h2 @ P["W3"].T + P["b3"]
    return y, (xs, z1, h1, z2, h2)


def mlp_loss_grad(params, X, T, M):
    """Masked mean squared error 0.5*sum(M*(y-T)^2)/B and its gradient w.r.t. the trainable
    parameters (everything after mean/inv_std), in the parameter-vector layout."""
    P = mlp_split(params)
    B = X.shape[0]
    y, (xs, z1, h1, z2, h2) = mlp_forward(params, X)
    d = np.asarray(M, np.float64) * (y - T) / B
    loss = 0.5 * float(np.sum(np.asarray(M, np.float64) * (y - T) ** 2)) / B
    gW3 = d.T @ h2
    gb3 = d.sum(0)
    dh2 = (d @ P["W3"]) * (z2 > 0)
    gW2 = dh2.T @ h1
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ P["W2"]) * (z1 > 0)
    gW1 = dh1.T @ xs
    gb1 = dh1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g
