"""ctypes binding of libttvx.so (include/ttvx.h).

The library must have been built in-tree (``python -m ttv_amd.build``); there is no fallback:
importing a function whose library is missing raises ``RuntimeError``.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libttvx.so")

# element types / flags / selection modes (ttvx.h)
F32, F64, BF16, FP8E4M3 = 0, 1, 2, 3
FLAG_KAHAN, FLAG_FP64_FINAL = 1, 2
SELECT_HEURISTIC, SELECT_LEARNED, SELECT_FORCED, SELECT_AUTO = 0, 1, 2, 3
NUM_FEATURES = 24

# error codes
OK, ERR_ARGUMENT, ERR_LAYOUT, ERR_DEVICE, ERR_UNSUPPORTED = 0, 1, 2, 3, 4
FAULT_ALLOC, FAULT_COPY, FAULT_LAUNCH = 0, 1, 2        # ttvx_debug_fail_next sites

_sz = ctypes.c_size_t
_szp = ctypes.POINTER(ctypes.c_size_t)
_vp = ctypes.c_void_p
_int = ctypes.c_int
_uint = ctypes.c_uint

_SIGS = {
    "ttvx_last_error": ([], ctypes.c_char_p),
    "ttvx_version": ([], _int),
    "ttvx_device_count": ([], _int),
    "ttvx_device_alloc": ([_sz, ctypes.POINTER(ctypes.c_void_p)], _int),
    "ttvx_device_free": ([_vp], _int),
    "ttvx_copy": ([_vp, _vp, _sz], _int),
    "ttvx_ttv_f32": ([_int, _int, _int, _sz, _sz, _vp, _szp, _szp, _szp, _vp, _szp, _vp, _szp, _szp, _szp], _int),
    "ttvx_ttv_f64": ([_int, _int, _int, _sz, _sz, _vp, _szp, _szp, _szp, _vp, _szp, _vp, _szp, _szp, _szp], _int),
    "ttvx_ttv_f32_async": ([_sz, _sz, _vp, _szp, _szp, _szp, _vp, _szp, _vp, _szp, _szp, _szp, _uint, _vp], _int),
    "ttvx_ttv_f64_async": ([_sz, _sz, _vp, _szp, _szp, _szp, _vp, _szp, _vp, _szp, _szp, _szp, _uint, _vp], _int),
    "ttvx_ttv_f32_async_noalloc": ([_sz, _sz, _vp, _szp, _szp, _szp, _vp, _szp, _vp, _szp, _szp, _szp, _uint, _vp],
                                   _int),
    "ttvx_mtv_f32": ([_vp, _vp, _vp, _sz, _sz, _sz, _sz, _int], _int),
    "ttvx_mtv_f64": ([_vp, _vp, _vp, _sz, _sz, _sz, _sz, _int], _int),
    "ttvx_quantize_bf16_async": ([_sz, _vp, _vp, _vp], _int),
    "ttvx_fp8_num_scales": ([_sz, _sz, _szp, _szp, _szp, _sz], _sz),
    "ttvx_quantize_fp8_async": ([_sz, _sz, _vp, _szp, _szp, _szp, _vp, _vp, _sz, _vp], _int),
    "ttvx_ttv_bf16_async": ([_sz, _sz, _vp, _szp, _szp, _szp, _vp, _szp, _vp, _szp, _szp, _szp, _uint, _vp], _int),
    "ttvx_ttv_fp8_async": ([_sz, _sz, _vp, _szp, _szp, _szp, _vp, _sz, _vp, _szp, _vp, _szp, _szp, _szp, _uint, _vp],
                           _int),
    "ttvx_num_variants": ([], _int),
    "ttvx_variant_name": ([_int], ctypes.c_char_p),
    "ttvx_set_selection": ([_int, _int], _int),
    "ttvx_plan_info": ([_int, _sz, _sz, _szp, _szp, _szp, ctypes.POINTER(ctypes.c_int64)], _int),
    "ttvx_variant_valid": ([_int, _sz, _sz, _szp, _szp, _szp, _int], _int),
    "ttvx_features": ([_int, _sz, _sz, _szp, _szp, _szp, ctypes.POINTER(ctypes.c_float)], _int),
    "ttvx_selector_num_params": ([], _sz),
    "ttvx_selector_set_weights": ([ctypes.POINTER(ctypes.c_float), _sz], _int),
    "ttvx_selector_set_table": ([ctypes.POINTER(ctypes.c_int64), _sz], _int),
    "ttvx_bench_read": ([_vp, _sz, _vp, _sz, _vp], _int),
    "ttvx_bench_copy": ([_vp, _vp, _sz, _vp], _int),
    "ttvx_bench_mix": ([_vp, _sz, _vp, _sz, _vp], _int),
    "ttvx_bench_sweep": ([_vp, _sz, _vp, _sz, _int, _vp, _vp], _int),
    "ttvx_kron_f32": ([_vp, _vp, _sz, _vp, _vp], _int),
    "ttvx_debug_fail_next": ([_int, _int], _int),
    "ttvx_kron_f64": ([_vp, _vp, _sz, _vp, _vp], _int),
    "ttvx_bench_spin": ([_int, _vp, _vp], _int),
    "ttvx_mlp_work_floats": ([_int], _sz),
    "ttvx_mlp_forward": ([_vp, _vp, _int, _vp, _vp, _vp], _int),
    "ttvx_mlp_backward": ([_vp, _vp, _vp, _vp, _int, _vp, _vp, _vp, _vp], _int),
    "ttvx_mlp_adam": ([_vp, _vp, _vp, _vp, _int, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                       ctypes.c_float, _vp], _int),
    "ttvx_batch_create": ([_sz, _vp, ctypes.POINTER(ctypes.c_void_p)], _int),
    "ttvx_batch_launch": ([_vp, _int, _vp], _int),
    "ttvx_batch_info": ([_vp, ctypes.POINTER(ctypes.c_int64)], _int),
    "ttvx_batch_variant": ([_vp, _sz], _int),
    "ttvx_batch_destroy": ([_vp], _int),
}


class BatchProblem(ctypes.Structure):
    """ttvx_batch_problem (ttvx.h)."""
    _fields_ = [("q", ctypes.c_size_t), ("p", ctypes.c_size_t), ("a", ctypes.c_void_p), ("na", _szp),
                ("wa", _szp), ("pia", _szp), ("b", ctypes.c_void_p), ("c", ctypes.c_void_p), ("wc", _szp),
                ("variant", ctypes.c_int)]

_lib = None


def lib():
    """The loaded libttvx.so with typed signatures (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libttvx.so is not built; run `python -m ttv_amd.build` (no CPU fallback exists)")
        # torch ships its own HIP runtime with the same SONAME (libamdhip64.so.7); loading torch
        # first makes libttvx bind to that one, so both share one runtime and one set of pointers.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def declared_symbols():
    """Every function name declared in include/*.h (parsed from the header text)."""
    import re
    inc = os.path.join(os.path.dirname(HERE), "include")
    names = []
    for fn in sorted(os.listdir(inc)):
        if not fn.endswith(".h"):
            continue
        txt = open(os.path.join(inc, fn)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(ttvx_[A-Za-z0-9_]+)\s*\(", txt, re.M):
            names.append(m.group(1))
    return sorted(set(names))


class TTVError(RuntimeError):
    """Raised with the library's message (the reference's std::runtime_error text)."""

    def __init__(self, status, message):
        super().__init__(message)
        self.status = status


def check(status):
    if status != OK:
        raise TTVError(status, lib().ttvx_last_error().decode())


def sizes(v):
    a = (ctypes.c_size_t * max(1, len(v)))()
    for i, x in enumerate(v):
        a[i] = int(x)
    return a
