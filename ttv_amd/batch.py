"""Batched mode-q products, one launch per group of problems sharing a tile body (ttvx_batch_* in ttvx.h).

    B = Batch([(q, A, b), ...])      # torch float32 GPU tensors; C tensors are allocated
    B.launch(graph=True)             # one launch per tile body (captured into a hipGraph on first use)
    B.outputs                        # the C tensors, in problem order
"""
import ctypes

from . import _lib
from ._lib import BatchProblem, check, lib, sizes
from .meta import generate_output_layout, generate_output_shape, generate_strides, layout_from_strides


class Batch:
    def __init__(self, problems, outputs=None, variants=None):
        """problems: (q, A, b) triples; outputs: C tensors to write (allocated when None);
        variants: per-problem kernel variant (None or -1 entries: the library's selection)."""
        import torch
        self._keep = []
        arr = (BatchProblem * len(problems))()
        self.outputs = []
        for n, (q, A, b) in enumerate(problems):
            if not (A.is_cuda and A.dtype == torch.float32 and b.is_cuda and b.dtype == torch.float32):
                raise _lib.TTVError(_lib.ERR_ARGUMENT, "ttv_amd.Batch: float32 GPU tensors required")
            na, wa = list(A.shape), list(A.stride())
            pia = layout_from_strides(na, wa)
            if outputs is None:
                nc = generate_output_shape(na, q)
                wc = generate_strides(nc, generate_output_layout(pia, q))
                C = torch.empty_strided(nc, wc, dtype=torch.float32, device=A.device)
            else:
                C = outputs[n]
                wc = list(C.stride())
            self.outputs.append(C)
            sa, sw, sp, sc = sizes(na), sizes(wa), sizes(pia), sizes(wc)
            self._keep += [A, b, C, sa, sw, sp, sc]
            v = -1 if variants is None or variants[n] is None else int(variants[n])
            arr[n] = BatchProblem(q, len(na), A.data_ptr(), sa, sw, sp, b.data_ptr(), C.data_ptr(), sc, v)
        self._arr = arr
        h = ctypes.c_void_p()
        check(lib().ttvx_batch_create(len(problems), ctypes.cast(arr, ctypes.c_void_p), ctypes.byref(h)))
        self._h = h

    def launch(self, graph=False):
        import torch
        check(lib().ttvx_batch_launch(self._h, 1 if graph else 0,
                                      ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))

    def info(self):
        out = (ctypes.c_int64 * 4)()
        check(lib().ttvx_batch_info(self._h, out))
        return dict(blocks=out[0], bytes=out[1], launches=out[2], problems=out[3])

    def variant(self, index):
        """Kernel variant running problem `index` (a single call forced to it is bit-identical)."""
        v = lib().ttvx_batch_variant(self._h, index)
        if v < 0:
            check(_lib.ERR_ARGUMENT)
        return v

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().ttvx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
