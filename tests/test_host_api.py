"""Host-side argument checks of the Python front end (no GPU needed: they raise before any
library call)."""
import numpy as np
import pytest


def test_ttvs_argument_errors():
    """The binding's messages for ttvs (ttvpy/src/wrapped_ttv.cpp:96-125)."""
    import ttv_amd
    A = np.zeros((3, 2, 4))
    with pytest.raises(ValueError, match="multiplication order"):
        ttv_amd.ttvs(1, A, [np.zeros(2), np.zeros(4)], "sideways")
    with pytest.raises(ValueError, match="number of input vectors"):
        ttv_amd.ttvs(1, A, [np.zeros(2)])
    with pytest.raises(ValueError, match="contraction mode"):
        ttv_amd.ttvs(4, A, [np.zeros(2), np.zeros(4)])
    with pytest.raises(ValueError, match="some of the input vectors"):
        ttv_amd.ttvs(1, A, [np.zeros((2, 1)), np.zeros(4)])
    with pytest.raises(ValueError, match="not compatible"):
        ttv_amd.ttvs(1, A, [np.zeros(3), np.zeros(4)])


def test_ttv_rejects_mixed_dtypes():
    import ttv_amd
    with pytest.raises(ttv_amd.TTVError):
        ttv_amd.ttv(1, np.zeros((2, 3), dtype=np.float32), np.zeros(2, dtype=np.float64))


def test_ttv_rejects_wrong_outputs_and_vectors():
    """A given `out` must have C's shape and A's element type; b must be a vector."""
    import ttv_amd
    A = np.zeros((2, 3, 4), dtype=np.float32)
    b = np.zeros(3, dtype=np.float32)
    with pytest.raises(ttv_amd.TTVError, match="out must be float32"):
        ttv_amd.ttv(2, A, b, out=np.zeros((2, 4), dtype=np.float64))
    with pytest.raises(ttv_amd.TTVError, match="out must have shape"):
        ttv_amd.ttv(2, A, b, out=np.zeros((4, 2), dtype=np.float32))
    with pytest.raises(ttv_amd.TTVError, match="1-D vector"):
        ttv_amd.ttv(2, A, np.zeros((3, 1), dtype=np.float32))


def test_reduced_precision_entry_points_check_operands():
    """ttv_bf16 / ttv_fp8 take GPU tensors of the storage type; anything else raises before a
    library call."""
    import torch

    import ttv_amd
    a16 = torch.zeros((2, 3, 4), dtype=torch.uint16)
    b = torch.zeros(3)
    with pytest.raises(ttv_amd.TTVError, match="uint16 GPU tensor"):
        ttv_amd.ttv_bf16(2, a16, b)
    with pytest.raises(ttv_amd.TTVError, match="uint16 GPU tensor"):
        ttv_amd.ttv_bf16(2, torch.zeros((2, 3, 4)), b)
    with pytest.raises(ttv_amd.TTVError, match="uint8 GPU tensor"):
        ttv_amd.ttv_fp8(2, a16, torch.ones(1), 8, b)


def test_fault_hook_arguments():
    """ttvx_debug_fail_next validates its site and count without a GPU; arming and disarming a
    valid site succeeds."""
    import ttv_amd
    from ttv_amd import _lib
    L = ttv_amd.lib()
    assert L.ttvx_debug_fail_next(3, 1) == _lib.ERR_ARGUMENT
    assert L.ttvx_debug_fail_next(-1, 1) == _lib.ERR_ARGUMENT
    assert L.ttvx_debug_fail_next(_lib.FAULT_COPY, -2) == _lib.ERR_ARGUMENT
    assert b"TTVX_FAULT" in L.ttvx_last_error()
    for site in (_lib.FAULT_ALLOC, _lib.FAULT_COPY, _lib.FAULT_LAUNCH):
        assert L.ttvx_debug_fail_next(site, 2) == _lib.OK
        assert L.ttvx_debug_fail_next(site, 0) == _lib.OK
