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
