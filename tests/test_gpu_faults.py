"""Error paths of the library on a healthy GPU, forced from the host with the fault-injection hook
(ttvx_debug_fail_next): an armed allocation, copy or launch failure must surface as TTVError with
HIP's message, write nothing to C, and leave the library working (the next call succeeds and is
correct).  The reference reports errors by exception (include/tlib/ttv.h:64-89); the C++ drop-in
header rethrows these statuses the same way."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture
def armed(ttvx):
    L = ttvx.lib()

    def arm(site, count=1):
        ttvx.check(L.ttvx_debug_fail_next(site, count))
    yield arm
    for site in (ttvx._lib.FAULT_ALLOC, ttvx._lib.FAULT_COPY, ttvx._lib.FAULT_LAUNCH):
        L.ttvx_debug_fail_next(site, 0)


def _problem(rng, na=(37, 21, 9), q=2):
    a = rng.uniform(-1, 1, int(np.prod(na))).astype(np.float32)
    A = a.reshape(na[::-1]).transpose()          # first-order (Fortran) layout
    b = rng.uniform(-1, 1, na[q - 1]).astype(np.float32)
    ref = O.naive(q, a, list(na), O.strides(list(na), [1, 2, 3]), b, pia=[1, 2, 3])
    return A, b, ref


def _close(got, ref):
    got = np.asarray(got, np.float64).reshape(-1, order="F")
    return np.linalg.norm(got - ref) <= 1e-5 * np.linalg.norm(ref)


def test_host_buffer_allocation_and_copy_failures(ttvx, device, armed):
    rng = np.random.default_rng(1)
    A, b, ref = _problem(rng)
    for site, text in ((ttvx._lib.FAULT_ALLOC, "out of memory"), (ttvx._lib.FAULT_COPY, "hipMemcpy")):
        armed(site)
        with pytest.raises(ttvx.TTVError, match=text) as e:
            ttvx.ttv(2, A, b)
        assert e.value.status == ttvx._lib.ERR_DEVICE
        assert _close(ttvx.ttv(2, A, b), ref), site          # the next call works


def test_launch_failure_leaves_output_untouched(ttvx, device, armed):
    import torch
    g = torch.Generator(device="cpu").manual_seed(2)
    A = (torch.rand(37, 21, 9, generator=g) * 2 - 1).to(device)
    b = (torch.rand(21, generator=g) * 2 - 1).to(device)
    want = torch.einsum("ijk,j->ik", A.double(), b.double())
    C = ttvx.ttv(2, A, b)
    C.fill_(-7.0)
    armed(ttvx._lib.FAULT_LAUNCH)
    with pytest.raises(ttvx.TTVError, match="launch") as e:
        ttvx.ttv(2, A, b, out=C)
    assert e.value.status == ttvx._lib.ERR_DEVICE
    torch.cuda.synchronize()
    assert torch.all(C == -7.0)
    ttvx.ttv(2, A, b, out=C)
    torch.cuda.synchronize()
    assert torch.linalg.vector_norm(C.double() - want) <= 1e-5 * torch.linalg.vector_norm(want)


def test_split_partials_allocation_failure(ttvx, device, armed):
    """A split contraction needs the per-stream partials buffer: a failed allocation reports the
    error and the next launch (which grows the buffer) is correct."""
    import torch
    names = ttvx.variant_names()
    na, q = [64, 1 << 16, 2], 2                 # I = 64, K = 65536, O = 2: few items, long contraction
    pia = [1, 2, 3]
    wa = O.strides(na, pia)
    if not ttvx.variant_valid(q, na, wa, pia, names.index("col_split")):
        pytest.skip("col_split does not apply to this shape on this device")
    rng = np.random.default_rng(3)
    a = rng.uniform(-1, 1, int(np.prod(na))).astype(np.float32)
    A = torch.from_numpy(a).to(device).as_strided(na, wa)
    b = torch.from_numpy(rng.uniform(-1, 1, na[1]).astype(np.float32)).to(device)
    ref = O.naive(q, a, na, wa, b.cpu().numpy(), pia=pia)
    ttvx.set_selection(ttvx.SELECT_FORCED, names.index("col_split"))
    try:
        with torch.cuda.stream(torch.cuda.Stream()):           # a stream with no partials buffer yet
            armed(ttvx._lib.FAULT_ALLOC)
            with pytest.raises(ttvx.TTVError, match="split partials"):
                ttvx.ttv(q, A, b)
            C = ttvx.ttv(q, A, b)
            torch.cuda.current_stream().synchronize()
    finally:
        ttvx.restore_default_selection()
    got = C.as_strided([C.numel()], [1]).cpu().numpy()
    assert np.linalg.norm(got - ref) <= 1e-5 * np.linalg.norm(ref)


def test_batch_and_device_allocation_failures(ttvx, device, armed):
    import ctypes
    import torch
    from ttv_amd.batch import Batch
    A = torch.rand(30, 20, 10, device=device)
    b = torch.rand(20, device=device)
    armed(ttvx._lib.FAULT_ALLOC)
    with pytest.raises(ttvx.TTVError, match="batch descriptors"):
        Batch([(2, A, b)])
    bt = Batch([(2, A, b)])
    bt.launch()
    torch.cuda.synchronize()
    want = torch.einsum("ijk,j->ik", A.double(), b.double())
    assert torch.allclose(bt.outputs[0].double(), want, rtol=1e-5, atol=1e-6)
    bt.close()
    L = ttvx.lib()
    p = ctypes.c_void_p()
    armed(ttvx._lib.FAULT_ALLOC)
    assert L.ttvx_device_alloc(1 << 20, ctypes.byref(p)) == ttvx._lib.ERR_DEVICE and not p.value
    assert L.ttvx_device_alloc(1 << 20, ctypes.byref(p)) == ttvx._lib.OK and p.value
    assert L.ttvx_device_free(p) == ttvx._lib.OK
