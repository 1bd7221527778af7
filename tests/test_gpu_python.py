"""Python front end (ttv_amd.ttv / ttv_amd.ttvs), written like the reference binding's tests
ttvpy/tests/test.py:5-80: integer-valued float64 arrays, results compared exactly with numpy.einsum
chains (every value is an integer below 2^53, so every summation order gives the same double).
The same cases run on torch GPU tensors (asynchronous entry points) and on non-C-order views,
which ttvpy cannot take (wrapped_ttv.cpp:44-45 forces the last-order layout)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

A3 = (3, 2, 4)
EINSUM3 = {1: "ijk,i->jk", 2: "ijk,j->ik", 3: "ijk,k->ij"}


@pytest.mark.parametrize("q", [1, 2, 3])
def test_ttv_modes_numpy(ttvx, q):
    """ttvpy/tests/test.py:7-27."""
    import ttv_amd
    A = np.arange(np.prod(A3), dtype=np.float64).reshape(A3)
    b = np.arange(A3[q - 1], dtype=np.float64)
    C = ttv_amd.ttv(q, A, b)
    assert np.array_equal(C, np.einsum(EINSUM3[q], A, b))


@pytest.mark.parametrize("q", [1, 2, 3])
def test_ttv_modes_torch_and_views(ttvx, device, q):
    import torch
    import ttv_amd
    A = np.arange(np.prod(A3), dtype=np.float64).reshape(A3)
    b = np.arange(A3[q - 1], dtype=np.float64)
    want = np.einsum(EINSUM3[q], A, b)
    for dt in (torch.float64, torch.float32):
        At = torch.tensor(A, dtype=dt, device=device)
        bt = torch.tensor(b, dtype=dt, device=device)
        C = ttv_amd.ttv(q, At, bt)
        torch.cuda.synchronize()
        assert np.array_equal(C.cpu().double().numpy(), want)
    # Fortran-order and permuted numpy views: same values, other layouts
    for view in (np.asfortranarray(A), A.transpose(2, 0, 1).copy().transpose(1, 2, 0)):
        assert np.array_equal(ttv_amd.ttv(q, view, b), want)


def _chain(A, q, B):
    """The einsum chain of ttvpy/tests/test.py:32-80: contract every mode but q, lowest first."""
    letters = "ijklm"[:A.ndim]
    D, cur = A, letters
    free = [m for m in range(A.ndim) if m != q - 1]
    for vec, m in zip(B, free):
        ch = letters[m]
        nxt = cur.replace(ch, "")
        D = np.einsum("%s,%s->%s" % (cur, ch, nxt), D, vec)
        cur = nxt
    return D


@pytest.mark.parametrize("q", [1, 2, 3, 4])
@pytest.mark.parametrize("shape", [(3, 2, 4, 5), (5, 5, 5, 5)])
def test_ttvs_every_order(ttvx, q, shape):
    """ttvpy/tests/test.py:30-80 (the 5^4 case of test_ttvs_mode1 extended to every q)."""
    import ttv_amd
    A = np.arange(np.prod(shape), dtype=np.float64).reshape(shape)
    B = [np.arange(shape[m], dtype=np.float64) for m in range(4) if m != q - 1]
    want = _chain(A, q, B)
    for order in ("forward", "backward", "optimal"):
        C = ttv_amd.ttvs(q, A, B, order)
        assert np.array_equal(np.asarray(C), want), order


def test_ttvs_torch_orders_agree(ttvx, device):
    """Random float64 GPU data: the three orders agree to rounding and match the einsum chain."""
    import torch
    import ttv_amd
    g = torch.Generator(device="cpu").manual_seed(7)
    shape = (17, 9, 33, 12, 5)
    A = torch.rand(shape, generator=g, dtype=torch.float64) * 2 - 1
    for q in (1, 3, 5):
        B = [torch.rand(shape[m], generator=g, dtype=torch.float64) * 2 - 1 for m in range(5) if m != q - 1]
        want = _chain(A.numpy(), q, [b.numpy() for b in B])
        for order in ("forward", "backward", "optimal"):
            C = ttv_amd.ttvs(q, A.to(device), [b.to(device) for b in B], order)
            torch.cuda.synchronize()
            got = C.cpu().numpy()
            assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want), (q, order)


def test_ttvs_fused(ttvx, device, monkeypatch):
    """order="fused" (order >= 5 with q stored in the middle: one row and one column product with
    Kronecker weights; otherwise the optimal chain): exact on integer data for orders 2-5, every q,
    first-order, last-order and a permuted layout, FP32 and FP64; within 1e-12 (FP64) of the einsum
    chain on random data; inputs it cannot fuse (a non-dense view, numpy arrays) take the optimal
    chain with the same result."""
    import torch
    import ttv_amd
    from oracle import oracle as O
    fused = []
    kron = ttv_amd._kron
    monkeypatch.setattr(ttv_amd, "_kron", lambda v, like: fused.append(len(v)) or kron(v, like))
    rng = np.random.default_rng(11)
    for shape in ((7, 5), (3, 4, 5), (4, 3, 5, 2), (3, 2, 4, 2, 3), (2, 3, 2, 4, 3)):
        p = len(shape)
        layouts = [list(range(1, p + 1)), list(range(p, 0, -1)), [int(x) + 1 for x in rng.permutation(p)]]
        for pia in layouts:
            wa = O.strides(list(shape), pia)
            span = O.span(list(shape), wa)
            for q in range(1, p + 1):
                a = rng.integers(-3, 4, span).astype(np.float64)
                B = [rng.integers(-3, 4, shape[m]).astype(np.float64) for m in range(p) if m != q - 1]
                An = np.lib.stride_tricks.as_strided(a, shape=shape, strides=[w * 8 for w in wa])
                want = _chain(np.array(An), q, B)
                for dt in (torch.float32, torch.float64):
                    A = torch.from_numpy(a).to(device=device, dtype=dt).as_strided(shape, wa)
                    C = ttv_amd.ttvs(q, A, [torch.from_numpy(b).to(device=device, dtype=dt) for b in B], "fused")
                    torch.cuda.synchronize()
                    assert np.array_equal(C.cpu().numpy().astype(np.float64), want), (shape, pia, q, dt)
    assert len(fused) >= 2 * 2 * 3 * 3, len(fused)     # the order-5 problems with q in the middle fused
    # random FP64 data
    shape = (17, 9, 33, 12, 5)
    A = torch.rand(shape, dtype=torch.float64) * 2 - 1
    for q in (1, 3, 5):
        B = [torch.rand(shape[m], dtype=torch.float64) * 2 - 1 for m in range(5) if m != q - 1]
        want = _chain(A.numpy(), q, [b.numpy() for b in B])
        C = ttv_amd.ttvs(q, A.to(device), [b.to(device) for b in B], "fused").cpu().numpy()
        assert np.linalg.norm(C - want) <= 1e-12 * np.linalg.norm(want), q
    # a non-dense view and numpy inputs take the chain
    big = torch.arange(6 * 8 * 5, dtype=torch.float64, device=device).reshape(6, 8, 5)
    V = big[1:5, ::2, :]
    B = [torch.arange(n, dtype=torch.float64, device=device) for n in (4, 5)]
    want = _chain(V.cpu().numpy(), 2, [b.cpu().numpy() for b in B])
    assert np.array_equal(ttv_amd.ttvs(2, V, B, "fused").cpu().numpy(), want)
    An = np.arange(60, dtype=np.float64).reshape(3, 4, 5)
    Bn = [np.arange(3.0), np.arange(5.0)]
    assert np.array_equal(np.asarray(ttv_amd.ttvs(2, An, Bn, "fused")), _chain(An, 2, Bn))
