"""Batched persistent launch (ttvx_batch_*): every problem equals its single-launch result bit for
bit (same tile bodies and configurations), equals the FP64 oracle within 1e-5, and graph
replays (with eager work and synchronisations between them) reproduce the eager result."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _flat(C):
    span = 1 + sum((n - 1) * w for n, w in zip(C.shape, C.stride()))
    return C.as_strided([span], [1])


def _make(torch, device, specs, seed):
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    probs = []
    for s in specs:
        span = O.span(s["na"], s["wa"])
        A = torch.empty(span, dtype=torch.float32, device=device).uniform_(-1, 1, generator=g)
        b = torch.empty(s["na"][s["q"] - 1], dtype=torch.float32, device=device).uniform_(-1, 1, generator=g)
        probs.append((s["q"], A.as_strided(s["na"], s["wa"]), b))
    return probs


def test_batch_matches_single_launches_and_oracle(ttvx, device):
    import torch
    from ttv_amd.batch import Batch
    from ttv_amd.workloads import batch_problems
    specs = batch_problems(seed=5, count=96, cap=1 << 21, lo=2, hi=160)
    specs += [dict(na=[37, 5, 3], q=2, pia=[1, 2, 3], wa=O.strides([37, 5, 3], [1, 2, 3])),
              dict(na=[300, 7], q=1, pia=[1, 2], wa=O.strides([300, 7], [1, 2])),
              dict(na=[3, 4096], q=2, pia=[2, 1], wa=O.strides([3, 4096], [2, 1]))]
    probs = _make(torch, device, specs, 6)
    bt = Batch(probs)
    info = bt.info()
    assert info["problems"] == len(probs) and info["tiles"] > 0
    bt.launch(graph=False)
    torch.cuda.synchronize()
    batched = [_flat(C).cpu().numpy().copy() for C in bt.outputs]
    try:
        rng = np.random.default_rng(0)
        check = set(rng.choice(len(probs), 24, replace=False).tolist()) | {len(probs) - 1, len(probs) - 2}
        for n, (q, A, b) in enumerate(probs):
            na, wa = list(A.shape), list(A.stride())
            pia = ttvx.layout_from_strides(na, wa)
            info = ttvx.plan_info(q, na, wa, pia)
            v = 6 if info["I"] > 1 else 2          # the batch's configurations: col_u2 / dot_u4
            ttvx.set_selection(ttvx.SELECT_FORCED, v)
            single = _flat(ttvx.ttv(q, A, b)).cpu().numpy()
            assert np.array_equal(single, batched[n]), (n, specs[n])
            if n in check:
                ref = O.naive(q, _flat(A).cpu().numpy(), na, wa, b.cpu().numpy(), pia=pia)
                err = np.linalg.norm(batched[n] - ref) / np.linalg.norm(ref)
                assert err <= 1e-5, (n, specs[n], err)
    finally:
        ttvx.restore_default_selection()
    bt.close()


def test_batch_graph_replays(ttvx, device):
    import torch
    from ttv_amd.batch import Batch
    from ttv_amd.workloads import batch_problems
    specs = batch_problems(seed=9, count=40, cap=1 << 20, lo=4, hi=120)
    probs = _make(torch, device, specs, 10)
    bt = Batch(probs)
    bt.launch(graph=False)
    torch.cuda.synchronize()
    eager = [_flat(C).cpu().numpy().copy() for C in bt.outputs]
    for rep in range(4):
        for C in bt.outputs:
            C.fill_(float("nan"))            # eager work between replays
        torch.cuda.synchronize()
        bt.launch(graph=True)
        torch.cuda.synchronize()
        for n, C in enumerate(bt.outputs):
            assert np.array_equal(_flat(C).cpu().numpy(), eager[n]), (rep, n)
    bt.close()


def test_batch_rejects_bad_problems(ttvx, device):
    import torch
    from ttv_amd.batch import Batch
    A = torch.zeros(4, 5, device=device)
    with pytest.raises(ttvx.TTVError, match="problem 0"):
        Batch([(3, A, torch.zeros(5, device=device))])
