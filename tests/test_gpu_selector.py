"""Selector network kernels on the GPU against the FP64 oracle restatement (oracle.mlp_*):
forward outputs and loss/gradient of the masked MSE; Adam decreases the loss; the GPU forward
pass and the library's host forward pass pick the same variants.  The shipped selection (tuning
table, then network) is timed against every valid variant on 48 held-out shapes."""
import numpy as np
import pytest

from oracle import oracle as O
from ttv_amd.selector import NV

pytestmark = pytest.mark.gpu


def _params(rng):
    from ttv_amd.selector import init_params
    X = rng.standard_normal((256, 24)).astype(np.float32) * 3 + 1
    p = init_params(X, seed=1)
    p[2 * 24:] += (rng.standard_normal(p.size - 48) * 0.05).astype(np.float32)   # non-zero biases too
    return p, X


def test_forward_and_gradient_vs_oracle(ttvx, device):
    import torch
    from ttv_amd.selector import GPUTrainer
    rng = np.random.default_rng(0)
    params, X = _params(rng)
    T = rng.standard_normal((256, NV)).astype(np.float32)
    M = (rng.uniform(size=(256, NV)) > 0.3).astype(np.float32)
    tr = GPUTrainer(params, 256)
    Xd, Td, Md = (torch.from_numpy(v).to(device) for v in (X, T, M))
    y = tr.forward(Xd).cpu().numpy()
    y_ref, _ = O.mlp_forward(params, X)
    assert np.max(np.abs(y - y_ref)) <= 1e-4 * max(1.0, np.max(np.abs(y_ref)))
    g, loss = tr.grad(Xd, Td, Md)
    l_ref, g_ref = O.mlp_loss_grad(params, X, T, M)
    g = g.cpu().numpy()
    assert abs(float(loss.item()) - l_ref) <= 1e-5 * max(1.0, l_ref)
    assert np.linalg.norm(g - g_ref) <= 1e-4 * np.linalg.norm(g_ref)


def test_adam_reduces_loss(ttvx, device):
    import torch
    from ttv_amd.selector import GPUTrainer
    rng = np.random.default_rng(1)
    params, X = _params(rng)
    W = rng.standard_normal((24, NV)).astype(np.float32)
    T = np.maximum(X @ W * 0.1, 0).astype(np.float32)
    M = np.ones_like(T)
    tr = GPUTrainer(params, 256)
    Xd, Td, Md = (torch.from_numpy(v).to(device) for v in (X, T, M))
    _, l0 = tr.grad(Xd, Td, Md)
    l0 = float(l0.item())
    for _ in range(300):
        tr.step(Xd, Td, Md, 3e-3)
    _, l1 = tr.grad(Xd, Td, Md)
    assert float(l1.item()) < 0.2 * l0


def test_gpu_and_host_inference_agree(ttvx, device):
    from ttv_amd.selector import predict_gpu
    from ttv_amd.workloads import selector_shapes
    rng = np.random.default_rng(2)
    params, _ = _params(rng)
    shapes = selector_shapes(seed=3, count=200)
    feats = np.stack([ttvx.features(s["q"], s["na"], s["wa"], s["pia"], ttvx.F32) for s in shapes])
    valid = np.array([[1.0 if ttvx.variant_valid(s["q"], s["na"], s["wa"], s["pia"], v) else np.nan
                       for v in range(NV)] for s in shapes])
    choice, _ = predict_gpu(params, feats, valid)
    ttvx.set_selector_weights(params)
    ttvx.set_selection(ttvx.SELECT_LEARNED)
    try:
        host = [ttvx.plan_info(s["q"], s["na"], s["wa"], s["pia"])["variant"] for s in shapes]
    finally:
        ttvx.restore_default_selection()
    agree = np.mean(np.array(host) == choice)
    assert agree >= 0.99, agree


def test_selection_regret_on_held_out_shapes(ttvx, device):
    """48 shapes of the held-out set (seed 202; orders 2-10, 2^19..2^28 elements, all four storage
    types), every valid variant timed HBM-cold on this GPU: the variants the default selection
    picks are on average within 5% of the fastest, and no slower than the fixed rule's picks."""
    from ttv_amd.selector import DT_CODE, collect_timings, regret
    from ttv_amd.workloads import selector_shapes
    shapes = selector_shapes(seed=202, count=48)
    ttvx.restore_default_selection()
    picks = np.array([ttvx.plan_info(s["q"], s["na"], s["wa"], s["pia"], DT_CODE[s["dtype"]])["variant"]
                      for s in shapes])
    _, times, heur = collect_timings(shapes, reps=5, log=None)
    ttvx.restore_default_selection()
    r_sel, r_fixed = regret(picks, times), regret(heur, times)
    print("held-out regret: selection %.4f, fixed rule %.4f" % (r_sel, r_fixed), flush=True)
    assert np.isfinite(r_sel) and r_sel <= 0.05, (r_sel, r_fixed)
    assert r_sel <= r_fixed, (r_sel, r_fixed)
