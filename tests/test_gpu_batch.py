"""Batched persistent launch (ttvx_batch_*): every problem runs the tiles of its own selected
variant, so it equals a single launch forced to that variant bit for bit; results equal the FP64
oracle within 1e-5, and graph replays (with eager work and synchronisations between them)
reproduce the eager result.  The full workload is the 512 seeded problems of extents 16..512 and at
most 2^26 elements (52 GiB of A), plus problems whose rows take 1- and 2-element vectors."""
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


def _spec(na, q, pia):
    return dict(na=na, q=q, pia=pia, wa=O.strides(na, pia))


# inner extents 1, 2, 3: FP32 vectors of one and two elements (three is clamped to two)
NARROW = [_spec([1, 3000, 7], 1, [1, 2, 3]), _spec([2, 3000, 7], 2, [1, 2, 3]),
          _spec([3, 2000, 5], 2, [1, 2, 3]), _spec([3, 40000], 1, [1, 2]), _spec([2, 512, 33], 3, [3, 1, 2])]


def _kind_width(ttvx, q, A, v):
    """('col' | 'row', FP32 elements per vector load) of a problem run by variant v."""
    na, wa = list(A.shape), list(A.stride())
    info = ttvx.plan_info(q, na, wa, ttvx.layout_from_strides(na, wa))
    names = ttvx.variant_names()
    if names[v] == "col_k_lanes":
        return "row", 1
    inner = info["I"] if info["I"] > 1 else info["K"]
    return ("col" if info["I"] > 1 else "row"), min(4, 1 << (int(inner).bit_length() - 1))


def _check_against_single(ttvx, bt, probs, batched):
    """Every problem bit-identical to a single launch forced to the batch's variant for it."""
    try:
        for n, (q, A, b) in enumerate(probs):
            ttvx.set_selection(ttvx.SELECT_FORCED, bt.variant(n))
            single = _flat(ttvx.ttv(q, A, b)).cpu().numpy()
            assert np.array_equal(single, batched[n]), (n, list(A.shape), list(A.stride()), q, bt.variant(n))
    finally:
        ttvx.restore_default_selection()


def _check_oracle(q, A, b, C, got):
    """FP64 oracle in C's own strides."""
    na, wa = list(A.shape), list(A.stride())
    ref = O.naive(q, _flat(A).cpu().numpy(), na, wa, b.cpu().numpy(), wc=list(C.stride()))
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err <= 1e-5, (na, wa, q, err)


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
    assert info["problems"] == len(probs) and info["blocks"] > 0 and 1 <= info["launches"] <= len(probs)
    bt.launch(graph=False)
    torch.cuda.synchronize()
    batched = [_flat(C).cpu().numpy().copy() for C in bt.outputs]
    _check_against_single(ttvx, bt, probs, batched)
    rng = np.random.default_rng(0)
    check = set(rng.choice(len(probs), 24, replace=False).tolist()) | {len(probs) - 1, len(probs) - 2}
    for n in sorted(check):
        q, A, b = probs[n]
        _check_oracle(q, A, b, bt.outputs[n], batched[n])
    bt.close()


@pytest.mark.timeout(600)
def test_batch_full_workload(ttvx, device):
    """The 512-problem workload (seed 4) plus the narrow-vector problems: bit identity with each
    problem's own variant for all of them, the FP64 oracle on a sample of 40 that holds both tile
    kinds (column, row) and the vector widths the default selection takes (2 and 4 elements; one
    element only in col_k_lanes, covered by test_batch_every_variant), and two graph replays."""
    import torch
    from ttv_amd.batch import Batch
    from ttv_amd.workloads import batch_problems
    specs = batch_problems(seed=4, count=512) + NARROW
    probs = _make(torch, device, specs, 44)
    bt = Batch(probs)
    assert bt.info()["problems"] == len(probs)
    print("batch: %d problems, %d blocks, %d launches" % (len(probs), bt.info()["blocks"], bt.info()["launches"]),
          flush=True)
    bt.launch(graph=False)
    torch.cuda.synchronize()
    batched = [_flat(C).cpu().numpy().copy() for C in bt.outputs]
    _check_against_single(ttvx, bt, probs, batched)
    kinds = [_kind_width(ttvx, q, A, bt.variant(n)) for n, (q, A, b) in enumerate(probs)]
    rng = np.random.default_rng(4)
    sample = set(range(512, len(probs)))
    for kw in sorted(set(kinds)):               # two of every (kind, width) present
        idx = [n for n in range(len(probs)) if kinds[n] == kw]
        sample |= set(rng.choice(idx, min(2, len(idx)), replace=False).tolist())
    rest = [n for n in range(len(probs)) if n not in sample]
    sample |= set(rng.choice(rest, 40 - len(sample), replace=False).tolist())
    covered = {kinds[n] for n in sample}
    assert {k for k, _ in covered} == {"col", "row"} and {2, 4} <= {w for _, w in covered}, covered
    for n in sorted(sample):
        q, A, b = probs[n]
        _check_oracle(q, A, b, bt.outputs[n], batched[n])
    for rep in range(2):
        for C in bt.outputs:
            C.fill_(float("nan"))
        torch.cuda.synchronize()
        bt.launch(graph=True)
        torch.cuda.synchronize()
        for n, C in enumerate(bt.outputs):
            assert np.array_equal(_flat(C).cpu().numpy(), batched[n]), (rep, n)
    bt.close()


def test_batch_every_variant(ttvx, device):
    """Each unsplit variant forced in turn (the batch takes the forced variant like a single call):
    a batch of the mixed problems it can run, with inner extents 1-3 among them (vectors of 1 and 2
    elements), is bit-identical to single launches of that variant and within 1e-5 of the oracle.
    Covers every tile body of the batched kernel: column items of 1, 2 and 4 slab units, col_wide,
    col_k_lanes, 256-byte column chunks (col_narrow), one-vector rows and longer rows (dot_u1..u8)."""
    import torch
    from ttv_amd.batch import Batch
    from ttv_amd.workloads import batch_problems
    names = ttvx.variant_names()
    specs = batch_problems(seed=7, count=24, cap=1 << 20, lo=2, hi=64) + NARROW
    specs += [_spec([5, 2, 900], 2, [2, 1, 3]), _spec([3, 700, 6], 1, [1, 3, 2]), _spec([1000, 7, 3], 2, [1, 2, 3]),
              _spec([64, 9, 70], 1, [1, 2, 3]), _spec([300, 70], 2, [2, 1])]
    probs_all = _make(torch, device, specs, 77)
    ran = set()
    for v, name in enumerate(names):
        if name in ("dot_split", "col_split", "strided"):
            continue
        probs = []
        for q, A, b in probs_all:
            na, wa = list(A.shape), list(A.stride())
            if ttvx.variant_valid(q, na, wa, ttvx.layout_from_strides(na, wa), v):
                probs.append((q, A, b))
        if not probs:
            continue
        ttvx.set_selection(ttvx.SELECT_FORCED, v)
        try:
            bt = Batch(probs)
            assert all(bt.variant(n) == v for n in range(len(probs))), name
            bt.launch(graph=False)
            torch.cuda.synchronize()
            batched = [_flat(C).cpu().numpy().copy() for C in bt.outputs]
        finally:
            ttvx.restore_default_selection()
        _check_against_single(ttvx, bt, probs, batched)
        for n, (q, A, b) in enumerate(probs):
            _check_oracle(q, A, b, bt.outputs[n], batched[n])
        ran.add(name)
        bt.close()
    assert {"dot_u1", "dot_u2", "dot_u4", "dot_u8", "col_u1", "col_u2", "col_u4", "col_narrow", "col_k_lanes",
            "col_wide"} <= ran, ran


def test_batch_per_problem_variants(ttvx, device):
    """Variants fixed per problem (the counterpart of the reference's per-call policy tags): each
    problem runs the one it names (split variants in their unsplit form), bit-identical to a single
    launch forced to it; a variant that cannot run its problem is refused with the problem's index."""
    import torch
    from ttv_amd.batch import Batch
    names = ttvx.variant_names()
    specs = [_spec([300, 40, 9], 2, [1, 2, 3]), _spec([64, 9, 70], 1, [1, 2, 3]), _spec([1000, 7, 3], 1, [1, 2, 3]),
             _spec([17, 33, 21], 3, [3, 1, 2]), _spec([256, 64, 8], 2, [1, 2, 3])]
    probs = _make(torch, device, specs, 12)
    want = []
    for q, A, b in probs:
        na, wa = list(A.shape), list(A.stride())
        ok = [v for v, n in enumerate(names) if n != "strided" and
              ttvx.variant_valid(q, na, wa, ttvx.layout_from_strides(na, wa), v)]
        want.append(ok[len(want) % len(ok)])
    bt = Batch(probs, variants=want)
    got = [bt.variant(n) for n in range(len(probs))]
    unsplit = {"col_split": "col_u1", "dot_split": "dot_u2"}
    assert got == [names.index(unsplit.get(names[v], names[v])) for v in want], (got, want)
    bt.launch(graph=False)
    torch.cuda.synchronize()
    batched = [_flat(C).cpu().numpy().copy() for C in bt.outputs]
    _check_against_single(ttvx, bt, probs, batched)
    for n, (q, A, b) in enumerate(probs):
        _check_oracle(q, A, b, bt.outputs[n], batched[n])
    bt.close()
    col = probs[0]           # a column problem: no row variant can run it
    with pytest.raises(ttvx.TTVError, match="problem 1: variant dot_u2 cannot run"):
        Batch([probs[1], col], variants=[None, names.index("dot_u2")])
    with pytest.raises(ttvx.TTVError, match="problem 0: no kernel variant"):
        Batch([col], variants=[len(names)])


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
