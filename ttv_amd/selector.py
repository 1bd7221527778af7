"""Learned kernel selector: timing collection, GPU training and evaluation.

The network (layout in ttvx.h / csrc/mlp.hpp) maps the 24 features of a problem
(ttvx_features) to the predicted log2 slowdown of each of the 15 kernel variants relative to the
fastest; the planner takes the valid variant with the smallest prediction.  The exhaustive
autotuner is `collect_timings`: it times every valid variant of a shape.  Training runs on the
GPU through the library's own kernels (ttvx_mlp_forward / _backward / _adam).

Regret of a choice = t(chosen) / t(fastest) - 1, averaged over shapes.

Timing method: each shape's timed launches are enqueued behind a bounded busy-wait kernel, so the
host has queued them all before the GPU reaches them and every event interval holds kernel time
only; consecutive launches read disjoint regions of a 2 GiB pool, so no launch finds its operand
in the 256 MB Infinity Cache (HBM-cold timing).
"""
import ctypes
import os
import time

import numpy as np

from . import _lib
from ._lib import check, lib
from .meta import generate_output_layout, generate_output_shape, generate_strides

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS_PATH = os.path.join(HERE, "data", "selector_gfx950.npz")
WEIGHTS_BIN = os.path.join(HERE, "data", "selector_gfx950.bin")     # raw float32, loaded by libttvx itself
TABLE_BIN = os.path.join(HERE, "data", "selector_table_gfx950.bin")  # raw int64 rows, loaded by libttvx itself
STRUCT_WEIGHT = 8     # the benchmark shapes (equal extents, order 8) count this many times in the training loss
NV = 15
NF = 24
DT_CODE = {"f32": _lib.F32, "bf16": _lib.BF16, "f64": _lib.F64, "fp8": _lib.FP8E4M3}
DT_NAME = {v: k for k, v in DT_CODE.items()}
ELEM = {"f32": 4, "bf16": 2, "f64": 8, "fp8": 1}


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


# ---- exhaustive timing ---------------------------------------------------------------------

class TimingPool:
    def __init__(self, pool_bytes=2 << 30, b_elems=1 << 28, c_elems=1 << 28, seed=0):
        import torch
        dev = torch.device("cuda:0")
        g = torch.Generator(device=dev)
        g.manual_seed(seed)
        self.pool_bytes = pool_bytes
        self.A = torch.empty(pool_bytes // 4, dtype=torch.float32, device=dev).uniform_(-1, 1, generator=g)
        # bf16 view: the upper halves of the floats are valid bfloat16 values in [-1, 1]; the FP64
        # and FP8 views reinterpret the same bytes (the time of a product does not depend on values)
        self.A16 = self.A.view(torch.uint16)
        self.A64 = self.A.view(torch.float64)
        self.A8 = self.A.view(torch.uint8)
        self.b = torch.empty(b_elems, dtype=torch.float32, device=dev).uniform_(-1, 1, generator=g)
        self.b64 = torch.empty(b_elems // 2, dtype=torch.float64, device=dev).uniform_(-1, 1, generator=g)
        self.C = torch.empty(c_elems, dtype=torch.float32, device=dev)
        self.C64 = self.C.view(torch.float64)
        self.S = torch.ones(c_elems, dtype=torch.float32, device=dev)     # FP8 scales
        self.sink = torch.zeros(1, dtype=torch.int32, device=dev)
        self.offset = 0

    @classmethod
    def for_shapes(cls, shapes, seed=0):
        """A pool large enough for every shape's tensor (at least 2 GiB) and output."""
        a_bytes, c_words = 2 << 30, 1 << 28
        for s in shapes:
            n = int(np.prod(s["na"], dtype=np.int64))
            a_bytes = max(a_bytes, n * ELEM[s["dtype"]])
            c_words = max(c_words, n // s["na"][s["q"] - 1] * (2 if s["dtype"] == "f64" else 1))
        return cls(pool_bytes=(a_bytes + 255) // 256 * 256, c_elems=c_words, seed=seed)

    def take(self, nbytes):
        """Byte offset of a region of nbytes not used by the previous launches (256-byte aligned)."""
        nbytes = (nbytes + 255) // 256 * 256
        if self.offset + nbytes > self.pool_bytes:
            self.offset = 0
        off = self.offset
        self.offset += nbytes
        return off


def _launcher(pool, shape):
    """Returns run(offset) that enqueues one product of `shape` on the pool region at offset."""
    import ttv_amd
    na, q, pia, wa = shape["na"], shape["q"], shape["pia"], shape["wa"]
    nc = generate_output_shape(na, q)
    wc = generate_strides(nc, generate_output_layout(pia, q))
    ncount = int(np.prod(nc, dtype=np.int64))
    C = pool.C[:ncount].as_strided(nc, wc)
    b = pool.b[:na[q - 1]]
    dt = shape["dtype"]
    if dt == "f32":
        def run(off):
            A = pool.A[off // 4:].as_strided(na, wa)
            ttv_amd.ttv(q, A, b, out=C)
    elif dt == "bf16":
        def run(off):
            A = pool.A16[off // 2:].as_strided(na, wa)
            ttv_amd.ttv_bf16(q, A, b, out=C)
    elif dt == "f64":
        C64 = pool.C64[:ncount].as_strided(nc, wc)
        b64 = pool.b64[:na[q - 1]]

        def run(off):
            A = pool.A64[off // 8:].as_strided(na, wa)
            ttv_amd.ttv(q, A, b64, out=C64)
    else:
        fps = ttv_amd.default_fibers_per_scale(na[q - 1])
        S = pool.S[:int(lib().ttvx_fp8_num_scales(q, len(na), _lib.sizes(na), _lib.sizes(wa), _lib.sizes(pia), fps))]

        def run(off):
            A = pool.A8[off:].as_strided(na, wa)
            ttv_amd.ttv_fp8(q, A, S, fps, b, out=C)
    return run


STRIDED = 11


def valid_variants(shape):
    """Variants that can run the shape; the strided fallback only when nothing else can."""
    import ttv_amd
    code = DT_CODE[shape["dtype"]]
    vs = [v for v in range(NV) if ttv_amd.variant_valid(shape["q"], shape["na"], shape["wa"], shape["pia"], v, code)]
    return [v for v in vs if v != STRIDED] or vs


def time_shape(pool, shape, reps=3, spin_us_per_launch=120.0):
    """Median time (ms) of every valid variant (NaN for invalid ones).  Every launch reads a region
    of the pool that no recent launch touched (at least 2 GiB rotate), so no launch finds its
    operand in the 256 MB last-level cache: HBM-cold times of a workload streaming through many
    tensors."""
    import torch
    import ttv_amd
    elem = ELEM[shape["dtype"]]
    nbytes = int(np.prod(shape["na"], dtype=np.int64)) * elem
    run = _launcher(pool, shape)
    vs = valid_variants(shape)
    out = np.full(NV, np.nan)
    stream = torch.cuda.current_stream()
    # one untimed launch per variant (code objects resident, C allocated)
    for v in vs:
        ttv_amd.set_selection(ttv_amd.SELECT_FORCED, v)
        run(pool.take(nbytes))
    n_launch = len(vs) * reps
    iters = int(min(100000, max(8, n_launch * spin_us_per_launch / 3.4)))
    check(lib().ttvx_bench_spin(iters, _ptr(pool.sink), _stream()))
    evs = []
    for v in vs:
        ttv_amd.set_selection(ttv_amd.SELECT_FORCED, v)
        for _ in range(reps):
            off = pool.take(nbytes)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            run(off)
            e1.record(stream)
            evs.append((v, e0, e1))
    torch.cuda.synchronize()
    ttv_amd.set_selection(ttv_amd.SELECT_AUTO)
    per = {}
    for v, e0, e1 in evs:
        per.setdefault(v, []).append(e0.elapsed_time(e1))
    for v, ts in per.items():
        out[v] = float(np.median(ts))
    return out


def fixed_rule_variant(shape):
    """The variant the fixed rule picks (whatever the process's current selection mode)."""
    import ttv_amd
    ttv_amd.set_selection(ttv_amd.SELECT_HEURISTIC)
    try:
        return ttv_amd.plan_info(shape["q"], shape["na"], shape["wa"], shape["pia"], DT_CODE[shape["dtype"]])["variant"]
    finally:
        ttv_amd.set_selection(ttv_amd.SELECT_AUTO)


def collect_timings(shapes, reps=3, log=print, pool=None):
    import ttv_amd
    pool = pool or TimingPool.for_shapes(shapes)
    feats = np.zeros((len(shapes), NF), np.float32)
    times = np.full((len(shapes), NV), np.nan)
    heur = np.zeros(len(shapes), np.int32)
    t0 = time.time()
    for n, s in enumerate(shapes):
        code = DT_CODE[s["dtype"]]
        feats[n] = ttv_amd.features(s["q"], s["na"], s["wa"], s["pia"], code)
        heur[n] = fixed_rule_variant(s)
        times[n] = time_shape(pool, s, reps=reps)
        if log and (n + 1) % 500 == 0:
            log("timed %d/%d shapes in %.0f s" % (n + 1, len(shapes), time.time() - t0))
    return feats, times, heur


def shape_arrays(shapes, pmax=10):
    """Shapes as fixed-width arrays (for storing next to their timings)."""
    n = len(shapes)
    na = np.zeros((n, pmax), np.int64)
    pia = np.zeros((n, pmax), np.int64)
    meta = np.zeros((n, 3), np.int64)          # p, q, dtype code
    for k, s in enumerate(shapes):
        p = len(s["na"])
        na[k, :p] = s["na"]
        pia[k, :p] = s["pia"]
        meta[k] = (p, s["q"], DT_CODE[s["dtype"]])
    return dict(na=na, pia=pia, meta=meta)


def shapes_from_arrays(z):
    out = []
    meta, na_all, pia_all = z["meta"], z["na"], z["pia"]      # read each (compressed) array once
    for k in range(meta.shape[0]):
        p, q, code = (int(x) for x in meta[k])
        na = [int(x) for x in na_all[k, :p]]
        pia = [int(x) for x in pia_all[k, :p]]
        out.append(dict(na=na, q=q, pia=pia, wa=generate_strides(na, pia), dtype=DT_NAME[code]))
    return out


# ---- training on the GPU --------------------------------------------------------------------

def targets(times):
    """log2 slowdown vs the fastest variant, and the validity mask."""
    best = np.nanmin(times, axis=1, keepdims=True)
    mask = np.isfinite(times).astype(np.float32)
    t = np.where(mask > 0, np.log2(np.where(mask > 0, times, 1.0) / best), 0.0).astype(np.float32)
    return t, mask


def init_params(feats, seed=0, hidden=64):
    rng = np.random.default_rng(seed)
    mean = feats.mean(0)
    std = feats.std(0)
    inv = np.where(std > 1e-6, 1.0 / np.maximum(std, 1e-6), 0.0)
    W1 = rng.standard_normal((hidden, NF)) * np.sqrt(2.0 / NF)
    W2 = rng.standard_normal((hidden, hidden)) * np.sqrt(2.0 / hidden)
    W3 = rng.standard_normal((NV, hidden)) * np.sqrt(1.0 / hidden)
    return np.concatenate([mean, inv, W1.ravel(), np.zeros(hidden), W2.ravel(), np.zeros(hidden), W3.ravel(),
                           np.zeros(NV)]).astype(np.float32)


class GPUTrainer:
    """Minibatch Adam on the GPU through ttvx_mlp_forward / _backward / _adam."""

    def __init__(self, params, batch):
        import torch
        dev = torch.device("cuda:0")
        self.batch = batch
        self.P = torch.from_numpy(np.ascontiguousarray(params, np.float32)).to(dev)
        ntr = self.P.numel() - 2 * NF
        self.g = torch.zeros(ntr, dtype=torch.float32, device=dev)
        self.m1 = torch.zeros_like(self.g)
        self.m2 = torch.zeros_like(self.g)
        self.work = torch.zeros(int(lib().ttvx_mlp_work_floats(batch)), dtype=torch.float32, device=dev)
        self.y = torch.zeros(batch * NV, dtype=torch.float32, device=dev)
        self.loss = torch.zeros(1, dtype=torch.float32, device=dev)
        self.step_no = 0

    def forward(self, X):
        check(lib().ttvx_mlp_forward(_ptr(self.P), _ptr(X), X.shape[0], _ptr(self.y), _ptr(self.work), _stream()))
        return self.y[:X.shape[0] * NV].view(X.shape[0], NV)

    def grad(self, X, T, M):
        self.forward(X)
        check(lib().ttvx_mlp_backward(_ptr(self.P), _ptr(self.y), _ptr(T), _ptr(M), X.shape[0], _ptr(self.work),
                                      _ptr(self.g), _ptr(self.loss), _stream()))
        return self.g, self.loss

    def step(self, X, T, M, lr, wd=1e-5):
        self.grad(X, T, M)
        self.step_no += 1
        check(lib().ttvx_mlp_adam(_ptr(self.P), _ptr(self.g), _ptr(self.m1), _ptr(self.m2), self.step_no, lr, 0.9,
                                  0.999, 1e-8, wd, _stream()))


def train(feats, times, epochs=300, batch=512, lr=2e-3, seed=0, log=None, n_structured=0, struct_weight=1):
    """Adam on the GPU kernels; the last n_structured rows (benchmark shapes) are repeated
    struct_weight times, so the common benchmark shapes weigh more than one random shape each."""
    import torch
    dev = torch.device("cuda:0")
    if n_structured and struct_weight > 1:
        feats = np.concatenate([feats] + [feats[-n_structured:]] * (struct_weight - 1))
        times = np.concatenate([times] + [times[-n_structured:]] * (struct_weight - 1))
    T, M = targets(times)
    params = init_params(feats, seed)
    tr = GPUTrainer(params, batch)
    X = torch.from_numpy(feats).to(dev)
    Td = torch.from_numpy(T).to(dev)
    Md = torch.from_numpy(M).to(dev)
    rng = np.random.default_rng(seed)
    n = feats.shape[0]
    nb = n // batch
    for ep in range(epochs):
        perm = torch.from_numpy(rng.permutation(n)).to(dev)
        lr_ep = lr * 0.5 * (1 + np.cos(np.pi * ep / epochs))          # cosine schedule
        for k in range(nb):
            idx = perm[k * batch:(k + 1) * batch]
            tr.step(X[idx].contiguous(), Td[idx].contiguous(), Md[idx].contiguous(), float(lr_ep))
        if log and (ep + 1) % 50 == 0:
            torch.cuda.synchronize()
            log("epoch %d loss %.4f" % (ep + 1, float(tr.loss.item())))
    torch.cuda.synchronize()
    return tr.P.cpu().numpy()


def predict_gpu(params, feats, times_or_valid):
    """Variant choice per shape with the GPU forward pass (argmin over valid variants)."""
    import torch
    dev = torch.device("cuda:0")
    B = feats.shape[0]
    tr = GPUTrainer(params, B)
    y = tr.forward(torch.from_numpy(np.ascontiguousarray(feats)).to(dev)).cpu().numpy()
    valid = np.isfinite(times_or_valid)
    y = np.where(valid, y, np.inf)
    return np.argmin(y, axis=1), y


def regret(choice, times):
    best = np.nanmin(times, axis=1)
    t = times[np.arange(len(choice)), choice]
    return float(np.mean(t / best - 1.0))


def save(path, params, meta):
    """Weights as .npz (with the report) and as the raw float32 .bin the library loads by itself."""
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez_compressed(path, params=params.astype(np.float32), **{k: np.asarray(v) for k, v in meta.items()})
    params.astype("<f4").tofile(os.path.splitext(path)[0] + ".bin")


def tuning_table(shapes, times):
    """Rows (dtype, O, K, I, variant) of the measured fastest variant per folded problem; shapes
    that fold to the same problem pool their timings (sum of slowdowns vs each one's fastest)."""
    import ttv_amd
    acc = {}
    for s, t in zip(shapes, times):
        if not np.isfinite(t).any():
            continue
        code = DT_CODE[s["dtype"]]
        info = ttv_amd.plan_info(s["q"], s["na"], s["wa"], s["pia"], code)
        if not info["three_d"]:
            continue
        key = (code, info["O"], info["K"], info["I"])
        rel = np.where(np.isfinite(t), t / np.nanmin(t), np.inf)
        acc[key] = acc.get(key, 0.0) + rel
    rows = [list(k) + [int(np.argmin(v))] for k, v in sorted(acc.items())]
    return np.array(rows, dtype=np.int64).reshape(-1, 5)


def load_table(path=TABLE_BIN):
    if not os.path.exists(path):
        return None
    return np.fromfile(path, dtype="<i8").reshape(-1, 5)


def load_weights(path=WEIGHTS_PATH):
    """The shipped weights, or None when absent or shaped for another network (another number of
    kernel variants)."""
    if not os.path.exists(path):
        return None
    with np.load(path) as z:
        p = z["params"].astype(np.float32)
    return p if p.size == int(_lib.lib().ttvx_selector_num_params()) else None


def main():
    """Collects timings of train/held-out shapes, trains on the GPU, reports regrets, saves the
    weights and the timing tables:
        python -m ttv_amd.selector --out DIR [--train N --test M]
        python -m ttv_amd.selector --out DIR --tables DIR0     (retrain on stored tables)"""
    import argparse
    import json
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=20000)
    ap.add_argument("--test", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default="selector_out")
    ap.add_argument("--tables", help="directory with timings_train.npz / timings_test.npz to train on")
    ap.add_argument("--retime-structured", action="store_true",
                    help="with --tables: time the benchmark shapes again first")
    a = ap.parse_args()
    import torch
    from .workloads import order8_shapes, selector_shapes, structured_shapes
    ztr = None

    def log(*x):
        print(*x, flush=True)

    os.makedirs(a.out, exist_ok=True)
    t0 = time.time()
    if a.tables:
        def table(kind):
            for name in ("timings_%s.npz" % kind, "selector_timings_%s_gfx950.npz" % kind):
                if os.path.exists(os.path.join(a.tables, name)):
                    return np.load(os.path.join(a.tables, name))
            raise FileNotFoundError("no %s timing table under %s" % (kind, a.tables))
        ztr, zte = table("train"), table("test")
        f_tr, t_tr, h_tr, n_struct = ztr["feats"], ztr["times"].copy(), ztr["heur"], int(ztr["n_structured"])
        f_te, t_te, h_te = zte["feats"], zte["times"], zte["heur"]
        n_tr = len(f_tr) - n_struct
        if a.retime_structured:
            st_shapes = shapes_from_arrays(ztr)[n_tr:]
            _, t_st, _ = collect_timings(st_shapes, reps=5, log=log)
            t_tr[n_tr:] = t_st
            log("benchmark shapes timed again in %.0f s" % (time.time() - t0))
            np.savez_compressed(os.path.join(a.out, "timings_train.npz"), feats=f_tr, times=t_tr, heur=h_tr,
                                n_structured=n_struct, **{k: ztr[k] for k in ("na", "pia", "meta")})
            np.savez_compressed(os.path.join(a.out, "timings_test.npz"), **{k: zte[k] for k in zte.files})
    else:
        shapes_te = selector_shapes(seed=202, count=a.test)
        shapes_st = structured_shapes() + order8_shapes()
        shapes_tr = selector_shapes(seed=101, count=a.train) + shapes_st
        n_tr, n_struct = a.train, len(shapes_st)
        pool = TimingPool.for_shapes(shapes_tr)
        f_te, t_te, h_te = collect_timings(shapes_te, reps=a.reps, log=log, pool=pool)
        log("held-out timings done in %.0f s" % (time.time() - t0))
        np.savez_compressed(os.path.join(a.out, "timings_test.npz"), feats=f_te, times=t_te, heur=h_te,
                            **shape_arrays(shapes_te))
        f_r, t_r, h_r = collect_timings(shapes_tr[:n_tr], reps=a.reps, log=log, pool=pool)
        f_s, t_s, h_s = collect_timings(shapes_tr[n_tr:], reps=5, log=log, pool=pool)
        f_tr, t_tr, h_tr = np.concatenate([f_r, f_s]), np.concatenate([t_r, t_s]), np.concatenate([h_r, h_s])
        log("training timings done in %.0f s" % (time.time() - t0))
        np.savez_compressed(os.path.join(a.out, "timings_train.npz"), feats=f_tr, times=t_tr, heur=h_tr,
                            n_structured=n_struct, **shape_arrays(shapes_tr))
        del pool
        torch.cuda.empty_cache()
    params = train(f_tr, t_tr, epochs=a.epochs, log=log, n_structured=n_struct, struct_weight=STRUCT_WEIGHT)
    ch_te, _ = predict_gpu(params, f_te, t_te)
    ch_tr, _ = predict_gpu(params, f_tr, t_tr)
    st = slice(n_tr, len(f_tr))
    rep = dict(regret_learned_test=regret(ch_te, t_te), regret_heuristic_test=regret(h_te, t_te),
               regret_learned_train=regret(ch_tr, t_tr), regret_heuristic_train=regret(h_tr, t_tr),
               regret_learned_structured=regret(ch_tr[st], t_tr[st]),
               regret_heuristic_structured=regret(h_tr[st], t_tr[st]),
               n_train=int(len(f_tr)), n_structured=int(n_struct), n_test=int(len(f_te)),
               seconds=time.time() - t0, device=torch.cuda.get_device_properties(0).gcnArchName, hip=torch.version.hip)
    for name, code in DT_CODE.items():
        sel = f_te[:, 6 + [_lib.F32, _lib.BF16, _lib.FP8E4M3, _lib.F64].index(code)] > 0.5
        if sel.any():
            rep["regret_learned_test_" + name] = regret(ch_te[sel], t_te[sel])
            rep["regret_heuristic_test_" + name] = regret(h_te[sel], t_te[sel])
    log(json.dumps(rep))
    save(os.path.join(a.out, "selector_gfx950.npz"), params, dict(report=json.dumps(rep)))
    # tuning table from the training shapes only (the held-out shapes stay unseen)
    tshapes = shapes_from_arrays(ztr) if a.tables else shapes_tr
    tuning_table(tshapes, t_tr).astype("<i8").tofile(os.path.join(a.out, "selector_table_gfx950.bin"))
    with open(os.path.join(a.out, "report.json"), "w") as f:
        json.dump(rep, f, indent=1)


def enable_learned(path=WEIGHTS_PATH):
    """Loads the shipped weights into the library and selects with them; False when absent."""
    import ttv_amd
    p = load_weights(path)
    if p is None:
        return False
    ttv_amd.set_selector_weights(p)
    ttv_amd.set_selection(ttv_amd.SELECT_LEARNED)
    return True


if __name__ == "__main__":
    main()
