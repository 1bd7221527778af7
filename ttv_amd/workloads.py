"""Synthetic problem generators for the batched and the selector workloads.

batch_problems(seed, count=512, cap=2**26)
    Orders 3 or 4 (50/50), extents uniform integers in [16, 512] redrawn until the problem has at
    most `cap` elements, q uniform over the modes, a uniformly random layout.
selector_shapes(seed, count)
    Orders 2..10, total elements log-uniform in [2^20, 2^28] (split into extents >= 2 by a random
    partition of log2 N), q uniform, random layout, element type FP32 or BF16 (50/50).
Both return plain Python descriptions (na, q, pia, dtype); strides are dense for the layout.
"""
import numpy as np

from .meta import generate_strides


def batch_problems(seed, count=512, cap=1 << 26, lo=16, hi=512):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        p = 3 if rng.uniform() < 0.5 else 4
        while True:
            na = [int(x) for x in rng.integers(lo, hi + 1, p)]
            if int(np.prod(na, dtype=np.int64)) <= cap:
                break
        q = int(rng.integers(1, p + 1))
        pia = [int(x) for x in rng.permutation(p) + 1]
        out.append(dict(na=na, q=q, pia=pia, wa=generate_strides(na, pia)))
    return out


def structured_shapes(dtypes=("f32", "bf16", "f64", "fp8"), sizes_log2=(20, 22, 24, 26, 28), pmax=10):
    """Equal extents n^p (the shapes of the reference's benchmarks, e.g. 128^4 and 16^7) for
    p = 2..pmax and N = 2^L, every q, first- and last-order layouts, every storage type."""
    out = []
    for p in range(2, pmax + 1):
        for L in sizes_log2:
            n = max(2, int(round(2.0 ** (L / p))))
            while n ** p > (1 << max(sizes_log2)):
                n -= 1
            na = [n] * p
            for pia in (list(range(1, p + 1)), list(range(p, 0, -1))):
                for q in range(1, p + 1):
                    for dt in dtypes:
                        out.append(dict(na=na, q=q, pia=pia, wa=generate_strides(na, pia), dtype=dt))
    return out


ORDER8 = [4096, 4, 8, 4, 8, 4, 2, 64]


def order8_shapes(dtypes=("f32", "bf16", "f64", "fp8")):
    """The order-8 configuration (extents 4096,4,8,4,8,4,2,64 = 2^31 elements) for every q in
    first-order, last-order and the layout that the benchmark draws (permutation of seed 31),
    every storage type."""
    seeded = [int(x) for x in np.random.default_rng(31).permutation(8) + 1]
    out = []
    for pia in (list(range(1, 9)), list(range(8, 0, -1)), seeded):
        for q in range(1, 9):
            for dt in dtypes:
                out.append(dict(na=list(ORDER8), q=q, pia=pia, wa=generate_strides(ORDER8, pia), dtype=dt))
    return out


def selector_shapes(seed, count, lo_log2=20, hi_log2=28, pmin=2, pmax=10, dtypes=("f32", "bf16", "f64", "fp8")):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        p = int(rng.integers(pmin, pmax + 1))
        L = rng.uniform(lo_log2, hi_log2)
        w = rng.dirichlet(np.ones(p))
        na = [max(2, int(round(2.0 ** (L * x)))) for x in w]
        n = int(np.prod(na, dtype=np.int64))
        if n > (1 << hi_log2) or n < (1 << (lo_log2 - 1)):
            continue
        q = int(rng.integers(1, p + 1))
        pia = [int(x) for x in rng.permutation(p) + 1]
        out.append(dict(na=na, q=q, pia=pia, wa=generate_strides(na, pia), dtype=dtypes[int(rng.integers(len(dtypes)))]))
    return out
