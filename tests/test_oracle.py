"""Pins the CPU oracle (oracle/) before it is trusted as the checker of the GPU path.

Each oracle function is checked against what the reference's own tests and files fix, against a
library routine, or against mathematics -- never against a retyped copy of its own formula:

* known-answer closed form C[j] = nq^2 j + nq(nq+1)/2 with the reference's ttv_init
  (test/src/gtest_tlib_ttv.cpp:27-71, :132) over ranks 2-4, shapes {2,4,8}^p, every layout, every q
* numpy.einsum on random data in arbitrary layouts, and a pure-Python brute force on tiny shapes
* the ttvpy test vectors (ttvpy/tests/test.py:5-27) and the example outputs
  (example/interface1.cpp, interface2.cpp, interface3.cpp: C = [15 18 21 24 51 54 57 60])
* the GEMV closed forms of test/src/gtest_tlib_mtv.cpp:70-77 through the 2-D cases
* fixtures written by the compiled reference (tests/golden/ttv_ref_fixtures.npz)
* the compiled reference itself (oracle/_ref, all 17 tested policies) when it is present
* the BF16 and E4M3 codecs against torch's bfloat16 / float8_e4m3fn conversions and the format's
  definition; the selector network's gradient against central finite differences
"""
import itertools
import os

import numpy as np
import pytest

from oracle import oracle as O

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = os.path.join(REPO, "tests", "golden", "ttv_ref_fixtures.npz")


@pytest.fixture(scope="module", autouse=True)
def _lib(oracle_lib):
    return oracle_lib


def _einsum_ref(q, a_flat, na, wa, b):
    """numpy.einsum on the strided view; returns C in the dense output layout of pia."""
    view = np.lib.stride_tricks.as_strided(a_flat, shape=na, strides=[w * a_flat.itemsize for w in wa])
    letters = "abcdefghij"[:len(na)]
    out = letters.replace(letters[q - 1], "")
    return np.einsum(f"{letters},{letters[q - 1]}->{out}", view.astype(np.float64), b.astype(np.float64))


def _to_dense_out(cnd, na, q, pia):
    """Writes an ndarray C (mode order) into the flat dense output layout of pia."""
    nc = O.output_shape(na, q)
    wc = O.strides(nc, O.output_layout(pia, q))
    flat = np.zeros(O.span(nc, wc))
    view = np.lib.stride_tricks.as_strided(flat, shape=nc, strides=[w * 8 for w in wc])
    view[...] = cnd
    return flat


@pytest.mark.parametrize("rank", [2, 3, 4])
def test_naive_known_answer_grid(rank):
    """Every shape {2,4,8}^p, every layout and q, b = 1 (gtest_tlib_ttv.cpp:146-189)."""
    n = 0
    for na in O.generate_shapes([2] * rank, [3] * rank):
        for pia in O.generate_permutations(rank):
            wa = O.strides(na, pia)
            for q in range(1, rank + 1):
                a = O.known_answer_input(q, na, wa, pia)
                b = np.ones(na[q - 1])
                c = O.naive(q, a, na, wa, b, pia=pia)
                nc = O.output_shape(na, q)
                exp = O.known_answer_expected(na[q - 1], int(np.prod(nc)))
                assert np.array_equal(c, exp), (na, pia, q)
                n += 1
    assert n > 0


def test_naive_vs_einsum_and_bruteforce():
    rng = np.random.default_rng(1)
    for p in range(1, 6):
        for _ in range(6):
            na = list(rng.integers(1, 5, p))
            pia = list(rng.permutation(p) + 1)
            wa = O.strides(na, pia)
            a = rng.uniform(-1, 1, O.span(na, wa))
            for q in range(1, p + 1):
                b = rng.uniform(-1, 1, na[q - 1])
                if p == 1:
                    continue
                c = O.naive(q, a, na, wa, b, pia=pia)
                exp = _to_dense_out(_einsum_ref(q, a, na, wa, b), na, q, pia)
                np.testing.assert_allclose(c, exp, rtol=1e-12, atol=1e-12)
                # brute force over every multi-index
                nc = O.output_shape(na, q)
                wc = O.strides(nc, O.output_layout(pia, q))
                bf = np.zeros_like(c)
                for idx in itertools.product(*[range(x) for x in na]):
                    off = sum(i * w for i, w in zip(idx, wa))
                    cidx = [i for r, i in enumerate(idx, 1) if r != q]
                    bf[sum(i * w for i, w in zip(cidx, wc))] += a[off] * b[idx[q - 1]]
                np.testing.assert_allclose(c, bf, rtol=1e-12, atol=1e-12)


def test_naive_general_strides():
    """A padded (non-dense) view and a strided C: the naive oracle takes any strides."""
    rng = np.random.default_rng(2)
    na, q = [3, 4, 5], 2
    wa = [1, 4, 20]                     # dense would be 1,3,12
    a = rng.uniform(-1, 1, O.span(na, wa))
    b = rng.uniform(-1, 1, 4)
    c = O.naive(q, a, na, wa, b, wc=[2, 7])
    exp = _einsum_ref(q, a, na, wa, b)
    got = np.lib.stride_tricks.as_strided(c, shape=[3, 5], strides=[16, 56])
    np.testing.assert_allclose(got, exp, rtol=1e-12)


def test_ttvpy_vectors_and_examples():
    """ttvpy/tests/test.py:5-27 (C-order arrays, exact) and the example programs' output."""
    A = np.arange(3 * 2 * 4, dtype=np.float64).reshape(3, 2, 4)
    na, wa, pia = [3, 2, 4], [8, 4, 1], [3, 2, 1]
    for q, spec in ((1, "ijk,i->jk"), (2, "ijk,j->ik"), (3, "ijk,k->ij")):
        b = np.arange(na[q - 1], dtype=np.float64)
        D = np.einsum(spec, A, b)
        c = O.naive(q, A.ravel(), na, wa, b, pia=pia)
        assert np.array_equal(c, D.ravel())
    # example/interface1.cpp: A = iota(1..24) of shape {4,3,2} first-order, q = 2, b = 1
    na = [4, 3, 2]
    a = np.arange(1, 25, dtype=np.float64)
    c = O.naive(2, a, na, O.strides(na, [1, 2, 3]), np.ones(3), pia=[1, 2, 3])
    assert c.tolist() == [15, 18, 21, 24, 51, 54, 57, 60]


def test_gemv_closed_forms():
    """gtest_tlib_mtv.cpp:70-97: row sums of the 1..M*N fill are fn(i) - fn(i-1)."""
    for m, n in ((2, 2), (3, 8), (16, 4), (7, 5)):
        fn = lambda i: (i * n * (i * n + 1)) // 2  # noqa: E731
        exp = np.array([fn(i) - fn(i - 1) for i in range(1, m + 1)], dtype=np.float64)
        # row-major (layout {2,1}): a[j + i*n] = j + 1 + i*n ; contract mode 2
        a_rm = np.array([j + 1 + i * n for i in range(m) for j in range(n)], dtype=np.float64)
        c = O.naive(2, a_rm, [m, n], [n, 1], np.ones(n), pia=[2, 1])
        assert np.array_equal(c, exp)
        # column-major (layout {1,2}): a[i + j*m] = i*n + j + 1
        a_cm = np.zeros(m * n)
        for j in range(n):
            for i in range(m):
                a_cm[i + j * m] = i * n + j + 1
        c = O.naive(2, a_cm, [m, n], [1, m], np.ones(n), pia=[1, 2])
        assert np.array_equal(c, exp)


@pytest.mark.parametrize("blas", [False, True])
def test_log_cpu_known_answers(blas):
    """The restated default CPU path (with OpenBLAS and with the non-BLAS loops) on the
    known-answer grid, FP64 and FP32 (all values are integers below 2^24, so exact)."""
    if blas and O.load_blas() is None:
        pytest.skip("no OpenBLAS library found")
    if not blas and O.lib().oracle_blas_loaded():
        pytest.skip("BLAS already bound in this process")
    for rank in (2, 3, 4):
        for na in O.generate_shapes([2] * rank, [3] * rank):
            for pia in O.generate_permutations(rank):
                wa = O.strides(na, pia)
                for q in range(1, rank + 1):
                    a = O.known_answer_input(q, na, wa, pia)
                    b = np.ones(na[q - 1])
                    exp = O.known_answer_expected(na[q - 1], int(np.prod(O.output_shape(na, q))))
                    for dt in (np.float64, np.float32):
                        c, case = O.log_cpu(q, a.astype(dt), na, wa, pia, b.astype(dt), threads=2)
                        assert np.array_equal(c.astype(np.float64), exp), (na, pia, q, dt)


def test_log_cpu_vs_naive_random():
    O.load_blas()
    rng = np.random.default_rng(3)
    for p in (2, 3, 4, 5):
        for _ in range(4):
            na = list(rng.integers(2, 9, p))
            pia = list(rng.permutation(p) + 1)
            wa = O.strides(na, pia)
            a = rng.uniform(-1, 1, O.span(na, wa))
            for q in range(1, p + 1):
                b = rng.uniform(-1, 1, na[q - 1])
                ref = O.naive(q, a, na, wa, b, pia=pia)
                c64, _ = O.log_cpu(q, a, na, wa, pia, b, threads=2)
                np.testing.assert_allclose(c64, ref, rtol=1e-12, atol=1e-13)
                c32, _ = O.log_cpu(q, a.astype(np.float32), na, wa, pia, b.astype(np.float32), threads=2)
                assert np.linalg.norm(c32 - ref) <= 1e-5 * np.linalg.norm(ref)


def test_oracle_vs_reference_fixtures():
    """Fixtures written by the compiled reference (tests/golden/make_golden.py)."""
    fx = np.load(FIX)
    n = 0
    for key in fx.files:
        if not key.endswith("_A"):
            continue
        tag, ps, pis = key.split("_")[:3]
        p = int(ps[1:])
        pia = [int(ch) for ch in pis[2:]]
        na = {2: [3, 5], 3: [3, 4, 5], 4: [2, 3, 4, 5]}[p]
        wa = O.strides(na, pia)
        a = fx[key]
        for q in range(1, p + 1):
            b = fx[f"{tag}_p{p}_q{q}_b"]
            cref = fx[f"{tag}_p{p}_pi{''.join(map(str, pia))}_q{q}_C"]
            cseq = fx[f"{tag}_p{p}_pi{''.join(map(str, pia))}_q{q}_Cseq"]
            c = O.naive(q, a, na, wa, b, pia=pia)
            tol = 1e-12 if tag == "f64" else 1e-5
            for r in (cref, cseq):
                assert np.linalg.norm(c - r) <= tol * np.linalg.norm(c), (key, q)
            n += 1
    assert n == 2 * sum(len(list(itertools.permutations(range(p)))) * p for p in (2, 3, 4))


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (needs the reference sources)")
def test_oracle_vs_compiled_reference_all_policies():
    rng = np.random.default_rng(4)
    for p in (2, 3, 4):
        na = list(rng.integers(2, 6, p))
        for pia in O.generate_permutations(p):
            wa = O.strides(na, pia)
            a = rng.uniform(-1, 1, O.span(na, wa))
            for q in range(1, p + 1):
                b = rng.uniform(-1, 1, na[q - 1])
                c = O.naive(q, a, na, wa, b, pia=pia)
                for pol in range(17):
                    r = O.ref_ttv(q, a, na, wa, pia, b, policy=pol)
                    np.testing.assert_allclose(r, c, rtol=1e-12, atol=1e-13)


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (needs the reference sources)")
def test_number_cores_matches_reference():
    assert O.number_cores() == int(O.ref().ref_number_cores())


def test_bf16_codec_vs_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.standard_normal(100000).astype(np.float32),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.01171875, 3.4e38, -3.4e38, 1e-40, np.inf, -np.inf],
                                 dtype=np.float32)])
    ours = O.bf16_rne(x)
    theirs = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(ours, theirs)
    assert np.array_equal(O.bf16_to_f32(ours), torch.from_numpy(x).to(torch.bfloat16).float().numpy())


def test_e4m3_codec():
    torch = pytest.importorskip("torch")
    codes = np.arange(256, dtype=np.uint8)
    dec = O.e4m3_decode(codes)
    # definition: bias 7, 3 mantissa bits, subnormals at exponent field 0, NaN at S.1111.111
    for c in range(256):
        s = -1.0 if c & 0x80 else 1.0
        e, m = (c >> 3) & 15, c & 7
        if (c & 0x7F) == 0x7F:
            assert np.isnan(dec[c])
            continue
        v = s * (np.ldexp(m, -9) if e == 0 else np.ldexp(8 + m, e - 10))
        assert dec[c] == np.float32(v), c
    assert dec[0x7E] == 448.0 and dec[0x08] == 2.0 ** -6 and dec[0x01] == 2.0 ** -9
    # torch's float8_e4m3fn decode of every code
    t = torch.from_numpy(codes).view(torch.float8_e4m3fn).float().numpy()
    fin = ~np.isnan(dec)
    assert np.array_equal(t[fin], dec[fin])
    # encode: exact on representable values, RNE on random in-range values (vs torch)
    assert np.array_equal(O.e4m3_encode(dec[fin]), codes[fin])
    rng = np.random.default_rng(6)
    x = (rng.standard_normal(200000) * rng.choice([1e-3, 0.1, 1, 30, 300], 200000)).astype(np.float32)
    x = np.clip(x, -448, 448)
    ours = O.e4m3_encode(x)
    theirs = torch.from_numpy(x).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    # +-0 may differ only in sign of zero results
    same = (ours == theirs) | ((ours & 0x7F) == 0) & ((theirs & 0x7F) == 0)
    assert same.all(), (x[~same][:5], ours[~same][:5], theirs[~same][:5])
    # saturation above the largest finite value
    assert O.e4m3_encode(np.array([500.0, -1e30, np.inf], np.float32)).tolist() == [0x7E, 0xFE, 0x7E]


def test_quantize_fp8_groups():
    """Scales are per group of fibers in C-offset order; decode(codes)*scale approximates A."""
    rng = np.random.default_rng(7)
    na, pia, q = [6, 5, 4], [2, 3, 1], 3
    wa = O.strides(na, pia)
    a = rng.standard_normal(O.span(na, wa)).astype(np.float32)
    codes, scales = O.quantize_fp8(q, a, na, wa, pia, fps=4)
    nf, offs = O.fiber_offsets(q, na, wa, pia)
    assert scales.shape[0] == -(-nf // 4)
    for g in range(scales.shape[0]):
        vals = a[offs[g * 4:(g + 1) * 4]]
        assert scales[g] == np.float32(np.abs(vals).max() / np.float32(448))
    deq = O.e4m3_decode(codes[offs]) * scales[np.arange(nf) // 4][:, None]
    rel = np.abs(deq - a[offs]) / np.maximum(np.abs(a[offs]), 1e-3)
    assert np.median(rel) < 0.05


def test_mlp_gradient_finite_differences():
    rng = np.random.default_rng(8)
    npar = 2 * O.MLP_F + O.MLP_H * O.MLP_F + O.MLP_H + O.MLP_H * O.MLP_H + O.MLP_H + O.MLP_V * O.MLP_H + O.MLP_V
    params = rng.standard_normal(npar) * 0.3
    params[O.MLP_F:2 * O.MLP_F] = 1.0
    params[:O.MLP_F] = 0.0
    X = rng.standard_normal((16, O.MLP_F))
    T = rng.standard_normal((16, O.MLP_V))
    M = (rng.uniform(size=(16, O.MLP_V)) > 0.3).astype(np.float64)
    loss, g = O.mlp_loss_grad(params, X, T, M)
    off = 2 * O.MLP_F
    for idx in rng.choice(g.size, 40, replace=False):
        e = np.zeros_like(params)
        e[off + idx] = 1e-6
        lp, _ = O.mlp_loss_grad(params + e, X, T, M)
        lm, _ = O.mlp_loss_grad(params - e, X, T, M)
        fd = (lp - lm) / 2e-6
        assert abs(fd - g[idx]) <= 1e-5 * max(1.0, abs(fd)), (idx, fd, g[idx])
