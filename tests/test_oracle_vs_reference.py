"""Pin the plain-C oracle against the real reference (oracle/_ref, compiled
from the reference's own headers) on seeded inputs, the reference's shape
sweep and every degree-window case.  Skipped where the reference driver was
never built."""
import numpy as np
import pytest

import oracle as O
from helpers import FT_FMA_CASES, FT_INPLACE_CASES, FT_MUL_CASES, SHAPES, damped_batch, random_batch, rel_err

pytestmark = pytest.mark.skipif(not O.has_ref(), reason="reference driver (oracle/_ref) not built here")

DTYPES = [np.float64, np.float32]
TOL = {np.float64: 1e-12, np.float32: 1e-5}


def check(a, b, dtype):
    # the oracle restates the reference's summation order term for term
    # (its tiled product, ft_multiply.hpp), each product and sum rounded on
    # its own, so the results agree bit for bit
    assert rel_err(a, b) <= TOL[dtype]
    assert np.array_equal(a, b), f"not bit-identical: max rel {rel_err(a, b):.3e}"


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", [s for s in SHAPES if s[0] ** s[1] <= 70000])
def test_ft_mul(shape, dtype):
    w, N = shape
    rng = np.random.default_rng(w + 13 * N)
    dim = O.tensor_dimension(w, N)
    lhs, rhs = damped_batch(rng, 3, w, N, dtype), damped_batch(rng, 3, w, N, dtype)
    a = np.full((3, dim), np.nan, dtype=dtype)
    b = a.copy()
    O.orc.ft_mul(w, N, a, lhs, rhs, beta=0.5)
    O.refimpl.ft_mul(w, N, b, lhs, rhs, beta=0.5)
    check(a, b, dtype)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("case", FT_MUL_CASES, ids=[c[0] for c in FT_MUL_CASES])
def test_ft_mul_windows(case, dtype):
    _, ow, lw, rw = case
    w, N = 3, 4
    rng = np.random.default_rng(1)
    dim = O.tensor_dimension(w, N)
    lhs, rhs = random_batch(rng, 2, dim, dtype), random_batch(rng, 2, dim, dtype)
    a = np.full((2, dim), np.nan, dtype=dtype)
    b = a.copy()
    O.orc.ft_mul(w, N, a, lhs, rhs, ow=ow, lw=lw, rw=rw)
    O.refimpl.ft_mul(w, N, b, lhs, rhs, ow=ow, lw=lw, rw=rw)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    check(np.nan_to_num(a), np.nan_to_num(b), dtype)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("case", FT_FMA_CASES, ids=[c[0] for c in FT_FMA_CASES])
def test_ft_fma_windows(case, dtype):
    _, ow, aw, bw, cw = case
    w, N = 2, 4
    rng = np.random.default_rng(2)
    dim = O.tensor_dimension(w, N)
    A, Bm, Cm = (random_batch(rng, 2, dim, dtype) for _ in range(3))
    a = np.full((2, dim), np.nan, dtype=dtype)
    b = a.copy()
    O.orc.ft_fma(w, N, a, A, Bm, Cm, ow=ow, aw=aw, bw=bw, cw=cw, alpha=0.5, beta=2.0)
    O.refimpl.ft_fma(w, N, b, A, Bm, Cm, ow=ow, aw=aw, bw=bw, cw=cw, alpha=0.5, beta=2.0)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    check(np.nan_to_num(a), np.nan_to_num(b), dtype)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("variant", [231, 123, 213])
@pytest.mark.parametrize("case", FT_INPLACE_CASES, ids=[c[0] for c in FT_INPLACE_CASES])
def test_ft_inplace_fma_windows(case, variant, dtype):
    _, aw, bw, cw = case
    w, N = 3, 4
    rng = np.random.default_rng(variant)
    dim = O.tensor_dimension(w, N)
    A, Bm, Cm = (random_batch(rng, 2, dim, dtype) for _ in range(3))
    a, b = A.copy(), A.copy()
    O.orc.ft_inplace_fma(variant, w, N, a, Bm, Cm, aw=aw, bw=bw, cw=cw, alpha=1.5, beta=0.25)
    O.refimpl.ft_inplace_fma(variant, w, N, b, Bm, Cm, aw=aw, bw=bw, cw=cw, alpha=1.5, beta=0.25)
    check(a, b, dtype)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("case", FT_INPLACE_CASES, ids=[c[0] for c in FT_INPLACE_CASES])
def test_ft_inplace_mul_windows(case, dtype):
    _, lw, rw, _ = case
    w, N = 2, 4
    rng = np.random.default_rng(8)
    dim = O.tensor_dimension(w, N)
    L, R = random_batch(rng, 2, dim, dtype), random_batch(rng, 2, dim, dtype)
    a, b = L.copy(), L.copy()
    O.orc.ft_inplace_mul(w, N, a, R, lw=lw, rw=rw, beta=3.0)
    O.refimpl.ft_inplace_mul(w, N, b, R, lw=lw, rw=rw, beta=3.0)
    check(a, b, dtype)


INTERMEDIATE_WINDOWS = [((0, 4), (0, 4), (0, 4)), ((0, 4), (1, 1), (0, 4)), ((0, 2), (2, 3), (1, 4)),
                        ((1, 3), (0, 4), (1, 2)), ((0, 0), (0, 4), (1, 4)), ((2, 4), (1, 4), (0, 1))]


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("win", INTERMEDIATE_WINDOWS)
@pytest.mark.parametrize("shape", [(1, 4), (2, 4), (4, 4), (3, 4)])
def test_intermediates(shape, win, dtype):
    ow, mw, aw = win
    w, N = shape
    rng = np.random.default_rng(4)
    dim = O.tensor_dimension(w, N)
    mult, arg = damped_batch(rng, 2, w, N, dtype), damped_batch(rng, 2, w, N, dtype)
    for op in ("ft_exp", "ft_log"):
        a = np.full((2, dim), np.nan, dtype=dtype)
        b = a.copy()
        getattr(O.orc, op)(w, N, a, arg, ow=ow, aw=aw)
        getattr(O.refimpl, op)(w, N, b, arg, ow=ow, aw=aw)
        assert np.array_equal(np.isnan(a), np.isnan(b)), op
        check(np.nan_to_num(a), np.nan_to_num(b), dtype)
    a = np.full((2, dim), np.nan, dtype=dtype)
    b = a.copy()
    O.orc.ft_fmexp(w, N, a, mult, arg, ow=ow, mw=mw, ew=aw)
    O.refimpl.ft_fmexp(w, N, b, mult, arg, ow=ow, mw=mw, ew=aw)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    check(np.nan_to_num(a), np.nan_to_num(b), dtype)


# Shapes whose top output degrees the reference forms by its right-tiled
# product (first tiled degree F: 7 at width 3, 5 at width 5, 4 at widths 6
# and 8, 11 at width 2, 3 at width 32, 2 above 32), where the summation order
# differs from ascending rhs degree.
TILED_SHAPES = [(3, 8), (5, 6), (6, 5), (8, 4), (2, 12), (32, 3), (33, 2)]


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", TILED_SHAPES)
def test_right_tiled_degrees_bit_identical(shape, dtype):
    """Products, exp, log and fmexp at shapes that reach the reference's
    right-tiled product order, bit for bit; with lhs[0] = 0 in one object
    (the tiled path skips that scalar term)."""
    w, N = shape
    rng = np.random.default_rng(w * 7 + N)
    dim = O.tensor_dimension(w, N)
    lhs, rhs = damped_batch(rng, 2, w, N, dtype), damped_batch(rng, 2, w, N, dtype)
    lhs[1, 0] = 0
    for ow, lw, rw in (((0, N), (0, N), (0, N)), ((1, N), (0, N - 1), (1, N))):
        a = np.full((2, dim), np.nan, dtype=dtype)
        b = a.copy()
        O.orc.ft_mul(w, N, a, lhs, rhs, ow=ow, lw=lw, rw=rw)
        O.refimpl.ft_mul(w, N, b, lhs, rhs, ow=ow, lw=lw, rw=rw)
        assert np.array_equal(np.isnan(a), np.isnan(b))
        check(np.nan_to_num(a), np.nan_to_num(b), dtype)
    for op in ("ft_exp", "ft_log"):
        a = np.zeros((2, dim), dtype=dtype)
        b = a.copy()
        getattr(O.orc, op)(w, N, a, rhs)
        getattr(O.refimpl, op)(w, N, b, rhs)
        check(a, b, dtype)
    a = np.zeros((2, dim), dtype=dtype)
    b = a.copy()
    O.orc.ft_fmexp(w, N, a, lhs, rhs)
    O.refimpl.ft_fmexp(w, N, b, lhs, rhs)
    check(a, b, dtype)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("windows", [((0, 4), (0, 4)), ((1, 3), (2, 4)), ((3, 4), (0, 2))])
def test_pairing(windows, dtype):
    fw, xw = windows
    rng = np.random.default_rng(6)
    dim = O.tensor_dimension(3, 4)
    f, x = random_batch(rng, 5, dim, dtype), random_batch(rng, 5, dim, dtype)
    a = np.full(5, np.nan, dtype=dtype)
    b = a.copy()
    O.orc.tensor_pairing(3, 4, a, f, x, fw=fw, xw=xw)
    O.refimpl.tensor_pairing(3, 4, b, f, x, fw=fw, xw=xw)
    check(a, b, dtype)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("side", ["lmul", "rmul"])
@pytest.mark.parametrize("windows", [((0, 4), (0, 4), (0, 4)), ((1, 3), (0, 2), (2, 4)), ((0, 2), (1, 1), (1, 4))])
def test_adjoints(side, windows, dtype):
    ow, pw, gw = windows
    rng = np.random.default_rng(10)
    w, N = 3, 4
    dim = O.tensor_dimension(w, N)
    op, arg = random_batch(rng, 3, dim, dtype), random_batch(rng, 3, dim, dtype)
    a = np.full((3, dim), np.nan, dtype=dtype)
    b = a.copy()
    getattr(O.orc, "ft_adj_" + side)(w, N, a, op, arg, ow=ow, pw=pw, gw=gw)
    getattr(O.refimpl, "ft_adj_" + side)(w, N, b, op, arg, ow=ow, pw=pw, gw=gw)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    check(np.nan_to_num(a), np.nan_to_num(b), dtype)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", [(5, 6), (3, 8), (8, 5), (6, 6)])
def test_signature_chain(shape, dtype):
    w, N = shape
    rng = np.random.default_rng(w * N)
    n, L = 2, 4
    inc = rng.normal(0, 0.2, size=(n, L, w)).astype(dtype)
    a = O.signature(w, N, inc.reshape(n * L, w), np.arange(n + 1) * L, dtype=dtype)
    b = O.ref_signature(w, N, inc)
    check(a, b, dtype)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("op", ["tensor_antipode", "tensor_reflect"])
@pytest.mark.parametrize("shape", [(2, 5), (3, 4), (4, 3), (5, 2)])
@pytest.mark.parametrize("windows", [((0, 99), (0, 99)), ((1, 3), (0, 99)), ((0, 99), (2, 3)), ((2, 2), (1, 2))])
def test_antipode_reflect(op, shape, windows, dtype):
    """tensor_antipode / tensor_reflect: bit for bit, every window pairing."""
    w, N = shape
    ow = (windows[0][0], min(windows[0][1], N))
    aw = (windows[1][0], min(windows[1][1], N))
    rng = np.random.default_rng(20 + w)
    dim = O.tensor_dimension(w, N)
    arg = random_batch(rng, 3, dim, dtype)
    a = np.full((3, dim), np.nan, dtype=dtype)
    b = a.copy()
    getattr(O.orc, op)(w, N, a, arg, ow=ow, aw=aw)
    getattr(O.refimpl, op)(w, N, b, arg, ow=ow, aw=aw)
    assert np.array_equal(np.isnan(a), np.isnan(b))
    assert np.array_equal(np.nan_to_num(a), np.nan_to_num(b))
