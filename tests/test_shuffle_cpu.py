"""The shuffle family on the CPU: st_mul, st_fma, st_inplace_fma, st_adj_mul,
left_hs_mul, left_hs_adj_lmul, left_hs_adj_rmul.

* The oracle against the real reference (oracle/_ref) bit for bit on the
  reference's own range tables (tests/rpp/ft_degree_range_cases.hpp,
  shuffle_adjoint_test_helper.hpp:19-36, half_shuffle_adjoint_test_helper.hpp
  :86-108) and on every window triple at depth <= 2, with NaN outside the
  input windows and guard values outside the output window.
* The oracle against the mathematics, on dyadic inputs so every comparison is
  exact: the word products of the reference's half-shuffle tests
  (left_hs_mul_tests.cpp:74-98), commutativity and the unit of the shuffle,
  the split x.y = x<y + y<x + x_0 y_0, the Zinbiel identity
  (x<y)<z = x<(y.z), and the three adjoint pairings
  <st_adj_mul(a, y), x> = <y, a.x>, <left_hs_adj_lmul(x, y), z> = <y, x<z>,
  <left_hs_adj_rmul(x, y), z> = <x, z<y>.
"""
import itertools

import numpy as np
import pytest

import oracle as O
from helpers import FT_FMA_CASES, FT_INPLACE_CASES, FT_MUL_CASES, SHUFFLE_ADJ_CASES

needs_ref = pytest.mark.skipif(not O.has_ref(), reason="reference driver (oracle/_ref) not built here")
DTYPES = [np.float64, np.float32]


def dyadic(rng, n, dim, dtype, denom=8):
    return (rng.integers(-8, 9, size=(n, dim)) / denom).astype(dtype)


def nan_outside(x, width, depth, win):
    db = O.degree_begin(width, depth)
    y = x.copy()
    y[:, :db[win[0]]] = np.nan
    y[:, db[win[1] + 1]:] = np.nan
    return y


def same(a, b):
    return np.array_equal(a, b, equal_nan=True)


def windows(depth):
    return [(a, b) for a in range(depth + 1) for b in range(a, depth + 1)]


# ------------------------------------------------------ oracle vs reference

def product_cases(depth):
    """(out, lhs, rhs) windows: the reference's table at depth 4, every triple
    at depth <= 2."""
    if depth <= 2:
        return list(itertools.product(windows(depth), windows(depth), windows(depth)))
    return [c[1:] for c in FT_MUL_CASES]


@needs_ref
@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", [(3, 4), (2, 4), (1, 4), (4, 4), (2, 2), (3, 1)])
def test_st_mul_and_left_hs_mul_windows(shape, dtype):
    w, N = shape
    rng = np.random.default_rng(w * 7 + N)
    dim = O.tensor_dimension(w, N)
    cases = product_cases(N) if N in (2, 4) else [((0, N), (0, N), (0, N))]
    for ow, lw, rw in cases:
        lhs = nan_outside(dyadic(rng, 2, dim, dtype), w, N, lw)
        rhs = nan_outside(dyadic(rng, 2, dim, dtype), w, N, rw)
        for op in ("st_mul", "left_hs_mul"):
            a = np.full((2, dim), 7.0, dtype=dtype)
            b = a.copy()
            getattr(O.orc, op)(w, N, a, lhs, rhs, ow, lw, rw, beta=-0.5)
            getattr(O.refimpl, op)(w, N, b, lhs, rhs, ow, lw, rw, beta=-0.5)
            assert same(a, b), (op, ow, lw, rw)


@needs_ref
@pytest.mark.parametrize("dtype", DTYPES)
def test_st_fma_and_inplace_windows(dtype):
    w, N = 3, 4
    rng = np.random.default_rng(5)
    dim = O.tensor_dimension(w, N)
    for _, ow, aw, bw, cw in FT_FMA_CASES:
        A = nan_outside(dyadic(rng, 2, dim, dtype), w, N, aw)
        Bm = nan_outside(dyadic(rng, 2, dim, dtype), w, N, bw)
        C = nan_outside(dyadic(rng, 2, dim, dtype), w, N, cw)
        a = np.full((2, dim), -7.0, dtype=dtype)
        b = a.copy()
        O.orc.st_fma(w, N, a, A, Bm, C, ow, aw, bw, cw, alpha=0.75, beta=-1.5)
        O.refimpl.st_fma(w, N, b, A, Bm, C, ow, aw, bw, cw, alpha=0.75, beta=-1.5)
        assert same(a, b), (ow, aw, bw, cw)
    for _, aw, bw, cw in FT_INPLACE_CASES:
        A = dyadic(rng, 2, dim, dtype)
        Bm = nan_outside(dyadic(rng, 2, dim, dtype), w, N, bw)
        C = nan_outside(dyadic(rng, 2, dim, dtype), w, N, cw)
        a, b = A.copy(), A.copy()
        O.orc.st_inplace_fma(w, N, a, Bm, C, aw, bw, cw, alpha=0.5, beta=2.0)
        O.refimpl.st_inplace_fma(w, N, b, Bm, C, aw, bw, cw, alpha=0.5, beta=2.0)
        assert same(a, b), (aw, bw, cw)


def adjoint_cases(depth):
    if depth <= 2:
        return list(itertools.product(windows(depth), windows(depth), windows(depth)))
    if depth == 4:
        return [c[1:] for c in SHUFFLE_ADJ_CASES]
    return [((0, depth), (0, depth), (0, depth))]


@needs_ref
@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", [(1, 4), (2, 2), (4, 4), (4, 0), (3, 3)])
@pytest.mark.parametrize("op", ["st_adj_mul", "left_hs_adj_lmul", "left_hs_adj_rmul"])
def test_adjoints_windows(op, shape, dtype):
    w, N = shape
    rng = np.random.default_rng(11 * w + N)
    dim = O.tensor_dimension(w, N)
    for ow, pw, gw in adjoint_cases(N):
        p = nan_outside(dyadic(rng, 2, dim, dtype), w, N, pw)
        q = nan_outside(dyadic(rng, 2, dim, dtype), w, N, gw)
        a = np.full((2, dim), 7.0, dtype=dtype)
        b = a.copy()
        getattr(O.orc, op)(w, N, a, p, q, ow, pw, gw)
        getattr(O.refimpl, op)(w, N, b, p, q, ow, pw, gw)
        assert same(a, b), (op, ow, pw, gw)


# -------------------------------------------------------- oracle vs maths

W, N = 3, 4
DIM = O.tensor_dimension(W, N)
DB = O.degree_begin(W, N)


def word(*letters):
    t = np.zeros((1, DIM))
    idx = 0
    for l in letters:
        idx = idx * W + l
    t[0, DB[len(letters)] + idx] = 1.0
    return t


def shuffle(x, y):
    return O.orc.st_mul(W, N, np.zeros((1, DIM)), x, y)


def half(x, y):
    return O.orc.left_hs_mul(W, N, np.zeros((1, DIM)), x, y)


def test_word_products_of_the_reference_tests():
    assert np.array_equal(half(word(0), word(1)), word(0, 1))
    assert np.array_equal(half(word(1), word(0)), word(1, 0))
    assert np.array_equal(half(word(0, 1), word(2)), word(0, 1, 2) + word(0, 2, 1))
    assert np.array_equal(half(word(0, 0), word(0, 0)), 3 * word(0, 0, 0, 0))
    f = dyadic(np.random.default_rng(1), 1, DIM, np.float64)
    pos = f.copy()
    pos[0, 0] = 0.0
    assert not half(word(), f).any()
    assert np.array_equal(half(f, word()), pos)
    assert np.array_equal(shuffle(word(0), word(1)), word(0, 1) + word(1, 0))
    assert np.array_equal(shuffle(word(0, 0), word(0)), 3 * word(0, 0, 0))
    assert np.array_equal(shuffle(word(0, 1), word(2, 0)),
                          word(0, 1, 2, 0) + word(0, 2, 1, 0) + word(0, 2, 0, 1) + word(2, 0, 1, 0)
                          + word(2, 0, 0, 1) + word(2, 0, 0, 1))


def test_shuffle_identities():
    rng = np.random.default_rng(2)
    f, g, h = (dyadic(rng, 1, DIM, np.float64) for _ in range(3))
    assert np.array_equal(shuffle(f, g), shuffle(g, f))
    assert np.array_equal(shuffle(word(), f), f)
    split = half(f, g) + half(g, f)
    split[0, 0] += f[0, 0] * g[0, 0]
    assert np.array_equal(shuffle(f, g), split)
    assert np.array_equal(half(half(f, g), h), half(f, shuffle(g, h)))
    # st_fma and st_inplace_fma against st_mul
    out = np.zeros((1, DIM))
    O.orc.st_fma(W, N, out, h, f, g, alpha=0.5, beta=2.0)
    assert np.array_equal(out, 0.5 * h + 2.0 * shuffle(f, g))
    a = h.copy()
    O.orc.st_inplace_fma(W, N, a, f, g, alpha=-1.0, beta=0.25)
    assert np.array_equal(a, -h + 0.25 * shuffle(f, g))


def test_adjoint_pairings():
    rng = np.random.default_rng(3)
    x, y, z, a = (dyadic(rng, 1, DIM, np.float64) for _ in range(4))
    pair = lambda u, v: float(np.sum(u * v))
    adj = O.orc.st_adj_mul(W, N, np.zeros((1, DIM)), a, y)
    assert pair(adj, x) == pair(y, shuffle(a, x))
    adj = O.orc.left_hs_adj_lmul(W, N, np.zeros((1, DIM)), x, y)
    assert pair(adj, z) == pair(y, half(x, z))
    adj = O.orc.left_hs_adj_rmul(W, N, np.zeros((1, DIM)), x, y)
    assert pair(adj, z) == pair(x, half(z, y))


def test_a_planted_mistake_is_caught():
    """The identities above are strong enough to see a swapped operand: the
    half-shuffle is not symmetric."""
    rng = np.random.default_rng(4)
    f, g = (dyadic(rng, 1, DIM, np.float64) for _ in range(2))
    assert not np.array_equal(half(f, g), half(g, f))


# ------------------------------------------------------- C-ABI validation

def test_shuffle_entry_points_validate_on_host():
    import ctypes as C

    import rpp_hip
    from rpp_hip import _capi as A

    lib = rpp_hip.lib()
    buf = (C.c_double * 256)()
    p = C.cast(buf, C.c_void_p).value
    t = A.TensorBatch(p, 40, 1, 0, 3)
    other = A.TensorBatch(p + 8 * 100, 40, 1, 0, 3)
    call = lambda name, d: getattr(lib, name)(C.byref(d), None)
    # the output shares memory with a factor
    assert call("rpp_hip_st_mul_f64", A.FtMulDesc(A.Basis(3, 3), 2, t, t, other, 1.0)) == 3
    assert call("rpp_hip_st_inplace_fma_f64", A.StInplaceFmaDesc(A.Basis(3, 3), 2, t, other, t, 1.0, 1.0)) == 3
    # st_fma may write into its addend only as the very same batch
    moved = A.TensorBatch(p, 41, 1, 0, 3)
    assert call("rpp_hip_st_fma_f64", A.FtFmaDesc(A.Basis(3, 3), 2, t, moved, other, other, 1.0, 1.0)) == 3
    # words beyond depth 16
    assert call("rpp_hip_left_hs_mul_f64", A.HalfShuffleDesc(A.Basis(1, 17), 1, t, other, other, 1.0)) == 8
    assert b"16" in lib.rpp_hip_last_error()
    # windows beyond the depth
    bad = A.TensorBatch(p, 40, 1, 0, 4)
    assert call("rpp_hip_st_adj_mul_f64", A.AdjointDesc(A.Basis(3, 3), 2, bad, other, other)) == 3
    # nothing to do
    empty = A.TensorBatch(p, 40, 1, 2, 1)
    assert call("rpp_hip_left_hs_adj_rmul_f64", A.HalfShuffleDesc(A.Basis(3, 3), 2, empty, other, other, 1.0)) == 0
