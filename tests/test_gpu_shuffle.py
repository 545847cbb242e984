"""GPU parity of the shuffle family against the oracle, bit for bit, on the
reference's range tables (NaN outside the input windows, guard values
outside the output window), both dtypes; the adjoint pairings and the
shuffle split evaluated entirely on the GPU at a larger shape; layouts."""
import itertools

import numpy as np
import pytest
import torch

import oracle as O
import rpp_hip
from helpers import FT_FMA_CASES, FT_INPLACE_CASES, FT_MUL_CASES, SHUFFLE_ADJ_CASES

pytestmark = pytest.mark.gpu

DTYPES = [np.float64, np.float32]


def dyadic(rng, n, dim, dtype):
    return (rng.integers(-8, 9, size=(n, dim)) / 8).astype(dtype)


def nan_outside(x, w, N, win):
    db = O.degree_begin(w, N)
    y = x.copy()
    y[:, :db[win[0]]] = np.nan
    y[:, db[win[1] + 1]:] = np.nan
    return y


def windows(depth):
    return [(a, b) for a in range(depth + 1) for b in range(a, depth + 1)]


def dev_t(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def same(a, b):
    return np.array_equal(a, b, equal_nan=True)


def product_cases(N):
    if N <= 2:
        return list(itertools.product(windows(N), windows(N), windows(N)))
    if N == 4:
        return [c[1:] for c in FT_MUL_CASES]
    return [((0, N), (0, N), (0, N))]


def adjoint_cases(N):
    if N <= 2:
        return list(itertools.product(windows(N), windows(N), windows(N)))
    if N == 4:
        return [c[1:] for c in SHUFFLE_ADJ_CASES]
    return [((0, N), (0, N), (0, N))]


SHAPES = [(3, 4), (2, 2), (1, 4), (4, 4), (4, 0), (2, 6)]


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", SHAPES)
def test_products_against_oracle(dev, shape, dtype):
    w, N = shape
    B = rpp_hip.TensorBasis(w, N)
    dim = B.size()
    rng = np.random.default_rng(w * 10 + N)
    n = 3
    for ow, lw, rw in product_cases(N):
        lhs = nan_outside(dyadic(rng, n, dim, dtype), w, N, lw)
        rhs = nan_outside(dyadic(rng, n, dim, dtype), w, N, rw)
        for name in ("st_mul", "left_hs_mul"):
            exp = np.full((n, dim), 7.0, dtype=dtype)
            getattr(O.orc, name)(w, N, exp, lhs, rhs, ow, lw, rw, beta=-0.5)
            out = dev_t(np.full((n, dim), 7.0, dtype=dtype), dev)
            getattr(rpp_hip, name)(rpp_hip.Batch(out, *ow), rpp_hip.Batch(dev_t(lhs, dev), *lw),
                                   rpp_hip.Batch(dev_t(rhs, dev), *rw), B, beta=-0.5)
            assert same(host(out), exp), (name, shape, ow, lw, rw)


@pytest.mark.parametrize("dtype", DTYPES)
def test_fma_variants_against_oracle(dev, dtype):
    w, N = 3, 4
    B = rpp_hip.TensorBasis(w, N)
    dim = B.size()
    rng = np.random.default_rng(44)
    n = 3
    for _, ow, aw, bw, cw in FT_FMA_CASES:
        A = nan_outside(dyadic(rng, n, dim, dtype), w, N, aw)
        Bm = nan_outside(dyadic(rng, n, dim, dtype), w, N, bw)
        C = nan_outside(dyadic(rng, n, dim, dtype), w, N, cw)
        exp = np.full((n, dim), -7.0, dtype=dtype)
        O.orc.st_fma(w, N, exp, A, Bm, C, ow, aw, bw, cw, alpha=0.75, beta=-1.5)
        out = dev_t(np.full((n, dim), -7.0, dtype=dtype), dev)
        rpp_hip.st_fma(rpp_hip.Batch(out, *ow), rpp_hip.Batch(dev_t(A, dev), *aw), rpp_hip.Batch(dev_t(Bm, dev), *bw),
                       rpp_hip.Batch(dev_t(C, dev), *cw), B, alpha=0.75, beta=-1.5)
        assert same(host(out), exp), ("st_fma", ow, aw, bw, cw)
    for _, aw, bw, cw in FT_INPLACE_CASES:
        A = dyadic(rng, n, dim, dtype)
        Bm = nan_outside(dyadic(rng, n, dim, dtype), w, N, bw)
        C = nan_outside(dyadic(rng, n, dim, dtype), w, N, cw)
        exp = A.copy()
        O.orc.st_inplace_fma(w, N, exp, Bm, C, aw, bw, cw, alpha=0.5, beta=2.0)
        a = dev_t(A, dev)
        rpp_hip.st_inplace_fma(rpp_hip.Batch(a, *aw), rpp_hip.Batch(dev_t(Bm, dev), *bw),
                               rpp_hip.Batch(dev_t(C, dev), *cw), B, alpha=0.5, beta=2.0)
        assert same(host(a), exp), ("st_inplace_fma", aw, bw, cw)
    # st_fma with out being the addend itself
    A = dyadic(rng, n, dim, dtype)
    Bm, C = dyadic(rng, n, dim, dtype), dyadic(rng, n, dim, dtype)
    exp = np.zeros((n, dim), dtype=dtype)
    O.orc.st_fma(w, N, exp, A, Bm, C, alpha=2.0, beta=0.25)
    a = dev_t(A, dev)
    rpp_hip.st_fma(a, a, dev_t(Bm, dev), dev_t(C, dev), B, alpha=2.0, beta=0.25)
    assert same(host(a), exp)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("op", ["st_adj_mul", "left_hs_adj_lmul", "left_hs_adj_rmul"])
def test_adjoints_against_oracle(dev, op, shape, dtype):
    w, N = shape
    B = rpp_hip.TensorBasis(w, N)
    dim = B.size()
    rng = np.random.default_rng(7 * w + N)
    n = 3
    kw = {} if op == "st_adj_mul" else {"beta": -2.0}
    for ow, pw, gw in adjoint_cases(N):
        p = nan_outside(dyadic(rng, n, dim, dtype), w, N, pw)
        q = nan_outside(dyadic(rng, n, dim, dtype), w, N, gw)
        exp = np.full((n, dim), 7.0, dtype=dtype)
        getattr(O.orc, op)(w, N, exp, p, q, ow, pw, gw, **kw)
        out = dev_t(np.full((n, dim), 7.0, dtype=dtype), dev)
        getattr(rpp_hip, op)(rpp_hip.Batch(out, *ow), rpp_hip.Batch(dev_t(p, dev), *pw),
                             rpp_hip.Batch(dev_t(q, dev), *gw), B, **kw)
        assert same(host(out), exp), (op, shape, ow, pw, gw)


def test_identities_on_the_gpu_at_width_4_depth_5(dev):
    """Size-independent checks, all arithmetic on the GPU, 64 objects of
    width 4, depth 5: x.y = x<y + y<x + x_0 y_0, and the three pairings
    <st_adj_mul(a, y), x> = <y, a.x>, <left_hs_adj_lmul(x, y), z> = <y, x<z>,
    <left_hs_adj_rmul(x, y), z> = <x, z<y> (dyadic inputs: exact)."""
    B = rpp_hip.TensorBasis(4, 5)
    dim, n = B.size(), 64
    rng = np.random.default_rng(9)
    x, y, z, a = (dev_t(dyadic(rng, n, dim, np.float64), dev) for _ in range(4))
    e = lambda: torch.zeros((n, dim), dtype=torch.float64, device=dev)
    s, h1, h2 = e(), e(), e()
    rpp_hip.st_mul(s, x, y, B)
    rpp_hip.left_hs_mul(h1, x, y, B)
    rpp_hip.left_hs_mul(h2, y, x, B)
    split = h1 + h2
    split[:, 0] += x[:, 0] * y[:, 0]
    assert torch.equal(s, split)

    def pair(u, v):
        out = torch.zeros(n, dtype=torch.float64, device=dev)
        rpp_hip.tensor_pairing(out, u, v, B)
        return out

    adj, prod = e(), e()
    rpp_hip.st_adj_mul(adj, a, y, B)
    rpp_hip.st_mul(prod, a, x, B)
    assert torch.equal(pair(adj, x), pair(y, prod))
    rpp_hip.left_hs_adj_lmul(adj, x, y, B)
    rpp_hip.left_hs_mul(prod, x, z, B)
    assert torch.equal(pair(adj, z), pair(y, prod))
    rpp_hip.left_hs_adj_rmul(adj, x, y, B)
    rpp_hip.left_hs_mul(prod, z, y, B)
    assert torch.equal(pair(adj, z), pair(x, prod))


def test_layouts_and_refusals(dev):
    """An interleaved output and a broadcast factor give the object-major
    result; aliasing and depths beyond 16 are refused."""
    w, N = 3, 3
    B = rpp_hip.TensorBasis(w, N)
    dim, n = B.size(), 33
    rng = np.random.default_rng(12)
    lhs = dyadic(rng, n, dim, np.float64)
    one = dyadic(rng, 1, dim, np.float64)
    exp = np.zeros((n, dim))
    O.orc.st_mul(w, N, exp, lhs, np.repeat(one, n, axis=0))
    il = torch.zeros((dim, n), dtype=torch.float64, device=dev)
    rpp_hip.st_mul(rpp_hip.interleaved(il, B), dev_t(lhs, dev), dev_t(one[0], dev), B, num_batches=n)
    assert np.array_equal(host(il).T, exp)
    x = dev_t(lhs, dev)
    with pytest.raises(rpp_hip.RppHipError) as e:
        rpp_hip.st_mul(x, x, x, B)
    assert e.value.code == 3
    with pytest.raises(rpp_hip.RppHipError) as e:
        rpp_hip.st_inplace_fma(x, x, dev_t(lhs, dev), B)
    assert e.value.code == 3
    deep = rpp_hip.TensorBasis(1, 17)
    t = torch.zeros((2, deep.size()), dtype=torch.float64, device=dev)
    with pytest.raises(rpp_hip.RppHipError) as e:
        rpp_hip.st_mul(t, t.clone(), t.clone(), deep)
    assert e.value.code == 8
