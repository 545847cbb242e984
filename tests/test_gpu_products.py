"""GPU parity of the products (ft_mul, ft_fma, ft_inplace_fma, ft_inplace_mul)
against the oracle on identical seeded inputs, through the C-ABI."""
import numpy as np
import pytest
import torch

import oracle as O
import rpp_hip
from helpers import (FT_FMA_CASES, FT_INPLACE_CASES, FT_MUL_CASES, SHAPES, assert_exact, damped_batch,
                     random_batch, window_mask)

pytestmark = pytest.mark.gpu

DTYPES = [np.float64, np.float32]


def to_dev(x, dev):
    return torch.tensor(np.ascontiguousarray(x), device=dev)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", SHAPES)
def test_ft_mul_shapes(dev, shape, dtype):
    w, N = shape
    rng = np.random.default_rng(100 + w * 31 + N)
    basis = rpp_hip.TensorBasis(w, N)
    dim = basis.size()
    n = 6
    lhs = damped_batch(rng, n, w, N, dtype)
    rhs = damped_batch(rng, n, w, N, dtype)
    expected = np.full((n, dim), np.nan, dtype=dtype)
    O.orc.ft_mul(w, N, expected, lhs, rhs, beta=0.75)
    out = to_dev(np.full((n, dim), np.nan, dtype=dtype), dev)
    rpp_hip.ft_mul(out, to_dev(lhs, dev), to_dev(rhs, dev), basis, beta=0.75)
    assert_exact(out.cpu().numpy(), expected, dtype, f"ft_mul {shape}")


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("case", FT_MUL_CASES, ids=[c[0] for c in FT_MUL_CASES])
@pytest.mark.parametrize("width", [2, 3, 7])
def test_ft_mul_windows(dev, case, width, dtype):
    name, ow, lw, rw = case
    N = 4
    rng = np.random.default_rng(7)
    basis = rpp_hip.TensorBasis(width, N)
    dim = basis.size()
    n = 3
    # storage outside the operand windows is populated on purpose (R3)
    lhs = random_batch(rng, n, dim, dtype)
    rhs = random_batch(rng, n, dim, dtype)
    init = np.full((n, dim), np.nan, dtype=dtype)
    expected = init.copy()
    O.orc.ft_mul(width, N, expected, lhs, rhs, ow=ow, lw=lw, rw=rw)
    out = to_dev(init, dev)
    rpp_hip.ft_mul(rpp_hip.Batch(out, *ow), rpp_hip.Batch(to_dev(lhs, dev), *lw),
                   rpp_hip.Batch(to_dev(rhs, dev), *rw), basis)
    got = out.cpu().numpy()
    m = window_mask(width, N, ow)
    assert np.all(np.isnan(got[:, ~m])), "wrote outside out's window"
    assert_exact(got[:, m], expected[:, m], dtype, name)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("case", FT_FMA_CASES, ids=[c[0] for c in FT_FMA_CASES])
def test_ft_fma_windows(dev, case, dtype):
    name, ow, aw, bw, cw = case
    w, N = 3, 4
    rng = np.random.default_rng(11)
    basis = rpp_hip.TensorBasis(w, N)
    dim = basis.size()
    n = 3
    a, b, c = (random_batch(rng, n, dim, dtype) for _ in range(3))
    init = np.full((n, dim), np.nan, dtype=dtype)
    expected = init.copy()
    O.orc.ft_fma(w, N, expected, a, b, c, ow=ow, aw=aw, bw=bw, cw=cw, alpha=0.5, beta=-1.25)
    out = to_dev(init, dev)
    rpp_hip.ft_fma(rpp_hip.Batch(out, *ow), rpp_hip.Batch(to_dev(a, dev), *aw),
                   rpp_hip.Batch(to_dev(b, dev), *bw), rpp_hip.Batch(to_dev(c, dev), *cw), basis,
                   alpha=0.5, beta=-1.25)
    got = out.cpu().numpy()
    m = window_mask(w, N, ow)
    assert np.all(np.isnan(got[:, ~m]))
    assert_exact(got[:, m], expected[:, m], dtype, name)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("variant", [231, 123, 213])
@pytest.mark.parametrize("case", FT_INPLACE_CASES, ids=[c[0] for c in FT_INPLACE_CASES])
def test_ft_inplace_fma_windows(dev, case, variant, dtype):
    name, aw, bw, cw = case
    w, N = 2, 4
    rng = np.random.default_rng(variant)
    basis = rpp_hip.TensorBasis(w, N)
    dim = basis.size()
    n = 4
    a, b, c = (random_batch(rng, n, dim, dtype) for _ in range(3))
    expected = a.copy()
    O.orc.ft_inplace_fma(variant, w, N, expected, b, c, aw=aw, bw=bw, cw=cw, alpha=1.5, beta=0.5)
    ad = to_dev(a, dev)
    rpp_hip.ft_inplace_fma(variant, rpp_hip.Batch(ad, *aw), rpp_hip.Batch(to_dev(b, dev), *bw),
                           rpp_hip.Batch(to_dev(c, dev), *cw), basis, alpha=1.5, beta=0.5)
    assert_exact(ad.cpu().numpy(), expected, dtype, f"{variant} {name}")


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("case", FT_INPLACE_CASES, ids=[c[0] for c in FT_INPLACE_CASES])
def test_ft_inplace_mul_windows(dev, case, dtype):
    name, lw, rw, _ = case
    w, N = 3, 4
    rng = np.random.default_rng(5)
    basis = rpp_hip.TensorBasis(w, N)
    dim = basis.size()
    n = 4
    lhs, rhs = (random_batch(rng, n, dim, dtype) for _ in range(2))
    expected = lhs.copy()
    O.orc.ft_inplace_mul(w, N, expected, rhs, lw=lw, rw=rw, beta=2.0)
    ld = to_dev(lhs, dev)
    rpp_hip.ft_inplace_mul(rpp_hip.Batch(ld, *lw), rpp_hip.Batch(to_dev(rhs, dev), *rw), basis, beta=2.0)
    assert_exact(ld.cpu().numpy(), expected, dtype, name)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("which", ["lhs", "rhs"])
def test_ft_mul_aliasing(dev, which, dtype):
    """out may alias either operand (ft_mul_tests.cpp:373-404)."""
    w, N = 4, 4
    rng = np.random.default_rng(3)
    basis = rpp_hip.TensorBasis(w, N)
    dim = basis.size()
    n = 5
    lhs, rhs = (random_batch(rng, n, dim, dtype) for _ in range(2))
    expected = np.zeros_like(lhs)
    O.orc.ft_mul(w, N, expected, lhs, rhs)
    ld, rd = to_dev(lhs, dev), to_dev(rhs, dev)
    out = ld if which == "lhs" else rd
    rpp_hip.ft_mul(out, ld, rd, basis)
    assert_exact(out.cpu().numpy(), expected, dtype, f"alias {which}")


@pytest.mark.parametrize("dtype", DTYPES)
def test_ft_mul_layouts(dev, dtype):
    """Batch-interleaved [dim, n] operands and a broadcast rhs give the same
    products as the object-major batch."""
    w, N = 5, 3
    rng = np.random.default_rng(9)
    basis = rpp_hip.TensorBasis(w, N)
    dim = basis.size()
    n = 37
    lhs = random_batch(rng, n, dim, dtype)
    rhs1 = random_batch(rng, 1, dim, dtype)[0]
    expected = np.zeros((n, dim), dtype=dtype)
    O.orc.ft_mul(w, N, expected, lhs, rhs1)
    out = torch.zeros((dim, n), dtype=getattr(torch, np.dtype(dtype).name), device=dev)
    rpp_hip.ft_mul(rpp_hip.interleaved(out, basis), rpp_hip.interleaved(to_dev(lhs.T, dev), basis),
                   to_dev(rhs1, dev), basis)
    assert_exact(out.cpu().numpy().T, expected, dtype, "interleaved + broadcast")


def test_c1_shape_ft_mul(dev):
    """1024 pairs, width 4, depth 4, U[-1,1] with coefficient 0 = 1, FP64."""
    rng = np.random.default_rng(0xC1)
    basis = rpp_hip.TensorBasis(4, 4)
    lhs = random_batch(rng, 1024, 341, np.float64)
    rhs = random_batch(rng, 1024, 341, np.float64)
    expected = np.zeros_like(lhs)
    O.orc.ft_mul(4, 4, expected, lhs, rhs)
    out = torch.empty((1024, 341), dtype=torch.float64, device=dev)
    rpp_hip.ft_mul(out, to_dev(lhs, dev), to_dev(rhs, dev), basis)
    assert_exact(out.cpu().numpy(), expected, np.float64, "C1")


def test_invalid_arguments_raise(dev):
    basis = rpp_hip.TensorBasis(3, 3)
    t = torch.zeros((2, basis.size()), dtype=torch.float64, device=dev)
    with pytest.raises(rpp_hip.RppHipError) as e:
        rpp_hip.ft_mul(rpp_hip.Batch(t, 0, 5), t, t, basis)
    assert e.value.code == 3
    with pytest.raises(rpp_hip.RppHipError):
        rpp_hip.ft_fmexp(t, t, t, basis)  # out aliases the multiplier (R7)
