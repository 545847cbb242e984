"""Aliasing on the GPU: views that share a buffer without sharing an element
(column slices of one object-major or interleaved buffer) are accepted and
give the bits of separate buffers; an output that is the same batch as an
input is served where the reference allows it; an output that reaches one
element into an input is refused before any launch."""
import numpy as np
import pytest
import torch

import oracle as O
import rpp_hip
from helpers import assert_exact, damped_batch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_disjoint_column_slices_accepted(dev, dtype):
    w, N, n = 3, 5, 9
    basis = rpp_hip.TensorBasis(w, N)
    dim = basis.size()
    rng = np.random.default_rng(41)
    arg = damped_batch(rng, n, w, N, dtype)
    arg[:, 0] = 1.0
    expected = np.zeros((n, dim), dtype=dtype)
    O.orc.ft_log(w, N, expected, arg)
    # object-major: arg in columns [0, dim), out in [dim, 2 dim) of one buffer
    buf = torch.zeros((n, 2 * dim), dtype=torch.from_numpy(arg).dtype, device=dev)
    buf[:, :dim] = torch.from_numpy(arg).to(dev)
    rpp_hip.ft_log(buf[:, dim:], buf[:, :dim], basis)
    got = buf.cpu().numpy()
    assert_exact(got[:, dim:], expected, dtype, "ft_log, object-major column slices")
    assert np.array_equal(got[:, :dim], arg), "the argument slice changed"
    # interleaved: arg in object columns [0, n), out in [n, 2n) of a [dim, 2n] buffer
    ibuf = torch.zeros((dim, 2 * n), dtype=buf.dtype, device=dev)
    ibuf[:, :n] = torch.from_numpy(np.ascontiguousarray(arg.T)).to(dev)
    rpp_hip.ft_log(rpp_hip.interleaved(ibuf[:, n:], basis), rpp_hip.interleaved(ibuf[:, :n], basis), basis)
    assert_exact(ibuf[:, n:].cpu().numpy().T, expected, dtype, "ft_log, interleaved column slices")


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_same_batch_products_and_linear_ops(dev, dtype):
    w, N, n = 4, 4, 7
    basis = rpp_hip.TensorBasis(w, N)
    dim = basis.size()
    rng = np.random.default_rng(43)
    lhs = damped_batch(rng, n, w, N, dtype)
    rhs = damped_batch(rng, n, w, N, dtype)
    expected = np.zeros((n, dim), dtype=dtype)
    O.orc.ft_mul(w, N, expected, lhs, rhs, beta=0.5)
    # ft_mul with out the very same batch as lhs
    x = torch.from_numpy(lhs).to(dev)
    rpp_hip.ft_mul(x, x, torch.from_numpy(rhs).to(dev), basis, beta=0.5)
    assert_exact(x.cpu().numpy(), expected, dtype, "ft_mul, out is lhs")
    # vector_add with out the same batch as lhs, on a column slice
    buf = torch.from_numpy(np.concatenate([lhs, rhs], axis=1)).to(dev)
    rpp_hip.vector_add(buf[:, :dim], buf[:, :dim], buf[:, dim:], basis, alpha=2.0, beta=-1.0)
    assert np.array_equal(buf[:, :dim].cpu().numpy(), (2.0 * lhs - rhs).astype(dtype))


def test_partial_overlap_refused(dev):
    w, N, n = 3, 4, 5
    basis = rpp_hip.TensorBasis(w, N)
    dim = basis.size()
    buf = torch.zeros((n, 2 * dim), dtype=torch.float64, device=dev)
    buf[:, 0] = 1.0
    with pytest.raises(rpp_hip.RppHipError) as e:
        rpp_hip.ft_log(buf[:, dim - 1:2 * dim - 1], buf[:, :dim], basis)
    assert e.value.code == 3 and "share memory" in str(e.value)
    with pytest.raises(rpp_hip.RppHipError) as e:
        rpp_hip.ft_mul(buf[1:, :dim], buf[:-1, :dim], buf[:-1, dim:], basis)
    assert e.value.code == 3
    torch.cuda.synchronize()
    assert torch.count_nonzero(buf).item() == n, "a refused call wrote"
