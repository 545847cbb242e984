"""GPU parity of the linear operations against the oracle, bit for bit:
every window pair on tensor and Lie gradings (NaN outside the input windows,
guard values around the output window), object-major and batch-interleaved
layouts and the tiled transposition between them, broadcast operands, the
aliasing the reference allows, and the rounding of vector_add's two products."""
import itertools

import numpy as np
import pytest
import torch

import oracle as O
import rpp_hip
from test_vector_ops_cpu import ROUNDING_CASE, check_rounding_case, storage, windows, with_nan_outside

pytestmark = pytest.mark.gpu

DTYPES = [np.float64, np.float32]
TORCH = {np.float64: torch.float64, np.float32: torch.float32}


def basis_of(kind, w, N):
    return rpp_hip.TensorBasis(w, N) if kind == "tensor" else rpp_hip.LieBasis(w, N)


def offsets_of(kind, w, N):
    return O.degree_begin(w, N) if kind == "tensor" else O.lie_degree_begin(w, N)


def same(a, b):
    return np.array_equal(a, b, equal_nan=True)


def dev_t(x, dev):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


GRADINGS = [("tensor", 3, 3), ("tensor", 1, 4), ("tensor", 4, 0), ("lie", 3, 4), ("lie", 2, 5)]


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("grading", GRADINGS, ids=lambda g: f"{g[0]}{g[1]}x{g[2]}")
def test_assign_inplace_add_scalar_ops_every_window(dev, grading, dtype):
    kind, w, N = grading
    B, start = basis_of(*grading), offsets_of(*grading)
    n = 3
    for ow, aw in itertools.product(windows(N), windows(N)):
        arg = with_nan_outside(storage(start, n, dtype, 1), start, aw)
        exp = np.full((n, start[-1]), 7.0, dtype=dtype)
        O.orc.vector_assign(start, N, exp, arg, ow=ow, aw=aw)
        out = dev_t(np.full((n, start[-1]), 7.0, dtype=dtype), dev)
        rpp_hip.vector_assign(rpp_hip.Batch(out, *ow), rpp_hip.Batch(dev_t(arg, dev), *aw), B)
        assert same(host(out), exp), ("assign", ow, aw)
        lhs0 = storage(start, n, dtype, 2)
        exp = lhs0.copy()
        O.orc.vector_inplace_add(start, N, exp, arg, lw=ow, rw=aw, alpha=-0.5)
        lhs = dev_t(lhs0, dev)
        rpp_hip.vector_inplace_add(rpp_hip.Batch(lhs, *ow), rpp_hip.Batch(dev_t(arg, dev), *aw), B, alpha=-0.5)
        assert same(host(lhs), exp), ("inplace_add", ow, aw)
    for vw in windows(N):
        for name, s in (("vector_scalar_multiply", -0.5), ("vector_set_constant", 0.25),
                        ("tensor_set_identity", 2.0), ("tensor_add_identity", 1.0)):
            v0 = storage(start, n, dtype, 5)
            exp = v0.copy()
            getattr(O.orc, name)(start, N, exp, s, vw=vw)
            v = dev_t(v0, dev)
            kw = {"scalar": s} if name != "vector_set_constant" else {"value": s}
            getattr(rpp_hip, name)(rpp_hip.Batch(v, *vw), B, **kw)
            assert same(host(v), exp), (name, vw)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("grading", [("tensor", 2, 3), ("lie", 3, 3)], ids=lambda g: f"{g[0]}{g[1]}x{g[2]}")
def test_vector_add_every_window_triple(dev, grading, dtype):
    kind, w, N = grading
    B, start = basis_of(*grading), offsets_of(*grading)
    n = 2
    for ow, lw, rw in itertools.product(windows(N), windows(N), windows(N)):
        lhs = with_nan_outside(storage(start, n, dtype, 3), start, lw)
        rhs = with_nan_outside(storage(start, n, dtype, 4), start, rw)
        exp = np.full((n, start[-1]), -7.0, dtype=dtype)
        O.orc.vector_add(start, N, exp, lhs, rhs, ow=ow, lw=lw, rw=rw, alpha=0.5, beta=-1.0)
        out = dev_t(np.full((n, start[-1]), -7.0, dtype=dtype), dev)
        rpp_hip.vector_add(rpp_hip.Batch(out, *ow), rpp_hip.Batch(dev_t(lhs, dev), *lw),
                           rpp_hip.Batch(dev_t(rhs, dev), *rw), B, alpha=0.5, beta=-1.0)
        assert same(host(out), exp), (ow, lw, rw)


def test_vector_add_rounds_each_product_as_the_reference(dev):
    c = ROUNDING_CASE
    out = torch.zeros((1, 3), dtype=torch.float64, device=dev)
    rpp_hip.vector_add(out, dev_t(c["lhs"], dev), dev_t(c["rhs"], dev), rpp_hip.TensorBasis(2, 1),
                       alpha=c["alpha"], beta=c["beta"])
    check_rounding_case(host(out))


# shapes: tensor (3,4) dim 121, (5,4) dim 781, Lie (3,8) true size 1319
LAYOUT_SHAPES = [("tensor", 3, 4), ("tensor", 5, 4), ("lie", 3, 8)]


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("n", [1, 5, 64, 129, 1000])
@pytest.mark.parametrize("grading", LAYOUT_SHAPES, ids=lambda g: f"{g[0]}{g[1]}x{g[2]}")
@pytest.mark.parametrize("wins", [((0, None), (0, None)), ((1, 3), (2, None)), ((2, 2), (0, 1))],
                         ids=["full", "overlap", "disjoint"])
def test_transposition_between_layouts(dev, grading, n, wins, dtype):
    """vector_assign object-major <-> interleaved (the tiled kernel for n > 1)
    and interleaved <-> interleaved, against the oracle's assign on the
    object-major form; the interleaved buffers carry padding columns that must
    stay untouched."""
    kind, w, N = grading
    B, start = basis_of(*grading), offsets_of(*grading)
    ow, aw = ((a, N if b is None else b) for a, b in wins)
    dim = start[-1]
    rng = np.random.default_rng(n + dim)
    arg = with_nan_outside(rng.standard_normal((n, dim)).astype(dtype), start, aw)
    exp = np.full((n, dim), 3.0, dtype=dtype)
    O.orc.vector_assign(start, N, exp, arg, ow=ow, aw=aw)
    pad = 7
    # object-major -> interleaved (padded [dim, n + pad] buffer)
    il = torch.full((dim, n + pad), 3.0, dtype=TORCH[dtype], device=dev)
    rpp_hip.vector_assign(rpp_hip.Batch(il[:, :n], *ow, interleaved=True), rpp_hip.Batch(dev_t(arg, dev), *aw), B,
                          num_batches=n)
    got = host(il)
    assert same(got[:, :n].T, exp)
    assert np.all(got[:, n:] == 3.0)
    # interleaved -> object-major
    arg_il = torch.full((dim, n + pad), np.nan, dtype=TORCH[dtype], device=dev)
    arg_il[:, :n] = dev_t(arg.T, dev)
    om = torch.full((n, dim), 3.0, dtype=TORCH[dtype], device=dev)
    rpp_hip.vector_assign(rpp_hip.Batch(om, *ow), rpp_hip.Batch(arg_il[:, :n], *aw, interleaved=True), B,
                          num_batches=n)
    assert same(host(om), exp)
    # interleaved -> interleaved (the element-wise kernel, object lanes)
    il2 = torch.full((dim, n), 3.0, dtype=TORCH[dtype], device=dev)
    rpp_hip.vector_assign(rpp_hip.Batch(il2, *ow, interleaved=True),
                          rpp_hip.Batch(arg_il[:, :n], *aw, interleaved=True), B, num_batches=n)
    assert same(host(il2).T, exp)


@pytest.mark.parametrize("dtype", DTYPES)
def test_large_round_trip_through_the_interleaved_layout(dev, dtype):
    """3000 signatures-sized objects of (3,8) (9841 coefficients): object-major
    -> interleaved -> object-major is the identity, and the interleaved form is
    the exact transpose."""
    B = rpp_hip.TensorBasis(3, 8)
    g = torch.Generator(device=dev)
    g.manual_seed(7)
    x = torch.randn((3000, B.size()), dtype=TORCH[dtype], device=dev, generator=g)
    il = rpp_hip.to_interleaved(x, B)
    assert torch.equal(il, x.t())
    back = rpp_hip.to_object_major(il, B)
    assert torch.equal(back, x)


@pytest.mark.parametrize("dtype", DTYPES)
def test_broadcast_operands_and_allowed_aliasing(dev, dtype):
    B = rpp_hip.TensorBasis(4, 3)
    start = O.degree_begin(4, 3)
    rng = np.random.default_rng(3)
    n, dim = 37, start[-1]
    one = rng.standard_normal(dim).astype(dtype)
    x0 = rng.standard_normal((n, dim)).astype(dtype)
    # assign from a broadcast object (batch stride 0)
    out = torch.zeros((n, dim), dtype=TORCH[dtype], device=dev)
    rpp_hip.vector_assign(out, dev_t(one, dev), B, num_batches=n)
    assert same(host(out), np.broadcast_to(one, (n, dim)))
    # out = lhs: x <- 2x + 0.5 one
    exp = x0.copy()
    O.orc.vector_add(start, 3, exp, x0, one, alpha=2.0, beta=0.5)
    x = dev_t(x0, dev)
    rpp_hip.vector_add(x, x, dev_t(one, dev), B, alpha=2.0, beta=0.5)
    assert same(host(x), exp)
    # lhs = rhs in place: x += -0.25 x
    exp = x0.copy()
    O.orc.vector_inplace_add(start, 3, exp, x0.copy(), alpha=-0.25)
    x = dev_t(x0, dev)
    rpp_hip.vector_inplace_add(x, x, B, alpha=-0.25)
    assert same(host(x), exp)
    # transposing in place is refused
    sq = torch.zeros((dim, dim), dtype=TORCH[dtype], device=dev)
    with pytest.raises(rpp_hip.RppHipError):
        rpp_hip.vector_assign(rpp_hip.Batch(sq, 0, 3, interleaved=True), rpp_hip.Batch(sq, 0, 3), B,
                              num_batches=dim)
