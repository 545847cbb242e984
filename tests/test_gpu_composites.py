"""GPU parity of ft_exp, ft_fmexp, ft_log, tensor_pairing and the adjoint
multiplications against the oracle."""
import numpy as np
import pytest
import torch

import oracle as O
import rpp_hip
from helpers import assert_close, assert_exact, damped_batch, degree_begin, random_batch, rel_err, window_mask

pytestmark = pytest.mark.gpu

DTYPES = [np.float64, np.float32]
SHAPES = [(1, 5), (2, 6), (3, 4), (4, 4), (5, 3), (6, 3), (7, 2), (16, 2)]


def to_dev(x, dev):
    return torch.tensor(np.ascontiguousarray(x), device=dev)


def deg1(rng, n, w, N, dtype, scale=0.5):
    dim = degree_begin(w, N)[-1]
    x = rng.uniform(-1, 1, size=(n, dim)).astype(dtype)  # junk outside degree 1
    x[:, 1:1 + w] = rng.normal(0.0, scale, size=(n, w))
    return x


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("ew", [(1, 1), (0, 1), None], ids=["deg1", "deg01", "full"])
def test_ft_fmexp(dev, shape, ew, dtype):
    w, N = shape
    rng = np.random.default_rng(w * 10 + N)
    basis = rpp_hip.TensorBasis(w, N)
    dim = basis.size()
    n = 5
    mult = damped_batch(rng, n, w, N, dtype)
    ex = deg1(rng, n, w, N, dtype) if ew is not None else damped_batch(rng, n, w, N, dtype)
    ew = ew or (0, N)
    expected = np.full((n, dim), np.nan, dtype=dtype)
    O.orc.ft_fmexp(w, N, expected, mult, ex, ew=ew)
    out = to_dev(np.full((n, dim), np.nan, dtype=dtype), dev)
    rpp_hip.ft_fmexp(out, to_dev(mult, dev), rpp_hip.Batch(to_dev(ex, dev), *ew), basis)
    assert_exact(out.cpu().numpy(), expected, dtype, f"fmexp {shape} {ew}")


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("ranges", [((0, 4), (1, 4)), ((0, 2), (1, 1)), ((1, 3), (1, 2)), ((2, 4), (0, 4)),
                                    ((0, 0), (1, 4)), ((0, 4), (3, 4))])
@pytest.mark.parametrize("out_max", [0, 2, 4])
def test_ft_fmexp_windows(dev, ranges, out_max, dtype):
    """Windows of multiplier and exponent, outputs of every prefix length
    (ft_intermediate_initialization_tests.cpp:24-109)."""
    mw, ew = ranges
    w, N = 4, 4
    rng = np.random.default_rng(111)
    basis = rpp_hip.TensorBasis(w, N)
    dim = basis.size()
    n = 3
    mult = damped_batch(rng, n, w, N, dtype)
    ex = damped_batch(rng, n, w, N, dtype)
    ow = (0, out_max)
    init = np.full((n, dim), np.nan, dtype=dtype)
    expected = init.copy()
    O.orc.ft_fmexp(w, N, expected, mult, ex, ow=ow, mw=mw, ew=ew)
    out = to_dev(init, dev)
    rpp_hip.ft_fmexp(rpp_hip.Batch(out, *ow), rpp_hip.Batch(to_dev(mult, dev), *mw),
                     rpp_hip.Batch(to_dev(ex, dev), *ew), basis)
    got = out.cpu().numpy()
    m = window_mask(w, N, ow)
    assert np.all(np.isnan(got[:, ~m]))
    assert_exact(got[:, m], expected[:, m], dtype, f"fmexp windows {ranges} {out_max}")


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("op", ["exp", "log"])
@pytest.mark.parametrize("aw", [(1, 1), (0, 99), (2, 3)], ids=["deg1", "full", "deg23"])
def test_ft_exp_log(dev, shape, op, aw, dtype):
    w, N = shape
    aw = (aw[0], min(aw[1], N))
    if aw[0] > N:
        pytest.skip("window beyond depth")
    rng = np.random.default_rng(17 * w + N)
    basis = rpp_hip.TensorBasis(w, N)
    dim = basis.size()
    n = 4
    arg = damped_batch(rng, n, w, N, dtype)
    expected = np.full((n, dim), np.nan, dtype=dtype)
    getattr(O.orc, "ft_" + op)(w, N, expected, arg, aw=aw)
    out = to_dev(np.full((n, dim), np.nan, dtype=dtype), dev)
    getattr(rpp_hip, "ft_" + op)(out, rpp_hip.Batch(to_dev(arg, dev), *aw), basis)
    # the reference's Horner steps with its product order (for a degree-one
    # argument, its single-term tiles): bit for bit
    assert_exact(out.cpu().numpy(), expected, dtype, f"{op} {shape} {aw}")


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("op", ["exp", "log"])
@pytest.mark.parametrize("ow", [(0, 2), (1, 3), (2, 4), (0, 0)])
def test_ft_exp_log_output_windows(dev, op, ow, dtype):
    w, N = 3, 4
    rng = np.random.default_rng(4)
    basis = rpp_hip.TensorBasis(w, N)
    dim = basis.size()
    n = 3
    arg = damped_batch(rng, n, w, N, dtype)
    init = np.full((n, dim), np.nan, dtype=dtype)
    expected = init.copy()
    getattr(O.orc, "ft_" + op)(w, N, expected, arg, ow=ow)
    out = to_dev(init, dev)
    getattr(rpp_hip, "ft_" + op)(rpp_hip.Batch(out, *ow), to_dev(arg, dev), basis)
    got = out.cpu().numpy()
    m = window_mask(w, N, ow)
    assert np.all(np.isnan(got[:, ~m]))
    assert_exact(got[:, m], expected[:, m], dtype, f"{op} out {ow}")


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("windows", [((0, 4), (0, 4)), ((1, 3), (2, 4)), ((3, 4), (0, 2)), ((2, 2), (2, 2))])
def test_tensor_pairing(dev, windows, dtype):
    fw, xw = windows
    w, N = 3, 4
    rng = np.random.default_rng(21)
    basis = rpp_hip.TensorBasis(w, N)
    dim = basis.size()
    n = 9
    f = random_batch(rng, n, dim, dtype)
    x = random_batch(rng, n, dim, dtype)
    expected = np.full(n, np.nan, dtype=dtype)
    O.orc.tensor_pairing(w, N, expected, f, x, fw=fw, xw=xw)
    out = to_dev(np.full(n, np.nan, dtype=dtype), dev)
    rpp_hip.tensor_pairing(out, rpp_hip.Batch(to_dev(f, dev), *fw), rpp_hip.Batch(to_dev(x, dev), *xw), basis)
    got = out.cpu().numpy()
    if fw[0] > xw[1] or xw[0] > fw[1]:
        assert np.all(got == 0)  # R8: empty overlap writes 0
    assert_close(got, expected, dtype, f"pairing {windows}")


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("side", ["lmul", "rmul"])
@pytest.mark.parametrize("windows", [((0, 4), (0, 4), (0, 4)), ((1, 3), (0, 2), (2, 4)),
                                     ((0, 2), (1, 1), (1, 4)), ((2, 4), (0, 0), (0, 4))])
def test_adjoint_multiplications(dev, side, windows, dtype):
    ow, pw, gw = windows
    w, N = 3, 4
    rng = np.random.default_rng(31)
    basis = rpp_hip.TensorBasis(w, N)
    dim = basis.size()
    n = 4
    op = random_batch(rng, n, dim, dtype)
    arg = random_batch(rng, n, dim, dtype)
    init = np.full((n, dim), np.nan, dtype=dtype)
    expected = init.copy()
    getattr(O.orc, "ft_adj_" + side)(w, N, expected, op, arg, ow=ow, pw=pw, gw=gw)
    out = to_dev(init, dev)
    getattr(rpp_hip, "ft_adj_" + side)(rpp_hip.Batch(out, *ow), rpp_hip.Batch(to_dev(op, dev), *pw),
                                       rpp_hip.Batch(to_dev(arg, dev), *gw), basis)
    got = out.cpu().numpy()
    m = window_mask(w, N, ow)
    assert np.all(np.isnan(got[:, ~m]))
    assert_exact(got[:, m], expected[:, m], dtype, f"adj_{side} {windows}")


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", [(5, 6, 64, 16), (6, 6, 24, 6), (8, 5, 24, 6), (3, 8, 64, 8), (4, 4, 300, 8)],
                         ids=["w5d6", "w6d6", "w8d5", "w3d8", "w4d4"])
def test_chained_fmexp_c2_shape(dev, shape, dtype):
    """The plain chain at every shape with a streaming tile (width 5,
    depth 6 is the headline shape): paths x increments of chained ft_fmexp
    against the oracle chain, bit for bit."""
    w, N, n, L = shape
    rng = np.random.default_rng(0xC2 + w * 10 + N)
    basis = rpp_hip.TensorBasis(w, N)
    dim = basis.size()
    inc = rng.normal(0.0, np.sqrt(1.0 / 128), size=(n, L, w))
    expected = O.signature(w, N, inc.reshape(n * L, w), np.arange(n + 1) * L, dtype=dtype)
    padded = np.zeros((n, 1 + L * w), dtype=dtype)
    padded[:, 1:] = inc.reshape(n, L * w)
    pad_d = to_dev(padded, dev)
    cur = torch.zeros((n, dim), dtype=pad_d.dtype, device=dev)
    cur[:, 0] = 1
    nxt = torch.empty_like(cur)
    for t in range(L):
        rpp_hip.ft_fmexp(nxt, cur, rpp_hip.Batch(pad_d[:, t * w: t * w + 1 + w], 1, 1), basis)
        cur, nxt = nxt, cur
    assert_exact(cur.cpu().numpy(), expected, dtype, "C2 chain")


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", [(5, 6), (6, 6), (8, 5), (3, 8), (4, 4)], ids=["w5d6", "w6d6", "w8d5", "w3d8", "w4d4"])
@pytest.mark.parametrize("mw", [(0, 0), (1, 0), (0, -1), (2, 0), (2, -1)],
                         ids=["full", "from1", "low", "from2", "mid"])
@pytest.mark.parametrize("n", [1, 7])
def test_ft_fmexp_streaming_tiles_multiplier_windows(dev, shape, mw, n, dtype):
    """The streaming tiles (subtree items or sibling-group wavefronts, by
    shape and dtype) with a degree-one exponent and a full output window, for
    multiplier windows that drop low, high and middle degrees (zero-extended,
    never read outside: NaN there), against the oracle bit for bit."""
    w, N = shape
    assert rpp_hip.fmexp_tile(w, N, torch.float64 if dtype == np.float64 else torch.float32) is not None
    lo, hi = mw[0], N + mw[1]
    rng = np.random.default_rng(700 + 10 * w + N + n)
    basis = rpp_hip.TensorBasis(w, N)
    dim = basis.size()
    mult = damped_batch(rng, n, w, N, dtype)
    mult[:, ~window_mask(w, N, (lo, hi))] = np.nan
    ex = deg1(rng, n, w, N, dtype)
    ex[:, 0] = np.nan
    expected = np.full((n, dim), np.nan, dtype=dtype)
    O.orc.ft_fmexp(w, N, expected, mult, ex, mw=(lo, hi), ew=(1, 1))
    out = to_dev(np.full((n, dim), np.nan, dtype=dtype), dev)
    rpp_hip.ft_fmexp(out, rpp_hip.Batch(to_dev(mult, dev), lo, hi), rpp_hip.Batch(to_dev(ex, dev), 1, 1), basis)
    assert_exact(out.cpu().numpy(), expected, dtype, f"fmexp tile {shape} mult {(lo, hi)}")


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", [(3, 4), (5, 6), (2, 7)])
def test_ft_fmexp_interleaved_and_broadcast(dev, shape, dtype):
    """Batch-interleaved multiplier/out with a broadcast degree-one exponent
    (the strided kernels) against the oracle."""
    w, N = shape
    rng = np.random.default_rng(50 + w)
    basis = rpp_hip.TensorBasis(w, N)
    dim = basis.size()
    n = 9
    mult = damped_batch(rng, n, w, N, dtype)
    ex = np.zeros(dim, dtype=dtype)
    ex[1:1 + w] = rng.normal(0, 0.4, size=w)
    expected = np.zeros((n, dim), dtype=dtype)
    O.orc.ft_fmexp(w, N, expected, mult, ex, ew=(1, 1))
    out = torch.zeros((dim, n), dtype=getattr(torch, np.dtype(dtype).name), device=dev)
    rpp_hip.ft_fmexp(rpp_hip.interleaved(out, basis), rpp_hip.interleaved(to_dev(mult.T, dev), basis),
                     rpp_hip.Batch(to_dev(ex, dev), 1, 1), basis)
    assert_exact(out.cpu().numpy().T, expected, dtype, f"interleaved fmexp {shape}")
    # object-major multiplier, broadcast exponent (Horner kernel)
    out2 = torch.zeros((n, dim), dtype=out.dtype, device=dev)
    rpp_hip.ft_fmexp(out2, to_dev(mult, dev), rpp_hip.Batch(to_dev(ex, dev), 1, 1), basis)
    assert_exact(out2.cpu().numpy(), expected, dtype, f"broadcast fmexp {shape}")


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", [(5, 6), (6, 6), (8, 5), (3, 8), (4, 4), (2, 6), (3, 4), (7, 2), (16, 2)])
@pytest.mark.parametrize("xw", [(1, 1), (0, 1)], ids=["x11", "x01"])
def test_ft_fmexp_adjoint(dev, shape, xw, dtype):
    """Adjoint of one ft_fmexp step (register tiles at the compiled shapes,
    the staged kernel elsewhere) against the oracle composed from the reference
    adjoints (ft_adj_lmul/ft_adj_rmul and the exp Jacobian)."""
    w, N = shape
    rng = np.random.default_rng(300 + 10 * w + N)
    basis = rpp_hip.TensorBasis(w, N)
    dim = basis.size()
    n = 5
    S = damped_batch(rng, n, w, N, dtype)
    G = damped_batch(rng, n, w, N, dtype)
    x = rng.normal(0, 0.5, size=(n, w)).astype(dtype)
    gm_ref, gx_ref = O.fmexp_vjp(w, N, S, x, G)
    ex = np.zeros((n, 1 + w), dtype=dtype)
    ex[:, 0] = 7.0  # junk at degree 0: the adjoint reads degree 1 only
    ex[:, 1:] = x
    gm = to_dev(np.full((n, dim), np.nan, dtype=dtype), dev)
    gx = to_dev(np.full((n, 1 + w), np.nan, dtype=dtype), dev)
    rpp_hip.ft_fmexp_adjoint(gm, rpp_hip.Batch(gx, *xw), to_dev(S, dev), rpp_hip.Batch(to_dev(ex, dev), *xw),
                             to_dev(G, dev), basis)
    got_x = gx.cpu().numpy()
    if xw[0] == 0:
        assert np.all(got_x[:, 0] == 0)
    else:
        assert np.all(np.isnan(got_x[:, 0]))
    assert_close(gm.cpu().numpy(), gm_ref, dtype, f"fmexp adjoint grad_mult {shape}")
    assert_close(got_x[:, 1:], gx_ref, dtype, f"fmexp adjoint grad_x {shape}")


def test_ft_fmexp_adjoint_rejects_aliasing_and_windows(dev):
    w, N = 3, 4
    basis = rpp_hip.TensorBasis(w, N)
    dim = basis.size()
    S = torch.zeros((2, dim), dtype=torch.float64, device=dev)
    G = torch.zeros_like(S)
    gm = torch.zeros_like(S)
    ex = torch.zeros((2, 1 + w), dtype=torch.float64, device=dev)
    gx = torch.zeros_like(ex)
    with pytest.raises(rpp_hip.RppHipError):
        rpp_hip.ft_fmexp_adjoint(G, rpp_hip.Batch(gx, 1, 1), S, rpp_hip.Batch(ex, 1, 1), G, basis)
    with pytest.raises(rpp_hip.RppHipError):
        rpp_hip.ft_fmexp_adjoint(gm, rpp_hip.Batch(gx, 1, 1), S, rpp_hip.Batch(S, 1, 2), G, basis)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", [(3, 4), (2, 6), (4, 4), (3, 8), (5, 3), (16, 2)])
@pytest.mark.parametrize("windows", [((1, 99), (0, 99)), ((0, 99), (0, 99)), ((2, 3), (0, 99)),
                                     ((1, 99), (1, 3)), ((1, 99), (0, 2))],
                         ids=["a1N", "a0N", "a23", "g13", "g02"])
def test_ft_log_adjoint(dev, shape, windows, dtype):
    """Adjoint of ft_log (reverse Horner on the GPU) against the oracle's
    term-by-term series derivative built from the reference adjoints."""
    w, N = shape
    aw, gw = windows
    aw = (aw[0], min(aw[1], N))
    gw = (gw[0], min(gw[1], N))
    if aw[0] > N or gw[0] > N:
        pytest.skip("window beyond depth")
    rng = np.random.default_rng(500 + 10 * w + N)
    basis = rpp_hip.TensorBasis(w, N)
    dim = basis.size()
    n = 4
    X = damped_batch(rng, n, w, N, dtype)
    X[:, 0] = 1
    G = damped_batch(rng, n, w, N, dtype)
    am, gm = window_mask(w, N, aw), window_mask(w, N, gw)
    expected = O.log_vjp(w, N, np.where(am, X, 0), np.where(gm, G, 0))
    got = to_dev(np.full((n, dim), np.nan, dtype=dtype), dev)
    rpp_hip.ft_log_adjoint(rpp_hip.Batch(got, *aw), rpp_hip.Batch(to_dev(X, dev), *aw),
                           rpp_hip.Batch(to_dev(G, dev), *gw), basis)
    g = got.cpu().numpy()
    assert np.all(np.isnan(g[:, ~am]))
    assert_close(g[:, am], expected[:, am], dtype, f"log adjoint {shape} {windows}")


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("op", ["tensor_antipode", "tensor_reflect"])
@pytest.mark.parametrize("shape", [(2, 5), (3, 4), (5, 6), (7, 3), (16, 2)])
@pytest.mark.parametrize("windows", [((0, 99), (0, 99)), ((1, 3), (0, 99)), ((0, 99), (2, 3))],
                         ids=["full", "out13", "arg23"])
def test_antipode_reflect(dev, op, shape, windows, dtype):
    w, N = shape
    ow = (windows[0][0], min(windows[0][1], N))
    aw = (windows[1][0], min(windows[1][1], N))
    rng = np.random.default_rng(700 + w * 10 + N)
    basis = rpp_hip.TensorBasis(w, N)
    dim = basis.size()
    n = 3
    arg = random_batch(rng, n, dim, dtype)
    init = np.full((n, dim), np.nan, dtype=dtype)
    expected = init.copy()
    getattr(O.orc, op)(w, N, expected, arg, ow=ow, aw=aw)
    out = to_dev(init, dev)
    getattr(rpp_hip, op)(rpp_hip.Batch(out, *ow), rpp_hip.Batch(to_dev(arg, dev), *aw), basis)
    got = out.cpu().numpy()
    m = window_mask(w, N, ow)
    assert np.all(np.isnan(got[:, ~m]))
    assert np.array_equal(got[:, m], expected[:, m]), f"{op} {shape} {windows}"


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", [(2, 8), (2, 10), (2, 12), (3, 6), (3, 7), (3, 8), (4, 5), (4, 6), (4, 7), (5, 4),
                                   (5, 5), (5, 6), (6, 4), (6, 5), (8, 4), (8, 5)])
@pytest.mark.parametrize("aw", [(0, 99), (1, 99), (2, 99), (1, 3)], ids=["a0N", "a1N", "a2N", "a13"])
def test_ft_log_lds_shapes(dev, shape, aw, dtype):
    """ft_log at the shapes with an LDS-resident kernel (full out window),
    argument windows included, against the oracle: bit for bit, as the kernel
    sums in the reference's CPU product order with contraction off."""
    w, N = shape
    aw = (aw[0], min(aw[1], N))
    rng = np.random.default_rng(900 + 10 * w + N)
    basis = rpp_hip.TensorBasis(w, N)
    dim = basis.size()
    n = 5
    arg = damped_batch(rng, n, w, N, dtype)
    expected = np.full((n, dim), np.nan, dtype=dtype)
    O.orc.ft_log(w, N, expected, arg, aw=aw)
    out = to_dev(np.full((n, dim), np.nan, dtype=dtype), dev)
    rpp_hip.ft_log(out, rpp_hip.Batch(to_dev(arg, dev), *aw), basis)
    got = out.cpu().numpy()
    assert_close(got, expected, dtype, f"log {shape} {aw}")
    assert np.array_equal(got, expected), f"log {shape} {aw}: not bit-identical, {rel_err(got, expected):.3e}"


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape,n", [((3, 8), 2600), ((2, 8), 9000), ((5, 6), 1100)], ids=["w3d8", "w2d8", "w5d6"])
def test_ft_log_lds_many_objects_per_work_group(dev, shape, n, dtype):
    """More objects than the compute units hold work-groups (at most 8 per
    unit at these shapes), so each work-group logs several objects in turn:
    the next object's staging loads, the alternating copies of P and the
    reuse of LDS across objects, against the oracle bit for bit."""
    w, N = shape
    rng = np.random.default_rng(1300 + 10 * w + N)
    basis = rpp_hip.TensorBasis(w, N)
    dim = basis.size()
    arg = damped_batch(rng, n, w, N, dtype)
    expected = np.full((n, dim), np.nan, dtype=dtype)
    O.orc.ft_log(w, N, expected, arg)
    out = to_dev(np.full((n, dim), np.nan, dtype=dtype), dev)
    rpp_hip.ft_log(out, to_dev(arg, dev), basis)
    got = out.cpu().numpy()
    assert np.array_equal(got, expected), f"log {shape} x {n}: not bit-identical, {rel_err(got, expected):.3e}"
