"""GPU parity of the batched Chen recursion (forward) and its backward pass."""
import numpy as np
import pytest
import torch

import oracle as O
import rpp_hip
from helpers import assert_close, rel_err

pytestmark = pytest.mark.gpu


def paths(rng, n, w, lo, hi, dtype=np.float64):
    lengths = rng.integers(lo, hi + 1, size=n)
    offsets = np.zeros(n + 1, dtype=np.int64)
    offsets[1:] = np.cumsum(lengths)
    inc = np.concatenate([rng.normal(0.0, np.sqrt(1.0 / max(L, 1)), size=(L, w)) for L in lengths]) \
        if offsets[-1] > 0 else np.zeros((0, w))
    return inc.astype(dtype), offsets


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(2, 5), (3, 4), (4, 4), (5, 3), (5, 6), (6, 3), (8, 3), (7, 3)])
def test_signature_uneven(dev, shape, dtype):
    w, N = shape
    rng = np.random.default_rng(w * 100 + N)
    inc, off = paths(rng, 13, w, 0, 9, dtype)
    basis = rpp_hip.TensorBasis(w, N)
    expected = O.signature(w, N, inc.astype(np.float64), off)
    got = rpp_hip.signature(torch.tensor(inc, device=dev), torch.tensor(off, device=dev), basis)
    assert_close(got.cpu().numpy(), expected.astype(dtype) if dtype == np.float64 else expected,
                 dtype, f"signature {shape}")


def test_signature_tiles_match_reference_chain_fp64(dev):
    """Every compiled forward tile at width 5, depth 6 (and the storage-generic
    kernel) against the reference's own chain of ft_fmexp, from the golden
    fixture it wrote (2 paths x 16 increments)."""
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "signature_w5d6.npz")
    with np.load(path) as z:
        inc, expected = z["increments"], z["out"]
    n, L, w = inc.shape
    basis = rpp_hip.TensorBasis(w, 6)
    inc_d = torch.tensor(inc.reshape(n * L, w), device=dev)
    off_d = torch.arange(n + 1, device=dev, dtype=torch.int64) * L
    tiles = [t for t in _tiles(torch.float64) if (t[0], t[1]) == (5, 6)] + [(5, 6, -1, 0, 0, 0)]
    assert len(tiles) > 2
    for (_, _, nt, split, kind, expand) in tiles:
        got = rpp_hip.signature(inc_d, off_d, basis, tile_threads=nt, tile_split=split, tile_kind=kind,
                                tile_expand=expand)
        assert_close(got.cpu().numpy(), expected, np.float64, f"tile {(nt, split, kind, expand)} vs reference")


@pytest.mark.parametrize("shape", [(2, 5), (3, 4), (4, 3), (5, 3), (6, 2)])
@pytest.mark.parametrize("ckpt", [0, 3])
def test_signature_backward_fp64(dev, shape, ckpt):
    w, N = shape
    rng = np.random.default_rng(5 + w * 7 + N)
    n = 6
    inc, off = paths(rng, n, w, 1, 8)
    basis = rpp_hip.TensorBasis(w, N)
    dim = basis.size()
    ell = rng.normal(size=dim)
    expected = O.signature_vjp(w, N, inc, off, ell)
    inc_d = torch.tensor(inc, device=dev)
    off_d = torch.tensor(off, device=dev)
    kwargs = {}
    if ckpt:
        lengths = np.diff(off)
        nck = lengths // ckpt
        ck_off = np.zeros(n + 1, dtype=np.int64)
        ck_off[1:] = np.cumsum(nck)
        ckpts = torch.zeros((max(int(ck_off[-1]), 1), dim), dtype=torch.float64, device=dev)
        kwargs = dict(checkpoint_every=ckpt, checkpoints=ckpts, ckpt_offsets=torch.tensor(ck_off, device=dev))
    sig = rpp_hip.signature(inc_d, off_d, basis, **kwargs)
    grad = rpp_hip.signature_backward(inc_d, off_d, basis, sig, torch.tensor(ell, device=dev), **kwargs)
    err = rel_err(grad.cpu().numpy(), expected)
    assert err < 1e-10, f"backward {shape} ckpt={ckpt}: {err:.3e}"


def test_signature_backward_finite_differences(dev):
    """Central differences of <ell, S(path)> in FP64 on the oracle forward."""
    w, N, L = 3, 4, 5
    rng = np.random.default_rng(99)
    inc = rng.normal(0, 0.4, size=(L, w))
    off = np.array([0, L], dtype=np.int64)
    basis = rpp_hip.TensorBasis(w, N)
    ell = rng.normal(size=basis.size())
    h = 1e-6
    fd = np.zeros_like(inc)
    for t in range(L):
        for a in range(w):
            p = inc.copy(); p[t, a] += h
            m = inc.copy(); m[t, a] -= h
            fd[t, a] = (O.signature(w, N, p, off)[0] @ ell - O.signature(w, N, m, off)[0] @ ell) / (2 * h)
    inc_d = torch.tensor(inc, device=dev)
    off_d = torch.tensor(off, device=dev)
    sig = rpp_hip.signature(inc_d, off_d, basis)
    grad = rpp_hip.signature_backward(inc_d, off_d, basis, sig, torch.tensor(ell, device=dev))
    assert rel_err(grad.cpu().numpy(), fd) < 1e-7


def test_signature_empty_paths(dev):
    w, N = 3, 3
    basis = rpp_hip.TensorBasis(w, N)
    off = torch.tensor([0, 0, 2, 2], dtype=torch.int64, device=dev)
    inc = torch.tensor(np.random.default_rng(0).normal(size=(2, w)), device=dev)
    sig = rpp_hip.signature(inc, off, basis).cpu().numpy()
    unit = np.zeros(basis.size()); unit[0] = 1
    assert np.array_equal(sig[0], unit) and np.array_equal(sig[2], unit)
    expected = O.signature(w, N, inc.cpu().numpy(), off.cpu().numpy())
    assert_close(sig, expected, np.float64, "empty paths")


def _tiles(dtype):
    return rpp_hip.signature_tiles(dtype)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
def test_every_register_tile_forward(dev, dtype):
    """Each compiled tile (and the storage-generic kernel) against the FP64
    oracle on uneven paths, lengths 0..6."""
    npd = np.float64 if dtype == torch.float64 else np.float32
    for (w, N, nt, split, kind, expand) in _tiles(dtype):
        rng = np.random.default_rng(w * 1000 + N * 10 + nt + split + kind)
        inc, off = paths(rng, 7, w, 0, 6)
        basis = rpp_hip.TensorBasis(w, N)
        expected = O.signature(w, N, inc, off)
        inc_d = torch.tensor(inc, dtype=dtype, device=dev)
        off_d = torch.tensor(off, device=dev)
        got = rpp_hip.signature(inc_d, off_d, basis, tile_threads=nt, tile_split=split, tile_kind=kind,
                                tile_expand=expand)
        assert_close(got.double().cpu().numpy(), expected, npd, f"tile {(w, N, nt, split, kind, expand)}")
        gen = rpp_hip.signature(inc_d, off_d, basis, tile_threads=-1)
        assert_close(gen.double().cpu().numpy(), expected, npd, f"generic {(w, N)}")


@pytest.mark.parametrize("ckpt", [0, 2])
def test_every_register_tile_backward_fp64(dev, ckpt):
    """Each backward tile (register tiles fixing at most one letter, and the
    thread-owned subtree tiles that have a backward) against the oracle's
    gradient, with reconstruction and with checkpoints."""
    for (w, N, nt, split, kind, expand) in _tiles(torch.float64):
        if not ((kind == 0 and split <= 1) or (kind == 2 and expand == 1)):
            continue
        rng = np.random.default_rng(w * 100 + N + nt + split)
        n = 4
        inc, off = paths(rng, n, w, 1, 5)
        basis = rpp_hip.TensorBasis(w, N)
        dim = basis.size()
        ell = rng.normal(size=dim)
        expected = O.signature_vjp(w, N, inc, off, ell)
        inc_d = torch.tensor(inc, device=dev)
        off_d = torch.tensor(off, device=dev)
        kw = {}
        if ckpt:
            nck = np.diff(off) // ckpt
            ck_off = np.zeros(n + 1, dtype=np.int64)
            ck_off[1:] = np.cumsum(nck)
            # NaN where a forward tile leaves a checkpoint word unwritten (the
            # top degree): no backward may depend on it
            kw = dict(checkpoint_every=ckpt,
                      checkpoints=torch.full((max(int(ck_off[-1]), 1), dim), float("nan"), dtype=torch.float64,
                                             device=dev),
                      ckpt_offsets=torch.tensor(ck_off, device=dev))
        sig = rpp_hip.signature(inc_d, off_d, basis, **kw)
        grad = rpp_hip.signature_backward(inc_d, off_d, basis, sig, torch.tensor(ell, device=dev),
                                          tile_threads=nt, tile_split=split, tile_expand=expand, tile_kind=kind,
                                          **kw)
        err = rel_err(grad.cpu().numpy(), expected)
        assert err < 1e-10, f"tile {(w, N, nt, split, kind, expand)} ckpt={ckpt}: {err:.3e}"
        if ckpt:
            gen = rpp_hip.signature_backward(inc_d, off_d, basis, sig, torch.tensor(ell, device=dev),
                                             tile_threads=-1, **kw)
            err = rel_err(gen.cpu().numpy(), expected)
            assert err < 1e-10, f"generic backward after tile forward {(w, N)} ckpt={ckpt}: {err:.3e}"


@pytest.mark.parametrize("device_order", [False, True], ids=["host_order", "device_order"])
def test_backward_paths_not_starting_at_row_zero(dev, device_order):
    """offsets[0] > 0 and rows after offsets[n]: every backward tile and the
    storage-generic kernel write the gradient rows of the paths (against the
    oracle's gradient) and leave the rows outside them untouched (a sentinel
    survives).  The tiled kernels index their partial sums by absolute row."""
    head, tail = 3, 2
    tiles = [(w, N, nt, split, kind, expand) for (w, N, nt, split, kind, expand) in _tiles(torch.float64)
             if (kind == 0 and split <= 1) or (kind == 2 and expand == 1)]
    tiles += [(w, N, -1, 0, 0, 0) for (w, N) in sorted({(t[0], t[1]) for t in tiles})]
    for (w, N, nt, split, kind, expand) in tiles:
        rng = np.random.default_rng(w * 13 + N + max(nt, 0) + split)
        inc, off = paths(rng, 4, w, 1, 5)
        total = int(off[-1])
        full = rng.normal(0, 0.5, size=(head + total + tail, w))
        full[head:head + total] = inc
        off_shift = off + head
        basis = rpp_hip.TensorBasis(w, N)
        ell = rng.normal(size=basis.size())
        expected = O.signature_vjp(w, N, inc, off, ell)
        inc_d = torch.tensor(full, device=dev)
        off_d = torch.tensor(off_shift, device=dev)
        order = rpp_hip.path_order(off_d) if device_order else None
        sig = rpp_hip.signature(inc_d, off_d, basis, order=order)
        assert_close(sig.cpu().numpy(), O.signature(w, N, inc, off), np.float64, f"forward {(w, N)}")
        sentinel = 12345.0
        grad = torch.full_like(inc_d, sentinel)
        rpp_hip.signature_backward(inc_d, off_d, basis, sig, torch.tensor(ell, device=dev), grad=grad,
                                   tile_threads=nt, tile_split=split, tile_expand=expand, tile_kind=kind,
                                   order=order)
        g = grad.cpu().numpy()
        tag = f"tile {(w, N, nt, split, kind, expand)}"
        assert np.all(g[:head] == sentinel) and np.all(g[head + total:] == sentinel), f"{tag}: rows outside written"
        err = rel_err(g[head:head + total], expected)
        assert err < 1e-10, f"{tag}: {err:.3e}"
        # the Python wrapper's own gradient buffer is zero outside the paths
        g2 = rpp_hip.signature_backward(inc_d, off_d, basis, sig, torch.tensor(ell, device=dev),
                                        tile_threads=nt, tile_split=split, tile_expand=expand, tile_kind=kind,
                                        order=order).cpu().numpy()
        assert np.all(g2[:head] == 0) and np.all(g2[head + total:] == 0), f"{tag}: wrapper rows outside"
        assert np.array_equal(g2[head:head + total], g[head:head + total]), f"{tag}: repeat differs"


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
def test_thread_backward_per_path_cotangent(dev, dtype):
    """The thread-owned subtree backward with a cotangent batch (one row per
    path, not a broadcast functional) on uneven paths, an empty one included,
    against the FP64 oracle's gradient path by path (FP32 within 1e-4)."""
    for (w, N, nt, split, kind, expand) in _tiles(dtype):
        if kind != 2 or expand != 1:
            continue
        rng = np.random.default_rng(w * 31 + N + split)
        lengths = np.array([3, 0, 6, 1, 5])
        n = lengths.size
        off = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
        inc = rng.normal(0.0, 0.5, size=(int(off[-1]), w))
        basis = rpp_hip.TensorBasis(w, N)
        dim = basis.size()
        cot = rng.normal(size=(n, dim))
        expected = np.zeros_like(inc)
        for p in range(n):
            a, b = int(off[p]), int(off[p + 1])
            if b > a:
                expected[a:b] = O.signature_vjp(w, N, inc[a:b], np.array([0, b - a]), cot[p])
        inc_d = torch.tensor(inc, device=dev, dtype=dtype)
        off_d = torch.tensor(off, device=dev)
        sig = rpp_hip.signature(inc_d, off_d, basis)
        grad = rpp_hip.signature_backward(inc_d, off_d, basis, sig, torch.tensor(cot, device=dev, dtype=dtype),
                                          tile_threads=nt, tile_split=split, tile_kind=2)
        err = rel_err(grad.double().cpu().numpy(), expected)
        tol = 1e-10 if dtype == torch.float64 else 1e-4
        assert err < tol, f"thread tile {(w, N, nt, split)} {dtype}: {err:.3e}"


def test_autograd_signature_and_ft_mul(dev):
    from rpp_hip import autograd as AG
    w, N, n = 3, 4, 5
    rng = np.random.default_rng(12)
    inc, off = paths(rng, n, w, 1, 7)
    dim = rpp_hip.tensor_dimension(w, N)
    ell = rng.normal(size=dim)
    expected = O.signature_vjp(w, N, inc, off, ell)
    for every in (0, 2):
        x = torch.tensor(inc, device=dev, requires_grad=True)
        sig = AG.signature(x, torch.tensor(off, device=dev), N, checkpoint_every=every)
        loss = (sig @ torch.tensor(ell, device=dev)).sum()
        loss.backward()
        assert rel_err(x.grad.cpu().numpy(), expected) < 1e-10
    # ft_mul: torch.autograd.gradcheck in FP64
    a = torch.tensor(rng.uniform(-1, 1, (2, dim)), device=dev, requires_grad=True)
    b = torch.tensor(rng.uniform(-1, 1, (2, dim)), device=dev, requires_grad=True)
    assert torch.autograd.gradcheck(lambda p, q: AG.ft_mul(p, q, w, N), (a, b), eps=1e-6, atol=1e-8)


@pytest.mark.parametrize("shape", [(3, 4), (5, 6), (2, 5)])
def test_autograd_fmexp_and_log_gradcheck(dev, shape):
    """torch.autograd.gradcheck (FP64) of the ft_fmexp and ft_log bindings,
    i.e. of their adjoint kernels, against finite differences of the forward
    kernels, plus a chained fmexp against the signature gradient."""
    from rpp_hip import autograd as AG
    w, N = shape
    dim = rpp_hip.tensor_dimension(w, N)
    rng = np.random.default_rng(77 + w)
    m = torch.tensor(rng.normal(0, 0.3, (2, dim)), device=dev, requires_grad=True)
    x = torch.tensor(rng.normal(0, 0.5, (2, w)), device=dev, requires_grad=True)
    # fast_mode: random projections of the Jacobian (two forward calls and
    # one backward call per check), so large tensors stay cheap
    fast = dim > 200
    assert torch.autograd.gradcheck(lambda p, q: AG.ft_fmexp(p, q, N), (m, x), eps=1e-6, atol=1e-7,
                                    fast_mode=fast)
    X = rng.normal(0, 0.2, (2, dim))
    X[:, 0] = 1
    a = torch.tensor(X, device=dev, requires_grad=True)
    assert torch.autograd.gradcheck(lambda p: AG.ft_log(p, w, N), (a,), eps=1e-6, atol=1e-7, fast_mode=fast)
    # chained per-op fmexp == the fused signature gradient
    L = 6
    inc = rng.normal(0, 0.4, (L, w))
    ell = torch.tensor(rng.normal(size=dim), device=dev)
    xs = [torch.tensor(inc[t:t + 1], device=dev, requires_grad=True) for t in range(L)]
    S = torch.zeros((1, dim), dtype=torch.float64, device=dev)
    S[0, 0] = 1
    for t in range(L):
        S = AG.ft_fmexp(S, xs[t], N)
    (S[0] @ ell).backward()
    got = torch.cat([v.grad for v in xs]).cpu().numpy()
    expected = O.signature_vjp(w, N, inc, np.array([0, L]), ell.cpu().numpy())
    assert rel_err(got, expected) < 1e-10


def test_chain_tiles_checkpoints_match(dev):
    """Checkpoints written by the two-barrier tiles and the thread-subtree
    tiles equal the signatures of the path prefixes below the top degree (the
    part a backward step reads; the thread-subtree tiles do not write the top
    degree)."""
    for (w, N, nt, split, kind, expand) in _tiles(torch.float64):
        if kind not in (1, 2):
            continue
        rng = np.random.default_rng(nt + split)
        inc, off = paths(rng, 3, w, 4, 7)
        basis = rpp_hip.TensorBasis(w, N)
        dim = basis.size()
        nck = np.diff(off) // 2
        ck_off = np.zeros(4, dtype=np.int64)
        ck_off[1:] = np.cumsum(nck)
        ckpts = torch.zeros((int(ck_off[-1]), dim), dtype=torch.float64, device=dev)
        rpp_hip.signature(torch.tensor(inc, device=dev), torch.tensor(off, device=dev), basis,
                          checkpoint_every=2, checkpoints=ckpts, ckpt_offsets=torch.tensor(ck_off, device=dev),
                          tile_threads=nt, tile_split=split, tile_kind=kind, tile_expand=expand)
        got = ckpts.cpu().numpy()
        below = dim - w ** N
        for p in range(3):
            for j in range(int(nck[p])):
                pre = inc[off[p]: off[p] + 2 * (j + 1)]
                expected = O.signature(w, N, pre, np.array([0, len(pre)]))[0]
                assert_close(got[ck_off[p] + j, :below], expected[:below], np.float64,
                             f"checkpoint tile {(w, N, nt, split)}")


@pytest.mark.parametrize("shape", [(4, 4), (3, 4), (5, 6)])
def test_signature_device_order(dev, shape):
    """A queue order computed on the device gives the same signatures and
    gradients as the library's own (host) ordering."""
    w, N = shape
    rng = np.random.default_rng(40 + w)
    inc, off = paths(rng, 11, w, 0, 12)
    basis = rpp_hip.TensorBasis(w, N)
    inc_d, off_d = torch.tensor(inc, device=dev), torch.tensor(off, device=dev)
    order = rpp_hip.path_order(off_d)
    assert sorted(order.cpu().tolist()) == list(range(11))
    a = rpp_hip.signature(inc_d, off_d, basis)
    b = rpp_hip.signature(inc_d, off_d, basis, order=order)
    assert_close(b.cpu().numpy(), a.cpu().numpy(), np.float64, "device order forward")
    ell = torch.tensor(rng.normal(size=basis.size()), device=dev)
    ga = rpp_hip.signature_backward(inc_d, off_d, basis, a, ell)
    gb = rpp_hip.signature_backward(inc_d, off_d, basis, a, ell, order=order)
    assert_close(gb.cpu().numpy(), ga.cpu().numpy(), np.float64, "device order backward")


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("shape,ckpt", [((4, 4), 0), ((3, 4), 0), ((5, 6), 0), ((4, 4), 4)])
def test_signature_graph_replays(dev, shape, ckpt, dtype):
    """Forward + backward captured in one hipGraph: every replay, with new
    increments and cotangents written into the static buffers and eager work
    (and synchronisations) in between, matches the eager calls."""
    w, N = shape
    rng = np.random.default_rng(60 + w * 3 + N)
    inc, off = paths(rng, 9, w, 1, 14)
    basis = rpp_hip.TensorBasis(w, N)
    dim = basis.size()
    off_d = torch.tensor(off, device=dev)
    inc_d = torch.tensor(inc, device=dev, dtype=dtype)
    ell = torch.tensor(rng.normal(size=dim), device=dev, dtype=dtype)
    g = rpp_hip.SignatureGraph(inc_d, off_d, basis, backward=True, cotangent=ell, checkpoint_every=ckpt)
    npdt = np.float64 if dtype == torch.float64 else np.float32
    for rep in range(3):
        noise = np.random.default_rng(rep).normal(0, 0.05, size=inc.shape)
        new_inc = torch.tensor(inc * (1 + 0.1 * rep) + (noise if rep else 0), device=dev, dtype=dtype)
        g.increments.copy_(new_inc)
        g.cotangent.copy_(ell * (rep + 1))
        junk = torch.randn(1 << 16, device=dev).sum()  # eager work between replays
        torch.cuda.synchronize()
        out, grad = g.replay()
        torch.cuda.synchronize()
        ref = rpp_hip.signature(new_inc, off_d, basis)
        ref_g = rpp_hip.signature_backward(new_inc, off_d, basis, ref, ell * (rep + 1))
        assert_close(out.cpu().numpy(), ref.cpu().numpy(), npdt, f"graph forward replay {rep}")
        tol = 1e-10 if dtype == torch.float64 else 1e-4
        assert rel_err(grad.cpu().numpy(), ref_g.cpu().numpy()) < tol, f"graph backward replay {rep}"
        del junk
    exp64 = O.signature(w, N, g.increments.double().cpu().numpy(), off)
    assert_close(out.cpu().numpy(), exp64, npdt, "graph vs oracle")


@pytest.mark.parametrize("ckpt", [0, 16])
def test_signature_backward_fd_c5_shape(dev, ckpt):
    """Width 6, depth 6, 32 paths of 256 increments N(0, 1/256), loss = <ell, S>
    with ell ~ N(0, 1): the backward pass (exp(-x) reconstruction, and re-anchored
    at sqrt(L) checkpoints) against FP64 central differences of the GPU forward
    at four random coordinates per path."""
    w, N, n, L = 6, 6, 32, 256
    rng = np.random.default_rng(0xC5)
    inc = rng.normal(0, np.sqrt(1.0 / L), size=(n * L, w))
    off = np.arange(n + 1, dtype=np.int64) * L
    basis = rpp_hip.TensorBasis(w, N)
    ell = torch.tensor(rng.normal(size=basis.size()), device=dev)
    inc_d, off_d = torch.tensor(inc, device=dev), torch.tensor(off, device=dev)
    kw = {}
    if ckpt:
        ck_off = torch.arange(n + 1, dtype=torch.int64, device=dev) * (L // ckpt)
        kw = dict(checkpoint_every=ckpt, checkpoints=torch.empty((n * (L // ckpt), basis.size()),
                                                                 dtype=torch.float64, device=dev),
                  ckpt_offsets=ck_off)
    sig = rpp_hip.signature(inc_d, off_d, basis, **kw)
    grad = rpp_hip.signature_backward(inc_d, off_d, basis, sig, ell, **kw).cpu().numpy()
    h = 1e-5
    worst = 0.0
    for _ in range(4):
        t = rng.integers(0, L, size=n)
        a = rng.integers(0, w, size=n)
        rows = np.arange(n) * L + t
        plus, minus = inc.copy(), inc.copy()
        plus[rows, a] += h
        minus[rows, a] -= h
        lp = rpp_hip.signature(torch.tensor(plus, device=dev), off_d, basis) @ ell
        lm = rpp_hip.signature(torch.tensor(minus, device=dev), off_d, basis) @ ell
        fd = ((lp - lm) / (2 * h)).cpu().numpy()
        worst = max(worst, rel_err(grad[rows, a], fd))
    assert worst < 1e-6, f"C5 shape finite differences: {worst:.3e}"


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64], ids=["f32", "f64"])
def test_backward_1024_increments_reconstruction_and_checkpoints(dev, dtype):
    """Width 5, depth 6, two paths of 1024 increments N(0, 1/1024): the
    gradient by exp(-x) reconstruction and by sqrt(L) = 32 checkpoints against
    the FP64 oracle (FP32: reconstruction within 5e-4, checkpoints within
    2e-4; FP64: both within 1e-10), and the
    autograd default ("auto": checkpoints every 32 in FP32, none in FP64)
    equal to the explicit choice bit for bit."""
    from rpp_hip import autograd as AG
    w, N, n, L = 5, 6, 2, 1024
    rng = np.random.default_rng(0x1024)
    inc = rng.normal(0, np.sqrt(1.0 / L), size=(n * L, w))
    off = np.arange(n + 1, dtype=np.int64) * L
    basis = rpp_hip.TensorBasis(w, N)
    dim = basis.size()
    ell = rng.normal(size=dim)
    expected = O.signature_vjp(w, N, inc, off, ell)
    inc_d = torch.tensor(inc, dtype=dtype, device=dev)
    off_d = torch.tensor(off, device=dev)
    ell_d = torch.tensor(ell, dtype=dtype, device=dev)
    errs, grads = {}, {}
    for every in (0, 32):
        kw = {}
        if every:
            kw = dict(checkpoint_every=every, checkpoints=torch.empty((n * (L // every), dim), dtype=dtype, device=dev),
                      ckpt_offsets=torch.arange(n + 1, dtype=torch.int64, device=dev) * (L // every))
        sig = rpp_hip.signature(inc_d, off_d, basis, **kw)
        grads[every] = rpp_hip.signature_backward(inc_d, off_d, basis, sig, ell_d, **kw)
        errs[every] = rel_err(grads[every].double().cpu().numpy(), expected)
    if dtype == torch.float32:
        assert errs[0] < 5e-4 and errs[32] < 2e-4, errs
    else:
        assert errs[0] < 1e-10 and errs[32] < 1e-10, errs
    auto_every = rpp_hip.checkpoint_every_for(dtype, L)
    assert auto_every == (32 if dtype == torch.float32 else 0)
    x = inc_d.clone().requires_grad_(True)
    (AG.signature(x, off_d, N) @ ell_d).sum().backward()
    assert torch.equal(x.grad, grads[auto_every]), "autograd default differs from the explicit checkpoint choice"


_C4_CACHE = {}


def _c4_paths():
    """16 paths of the width-8 depth-5 uneven workload's distribution with
    lengths spanning 16 .. 1024 (both ends included), and their FP64 oracle
    signatures (computed once per session)."""
    if not _C4_CACHE:
        w, N = 8, 5
        rng = np.random.default_rng(0xC4)
        lengths = np.concatenate([[16, 1024], rng.integers(16, 1025, size=14)])
        off = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
        inc = np.concatenate([rng.normal(0.0, np.sqrt(1.0 / Lp), size=(Lp, w)) for Lp in lengths])
        _C4_CACHE.update(lengths=lengths, off=off, inc=inc, expected=O.signature(w, N, inc, off))
    return _C4_CACHE


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64], ids=["f32", "f64"])
def test_uneven_graph_real_lengths(dev, dtype):
    """The uneven-batch path as the benchmark runs it (the fused kernel
    replayed from a hipGraph, autotuned tile) at width 8, depth 5 on paths of
    16 .. 1024 increments: every path within the reference tolerance of the
    FP64 oracle (1e-12 FP64, 1e-5 FP32), the 1024-increment one included, on
    two replays with a rewrite of the increments in between."""
    c = _c4_paths()
    basis = rpp_hip.TensorBasis(8, 5)
    inc_d = torch.tensor(c["inc"], dtype=dtype, device=dev)
    g = rpp_hip.SignatureGraph(inc_d, torch.tensor(c["off"], device=dev), basis)
    npd = np.float64 if dtype == torch.float64 else np.float32
    for rep in range(2):
        if rep:
            g.increments.zero_()
            torch.cuda.synchronize()
            g.replay()
            g.increments.copy_(inc_d)
        out = g.replay().double().cpu().numpy()
        for p, Lp in enumerate(c["lengths"]):
            assert_close(out[p], c["expected"][p], npd, f"replay {rep}, path {p} of {Lp} increments")
