"""The allocation-free signature backward (rpp_hip_signature_backward_noalloc_*)
through the C-ABI: on two calls in a row, the second with a smaller batch,
its gradients equal those of rpp_hip_signature_backward_* (the same thread-
subtree kernels, partial sums in the library's fixed scratch instead of its
workspace) bit for bit; calls it cannot serve are refused before any launch."""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle as O
import rpp_hip
from rpp_hip import _capi as A

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20


def _paths(rng, lengths, w, head=0):
    off = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64) + head
    inc = rng.normal(0.0, 0.4, size=(int(off[-1]) + 2, w))
    return inc, off


def _desc(basis, inc, off, sig, cot, grad, order, tile):
    nt, split = tile
    return A.SignatureBackwardDesc(basis._c(), off.numel() - 1, inc.data_ptr(), off.data_ptr(),
                                   rpp_hip.Batch(sig, 0, basis.depth)._c(), rpp_hip.Batch(cot, 0, basis.depth)._c(),
                                   grad.data_ptr(), 0, None, None, None, nt, split,
                                   order.data_ptr() if order is not None else None, inc.shape[0], 0, 2)


def _noalloc(dtype, desc):
    name = "rpp_hip_signature_backward_noalloc_" + ("f64" if dtype == torch.float64 else "f32")
    return getattr(A.lib(), name)(C.byref(desc), C.c_void_p(torch.cuda.current_stream().cuda_stream))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32], ids=["f64", "f32"])
def test_noalloc_backward_matches_backward(dev, dtype):
    done = 0
    for (w, N, nt, split, kind, expand) in rpp_hip.signature_tiles(dtype):
        if kind != 2 or expand != 1:
            continue
        basis = rpp_hip.TensorBasis(w, N)
        rng = np.random.default_rng(w * 10 + N)
        # first call: 24 paths; second call: 5 shorter paths starting at row 3
        for call, (lengths, head) in enumerate(((rng.integers(1, 30, size=24), 0), (rng.integers(1, 6, size=5), 3))):
            inc_np, off_np = _paths(rng, lengths, w, head)
            inc = torch.tensor(inc_np, dtype=dtype, device=dev)
            off = torch.tensor(off_np, device=dev)
            order = rpp_hip.path_order(off)
            cot = torch.tensor(rng.normal(size=(len(lengths), basis.size())), dtype=dtype, device=dev)
            sig = rpp_hip.signature(inc, off, basis, order=order)
            ref = rpp_hip.signature_backward(inc, off, basis, sig, cot, tile_threads=nt, tile_split=split,
                                             tile_kind=2, order=order)
            got = torch.zeros_like(inc)
            d = _desc(basis, inc, off, sig, cot, got, order, (nt, split))
            rc = _noalloc(dtype, d)
            assert rc == 0, A.lib().rpp_hip_last_error().decode()
            torch.cuda.synchronize()
            tag = f"tile {(w, N, nt, split)} call {call}"
            assert torch.equal(got, ref), f"{tag}: differs from rpp_hip_signature_backward"
            if dtype == torch.float64 and call == 1:
                rows = slice(int(off_np[0]), int(off_np[-1]))
                expected = np.zeros((rows.stop - rows.start, w))
                for p in range(len(lengths)):
                    a, b = int(off_np[p]) - rows.start, int(off_np[p + 1]) - rows.start
                    expected[a:b] = O.signature_vjp(w, N, inc_np[int(off_np[p]):int(off_np[p + 1])],
                                                    np.array([0, b - a]), cot[p].cpu().numpy())
                g = got.cpu().numpy()
                assert np.all(g[:rows.start] == 0) and np.all(g[rows.stop:] == 0), f"{tag}: rows outside written"
                err = np.max(np.abs(g[rows] - expected) / np.maximum(1, np.abs(expected)))
                assert err < 1e-10, f"{tag}: {err:.3e} against the oracle"
        done += 1
    assert done > 0


def test_noalloc_backward_refusals(dev):
    """Refused on the host before any launch: no order, a scratch buffer, a
    non-thread tile kind, and a batch whose partial sums exceed the fixed
    scratch."""
    dtype = torch.float64
    w, N, nt, split = next((t[0], t[1], t[2], t[3]) for t in rpp_hip.signature_tiles(dtype) if t[4] == 2 and t[5] == 1)
    basis = rpp_hip.TensorBasis(w, N)
    rng = np.random.default_rng(1)
    inc_np, off_np = _paths(rng, [3, 4], w)
    inc = torch.tensor(inc_np, device=dev)
    off = torch.tensor(off_np, device=dev)
    order = rpp_hip.path_order(off)
    sig = rpp_hip.signature(inc, off, basis, order=order)
    cot = torch.ones(basis.size(), dtype=dtype, device=dev)
    grad = torch.full_like(inc, 7.0)
    d = _desc(basis, inc, off, sig, cot, grad, None, (nt, split))
    assert _noalloc(dtype, d) == 3 and b"order" in A.lib().rpp_hip_last_error()
    d = _desc(basis, inc, off, sig, cot, grad, order, (nt, split))
    d.scratch = grad.data_ptr()
    assert _noalloc(dtype, d) == 3
    d = _desc(basis, inc, off, sig, cot, grad, order, (nt, split))
    d.tile_kind = 0
    assert _noalloc(dtype, d) == 8
    d = _desc(basis, inc, off, sig, cot, grad, order, (nt, split))
    d.num_increments = SCRATCH_BYTES // (w * w * 8) + 1
    assert _noalloc(dtype, d) == 3 and b"scratch" in A.lib().rpp_hip_last_error()
    torch.cuda.synchronize()
    assert torch.all(grad == 7.0), "a refused call wrote the gradient"
