"""GPU parity of sparse_matrix_vector_product, tensor_to_lie and lie_to_tensor."""
import numpy as np
import pytest
import torch

import oracle as O
import rpp_hip
from helpers import assert_close, damped_batch

pytestmark = pytest.mark.gpu


def to_dev(x, dev):
    return torch.tensor(np.ascontiguousarray(x), device=dev)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("fmt", ["csr", "csc"])
def test_spmv_view_relative(dev, fmt, dtype):
    """R9: row 0 <-> out.begin(), column 0 <-> arg.begin()
    (sparse_matrix_vector_product_tests.cpp:192-226)."""
    rng = np.random.default_rng(3)
    rows, cols, n = 7, 9, 5
    dense = rng.uniform(-1, 1, size=(rows, cols)) * (rng.uniform(size=(rows, cols)) < 0.4)
    if fmt == "csr":
        offs = np.concatenate([[0], np.cumsum((dense != 0).sum(axis=1))])
        idx = np.concatenate([np.nonzero(dense[r])[0] for r in range(rows)])
        vals = np.concatenate([dense[r][dense[r] != 0] for r in range(rows)])
    else:
        offs = np.concatenate([[0], np.cumsum((dense != 0).sum(axis=0))])
        idx = np.concatenate([np.nonzero(dense[:, c])[0] for c in range(cols)])
        vals = np.concatenate([dense[:, c][dense[:, c] != 0] for c in range(cols)])
    out_begin, arg_begin, width_out, width_arg = 3, 2, 12, 14
    arg = rng.uniform(-1, 1, size=(n, width_arg)).astype(dtype)
    init = np.full((n, width_out), 7.0, dtype=dtype)
    expected = init.copy()
    O.spmv(vals, idx, offs, rows if fmt == "csr" else cols, expected, arg, csc=(fmt == "csc"),
           out_offset=out_begin, out_size=rows, arg_offset=arg_begin, alpha=0.5)
    M = rpp_hip.SparseMatrix(fmt, to_dev(vals.astype(dtype), dev), to_dev(idx.astype(np.int64), dev),
                             to_dev(offs.astype(np.int64), dev), rows, cols)
    out = to_dev(init, dev)
    rpp_hip.sparse_matrix_vector_product(out, to_dev(arg, dev), M, out_begin=out_begin, arg_begin=arg_begin,
                                         out_size=rows, alpha=0.5)
    got = out.cpu().numpy()
    assert np.all(got[:, :out_begin] == 7) and np.all(got[:, out_begin + rows:] == 7)
    assert_close(got, expected, dtype, f"spmv {fmt}")
    # every row summed in the reference's order (CSR: entry order; CSC: column
    # order, alpha on the argument), contraction off: bit for bit
    assert np.array_equal(got, expected)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_spmv_csc_ragged_exact(dev, dtype):
    """A CSC matrix with empty columns, empty rows, a row named twice in one
    column and an out view longer than the matrix (zeroed past its rows, as the
    reference's set_constant does): equal to the oracle's column scatter bit
    for bit, on repeated calls (the cached CSR form) and after the cache is
    dropped."""
    rng = np.random.default_rng(23)
    rows, cols, n = 40, 300, 9
    offs, idx, vals = [0], [], []
    for c in range(cols):
        k = 0 if c % 7 == 3 else int(rng.integers(0, 4))
        r = list(rng.integers(0, rows - 3, size=k))  # the last three rows stay empty
        if c % 11 == 5 and r:
            r.append(r[0])  # the same row twice in one column
        idx += r
        vals += list(rng.uniform(-1, 1, size=len(r)))
        offs.append(len(idx))
    idx, vals, offs = (np.array(a) for a in (idx, vals, offs))
    arg = rng.uniform(-1, 1, size=(n, cols)).astype(dtype)
    out_size = rows + 5
    expected = np.full((n, out_size + 2), 7.0, dtype=dtype)
    O.spmv(vals, idx, offs, cols, expected, arg, csc=True, out_offset=1, out_size=out_size, alpha=-1.25)
    M = rpp_hip.SparseMatrix("csc", to_dev(vals.astype(dtype), dev), to_dev(idx.astype(np.int64), dev),
                             to_dev(offs.astype(np.int64), dev), rows, cols)
    arg_d = to_dev(arg, dev)
    for rep in range(3):
        if rep == 2:
            rpp_hip.sparse_matrix_forget(M)
        out = to_dev(np.full((n, out_size + 2), 7.0, dtype=dtype), dev)
        rpp_hip.sparse_matrix_vector_product(out, arg_d, M, out_begin=1, out_size=out_size, alpha=-1.25)
        assert np.array_equal(out.cpu().numpy(), expected), f"call {rep}"


def _hall_matrix(lie, which, dtype, dev):
    m = lie[which]
    ts, dim = lie["true_size"], lie["tensor_dim"]
    rows, cols = (ts - 1, dim) if which == "t2l" else (dim, ts - 1)
    return rpp_hip.SparseMatrix("csc", to_dev(m["data"].astype(dtype), dev), to_dev(m["indices"], dev),
                                to_dev(m["offsets"], dev), rows, cols)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_reference_csc_matrices_4099_objects(dev, dtype):
    """The Hall T2L and L2T in the CSC format of the reference's builders
    (byte for byte make_t2l_matrix / make_l2t_matrix, tests/test_lie_cpu.py)
    at width 3, depth 8 over 4099 objects: sparse_matrix_vector_product
    (alpha 1 and 0.5), and tensor_to_lie / lie_to_tensor given those matrices
    (served by the Lie kernel once recognised), equal the oracle's CSC column
    scatter bit for bit, and the library's own conversions; a matrix that
    differs from the Hall one in one value takes the sparse product and still
    equals the oracle."""
    w, N, n = 3, 8, 4099
    lie = rpp_hip.lie_basis(w, N)
    ts, dim = lie["true_size"], lie["tensor_dim"]
    basis = rpp_hip.TensorBasis(w, N)
    rng = np.random.default_rng(4099)
    x = damped_batch(rng, n, w, N, dtype)
    x_d = to_dev(x, dev)
    T2L, L2T = _hall_matrix(lie, "t2l", dtype, dev), _hall_matrix(lie, "l2t", dtype, dev)
    t2l, l2t = lie["t2l"], lie["l2t"]
    exp_lie = {}
    for alpha in (1.0, 0.5):
        e = np.full((n, ts), 7.0, dtype=dtype)
        O.spmv(t2l["data"].astype(dtype), t2l["indices"], t2l["offsets"], dim, e, x, csc=True, out_offset=1,
               out_size=ts - 1, alpha=alpha)
        out = to_dev(np.full((n, ts), 7.0, dtype=dtype), dev)
        rpp_hip.sparse_matrix_vector_product(out, x_d, T2L, out_begin=1, out_size=ts - 1, alpha=alpha)
        assert np.array_equal(out.cpu().numpy(), e), f"spmv T2L alpha {alpha}"
        exp_lie[alpha] = e
    lie_d = to_dev(np.full((n, ts), 7.0, dtype=dtype), dev)
    rpp_hip.tensor_to_lie(lie_d, x_d, basis, matrix=T2L)
    got = lie_d.cpu().numpy()
    assert np.all(got[:, 0] == 7)
    assert np.array_equal(got[:, 1:], exp_lie[1.0][:, 1:]), "tensor_to_lie with the reference CSC T2L"
    own = to_dev(np.zeros((n, ts), dtype=dtype), dev)
    rpp_hip.tensor_to_lie(own, x_d, basis)
    assert np.array_equal(own.cpu().numpy()[:, 1:], got[:, 1:]), "own T2L"
    e = np.full((n, dim), np.nan, dtype=dtype)
    O.spmv(l2t["data"].astype(dtype), l2t["indices"], l2t["offsets"], ts - 1, e, got, csc=True, out_offset=0,
           out_size=dim, arg_offset=1)
    t_d = to_dev(np.full((n, dim), np.nan, dtype=dtype), dev)
    rpp_hip.lie_to_tensor(t_d, lie_d, basis, matrix=L2T)
    assert np.array_equal(t_d.cpu().numpy(), e), "lie_to_tensor with the reference CSC L2T"
    # not the Hall matrix: one value changed
    vals = t2l["data"].astype(dtype).copy()
    vals[len(vals) // 2] *= 2
    P = rpp_hip.SparseMatrix("csc", to_dev(vals, dev), to_dev(t2l["indices"], dev), to_dev(t2l["offsets"], dev),
                             ts - 1, dim)
    e = np.full((n, ts), 7.0, dtype=dtype)
    O.spmv(vals, t2l["indices"], t2l["offsets"], dim, e, x, csc=True, out_offset=1, out_size=ts - 1)
    pl = to_dev(np.full((n, ts), 7.0, dtype=dtype), dev)
    rpp_hip.tensor_to_lie(pl, x_d, basis, matrix=P)
    assert np.array_equal(pl.cpu().numpy(), e), "tensor_to_lie with a non-Hall matrix"
    assert not np.array_equal(pl.cpu().numpy(), got)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_spmv_csr_staged_argument_exact(dev, dtype):
    """CSR product with enough rows for the LDS-staged kernel (argument in
    LDS, one thread per row): equal to the oracle's row loop bit for bit."""
    rng = np.random.default_rng(11)
    rows, cols, n = 300, 500, 7
    dense = rng.uniform(-1, 1, size=(rows, cols)) * (rng.uniform(size=(rows, cols)) < 0.05)
    offs = np.concatenate([[0], np.cumsum((dense != 0).sum(axis=1))])
    idx = np.concatenate([np.nonzero(dense[r])[0] for r in range(rows)])
    vals = np.concatenate([dense[r][dense[r] != 0] for r in range(rows)])
    arg = rng.uniform(-1, 1, size=(n, cols)).astype(dtype)
    expected = np.zeros((n, rows), dtype=dtype)
    O.spmv(vals, idx, offs, rows, expected, arg, csc=False, out_offset=0, out_size=rows, alpha=0.75)
    M = rpp_hip.SparseMatrix("csr", to_dev(vals.astype(dtype), dev), to_dev(idx.astype(np.int64), dev),
                             to_dev(offs.astype(np.int64), dev), rows, cols)
    out = to_dev(np.zeros((n, rows), dtype=dtype), dev)
    rpp_hip.sparse_matrix_vector_product(out, to_dev(arg, dev), M, out_begin=0, arg_begin=0, out_size=rows,
                                         alpha=0.75)
    assert np.array_equal(out.cpu().numpy(), expected)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(2, 4), (3, 4), (3, 8), (4, 5), (5, 3)])
def test_tensor_to_lie_and_back(dev, shape, dtype):
    w, N = shape
    rng = np.random.default_rng(w + N)
    basis = rpp_hip.TensorBasis(w, N)
    lie = rpp_hip.lie_basis(w, N)
    ts, dim = lie["true_size"], lie["tensor_dim"]
    n = 6
    x = damped_batch(rng, n, w, N, dtype)
    # oracle: CSC products with the (reference-identical) matrices
    t2l = lie["t2l"]
    exp_lie = np.zeros((n, ts), dtype=dtype)
    O.spmv(t2l["data"], t2l["indices"], t2l["offsets"], dim, exp_lie, x, csc=True, out_offset=1,
           out_size=ts - 1)
    lie_out = to_dev(np.zeros((n, ts), dtype=dtype), dev)
    rpp_hip.tensor_to_lie(lie_out, to_dev(x, dev), basis)
    got_lie = lie_out.cpu().numpy()
    assert_close(got_lie, exp_lie, dtype, f"t2l {shape}")
    # each row is summed in the reference's order: bit for bit
    assert np.array_equal(got_lie, exp_lie), f"t2l {shape} not bit-identical"
    l2t = lie["l2t"]
    exp_t = np.zeros((n, dim), dtype=dtype)
    O.spmv(l2t["data"], l2t["indices"], l2t["offsets"], ts - 1, exp_t, got_lie, csc=True, out_offset=0,
           out_size=dim, arg_offset=1)
    t_out = to_dev(np.full((n, dim), np.nan, dtype=dtype), dev)
    rpp_hip.lie_to_tensor(t_out, lie_out, basis)
    assert_close(t_out.cpu().numpy(), exp_t, dtype, f"l2t {shape}")
    assert np.array_equal(t_out.cpu().numpy(), exp_t), f"l2t {shape} not bit-identical"


def _lie_expected(lie, x_zero, to_lie, ts, dim, dtype):
    if to_lie:
        m = lie["t2l"]
        exp = np.zeros((x_zero.shape[0], ts), dtype=dtype)
        O.spmv(m["data"], m["indices"], m["offsets"], dim, exp, x_zero, csc=True, out_offset=1, out_size=ts - 1)
    else:
        m = lie["l2t"]
        exp = np.zeros((x_zero.shape[0], dim), dtype=dtype)
        O.spmv(m["data"], m["indices"], m["offsets"], ts - 1, exp, x_zero, csc=True, out_offset=0, out_size=dim,
               arg_offset=1)
    return exp


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(3, 4), (2, 6), (3, 8), (3, 9)])
@pytest.mark.parametrize("to_lie", [True, False], ids=["t2l", "l2t"])
def test_lie_conversions_on_windows(dev, shape, to_lie, dtype):
    """Both conversions on window pairs: only the output window is written,
    the argument is read only on its window (NaN elsewhere), degrees that the
    argument's window lacks come out 0.  The matrices are block diagonal by
    degree, so the expectation is the full product of the zero-extended
    argument restricted to the output window.  (3,9) in FP64 exercises the
    generic CSR path (its argument does not fit in LDS), in FP32 the packed
    row kernel at its largest shape."""
    w, N = shape
    basis = rpp_hip.TensorBasis(w, N)
    lie = rpp_hip.lie_basis(w, N)
    ts, dim = lie["true_size"], lie["tensor_dim"]
    tdb = O.degree_begin(w, N)
    ldb = [int(v) for v in lie["degree_begin"]]
    rng = np.random.default_rng(w * 17 + N)
    n = 3
    twins = [(0, N), (1, N), (2, N - 1), (0, 1), (N, N)]
    lwins = [(1, N), (2, N), (1, 2), (N, N), (2, 3)]
    pairs = [(twins[i], lwins[j]) for i in range(len(twins)) for j in range(len(lwins)) if (i + j) % 2 == 0]
    if shape == (3, 9):
        pairs = [((0, N), (1, N)), ((2, N - 1), (2, 3))]
    for tw, lw in pairs:
        if to_lie:
            x = damped_batch(rng, n, w, N, dtype)
            xz = x.copy()
            xz[:, :tdb[tw[0]]] = 0
            xz[:, tdb[tw[1] + 1]:] = 0
            xin = x.copy()
            xin[:, :tdb[tw[0]]] = np.nan
            xin[:, tdb[tw[1] + 1]:] = np.nan
            exp = _lie_expected(lie, xz, True, ts, dim, dtype)
            lo, hi = ldb[lw[0]], ldb[lw[1] + 1]
            out = to_dev(np.full((n, ts), 7.0, dtype=dtype), dev)
            rpp_hip.tensor_to_lie(rpp_hip.Batch(out, *lw), rpp_hip.Batch(to_dev(xin, dev), *tw), basis)
        else:
            y = rng.uniform(-1, 1, size=(n, ts)).astype(dtype)
            y[:, 0] = 0
            yz = y.copy()
            yz[:, :ldb[lw[0]]] = 0
            yz[:, ldb[lw[1] + 1]:] = 0
            yin = y.copy()
            yin[:, :ldb[lw[0]]] = np.nan
            yin[:, ldb[lw[1] + 1]:] = np.nan
            exp = _lie_expected(lie, yz, False, ts, dim, dtype)
            lo, hi = tdb[tw[0]], tdb[tw[1] + 1]
            out = to_dev(np.full((n, dim), 7.0, dtype=dtype), dev)
            rpp_hip.lie_to_tensor(rpp_hip.Batch(out, *tw), rpp_hip.Batch(to_dev(yin, dev), *lw), basis)
        got = out.cpu().numpy()
        assert np.all(got[:, :lo] == 7.0) and np.all(got[:, hi:] == 7.0), (tw, lw)
        what = f"{'t2l' if to_lie else 'l2t'} {shape} {tw} {lw}"
        assert_close(got[:, lo:hi], exp[:, lo:hi], dtype, what)
        assert np.array_equal(got[:, lo:hi], exp[:, lo:hi]), what + " not bit-identical"


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(3, 8, 4099), (5, 6, 1031)], ids=["w3d8", "w5d6"])
def test_lie_conversions_large_batch(dev, shape, dtype):
    """(3,8) over 4099 objects and (5,6) over 1031: more object groups than
    resident work-groups (each work-group takes several in turn) and a last
    group that is only partly filled, with one thread per row serving every
    staged object (2, 4 or 8 of them by shape and type); both conversions
    bit-identical to the oracle."""
    w, N, n = shape
    basis = rpp_hip.TensorBasis(w, N)
    lie = rpp_hip.lie_basis(w, N)
    ts, dim = lie["true_size"], lie["tensor_dim"]
    rng = np.random.default_rng(0x38 + w)
    x = rng.uniform(-1, 1, size=(n, dim)).astype(dtype)
    exp_lie = _lie_expected(lie, x, True, ts, dim, dtype)
    lie_out = to_dev(np.full((n, ts), 7.0, dtype=dtype), dev)
    rpp_hip.tensor_to_lie(lie_out, to_dev(x, dev), basis)
    got = lie_out.cpu().numpy()
    assert np.all(got[:, 0] == 7.0)
    assert np.array_equal(got[:, 1:], exp_lie[:, 1:])
    y = rng.uniform(-1, 1, size=(n, ts)).astype(dtype)
    y[:, 0] = np.nan  # slot 0 is never read
    exp_t = _lie_expected(lie, np.nan_to_num(y, nan=0.0), False, ts, dim, dtype)
    t_out = to_dev(np.full((n, dim), np.nan, dtype=dtype), dev)
    rpp_hip.lie_to_tensor(t_out, to_dev(y, dev), basis)
    assert np.array_equal(t_out.cpu().numpy(), exp_t)
