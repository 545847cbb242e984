"""The linear operations (vector_assign, vector_add, vector_inplace_add,
vector_scalar_multiply, vector_set_constant, tensor_set_identity,
tensor_add_identity) on the CPU:

* the oracle against the real reference (oracle/_ref) over every pair of
  degree windows, on tensor gradings of the reference's bound tests
  (tests/rpp/cpu/single_thread/ops/vector_bounds_tests.cpp:24-27) and on Lie
  gradings, with NaN outside the input windows (an accidental read shows)
  and guard values in the output (an accidental write shows), bit for bit;
* the Lie degree offsets (Witt's formula) of the oracle and of the
  library's Python API against the reference's own Hall-basis builder;
* the C-ABI's host-side validation (no HIP call is reached).
"""
import ctypes as C
import itertools

import numpy as np
import pytest

import oracle as O
import rpp_hip
from rpp_hip import _capi as A

needs_ref = pytest.mark.skipif(not O.has_ref(), reason="reference driver (oracle/_ref) not built here")

DTYPES = [np.float64, np.float32]
# (kind, width, depth): the reference's bound-test shapes plus Lie gradings
GRADINGS = [("tensor", 1, 4), ("tensor", 4, 4), ("tensor", 4, 0), ("tensor", 127, 1), ("tensor", 3, 3),
            ("lie", 3, 4), ("lie", 2, 5), ("lie", 5, 2)]


def offsets(kind, width, depth):
    return O.degree_begin(width, depth) if kind == "tensor" else O.lie_degree_begin(width, depth)


def windows(depth):
    return [(a, b) for a in range(depth + 1) for b in range(a, depth + 1)]


def storage(start, n, dtype, seed):
    """Exactly representable values, as the reference's bound tests use
    ((i % 5 - 2) / 4, vector_bounds_test_helper.hpp:12-21), per object."""
    dim = start[-1]
    i = np.arange(dim)
    base = ((i % 5) - 2) / 4.0
    return np.stack([np.roll(base, seed + 3 * b) for b in range(n)]).astype(dtype)


def with_nan_outside(x, start, win):
    y = x.copy()
    lo, hi = start[win[0]], start[win[1] + 1]
    y[:, :lo] = np.nan
    y[:, hi:] = np.nan
    return y


def same(a, b):
    return np.array_equal(a, b, equal_nan=True)


@needs_ref
@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("grading", GRADINGS, ids=lambda g: f"{g[0]}{g[1]}x{g[2]}")
def test_assign_and_inplace_add_every_window_pair(grading, dtype):
    start = offsets(*grading)
    N = grading[2]
    n = 2
    for ow, aw in itertools.product(windows(N), windows(N)):
        arg = with_nan_outside(storage(start, n, dtype, 1), start, aw)
        init = np.full((n, start[-1]), 7.0, dtype=dtype)
        a, b = init.copy(), init.copy()
        O.orc.vector_assign(start, N, a, arg, ow=ow, aw=aw)
        O.refimpl.vector_assign(start, N, b, arg, ow=ow, aw=aw)
        assert same(a, b), ("assign", ow, aw)
        lhs0 = storage(start, n, dtype, 2)
        for alpha in (0.0, 0.5, -1.0):
            a, b = lhs0.copy(), lhs0.copy()
            O.orc.vector_inplace_add(start, N, a, arg, lw=ow, rw=aw, alpha=alpha)
            O.refimpl.vector_inplace_add(start, N, b, arg, lw=ow, rw=aw, alpha=alpha)
            assert same(a, b), ("inplace_add", ow, aw, alpha)


@needs_ref
@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("grading", [g for g in GRADINGS if g[2] <= 3], ids=lambda g: f"{g[0]}{g[1]}x{g[2]}")
def test_vector_add_every_window_triple(grading, dtype):
    start = offsets(*grading)
    N = grading[2]
    n = 2
    for ow, lw, rw in itertools.product(windows(N), windows(N), windows(N)):
        lhs = with_nan_outside(storage(start, n, dtype, 3), start, lw)
        rhs = with_nan_outside(storage(start, n, dtype, 4), start, rw)
        init = np.full((n, start[-1]), -7.0, dtype=dtype)
        a, b = init.copy(), init.copy()
        O.orc.vector_add(start, N, a, lhs, rhs, ow=ow, lw=lw, rw=rw, alpha=0.5, beta=-1.0)
        O.refimpl.vector_add(start, N, b, lhs, rhs, ow=ow, lw=lw, rw=rw, alpha=0.5, beta=-1.0)
        assert same(a, b), (ow, lw, rw)


@needs_ref
@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("grading", GRADINGS, ids=lambda g: f"{g[0]}{g[1]}x{g[2]}")
def test_scalar_operations_every_window(grading, dtype):
    start = offsets(*grading)
    N = grading[2]
    ops = [("vector_scalar_multiply", (0.0, 0.5, -1.0)), ("vector_set_constant", (0.0, 0.5, -1.0)),
           ("tensor_set_identity", (1.0, -2.5)), ("tensor_add_identity", (1.0, 0.25))]
    for vw in windows(N):
        for name, scalars in ops:
            for s in scalars:
                v0 = storage(start, 3, dtype, 5)
                if name in ("vector_set_constant", "tensor_set_identity"):
                    # every overwritten coefficient starts as NaN
                    lo, hi = start[vw[0]], start[vw[1] + 1]
                    v0[:, lo:hi] = np.nan
                a, b = v0.copy(), v0.copy()
                getattr(O.orc, name)(start, N, a, s, vw=vw)
                getattr(O.refimpl, name)(start, N, b, s, vw=vw)
                assert same(a, b), (name, vw, s)


# vector_add rounds 0 + alpha*lhs, then + beta*rhs, with a rounding after each
# product (cpu/.../linalg/vector_add.hpp:52-60).  Coefficient 0: alpha*lhs =
# -(1 + 2^-29) exactly, beta*rhs = (1 + 2^-30)^2 = 1 + 2^-29 + 2^-60 rounds to
# 1 + 2^-29, so the sum is 0; a fused multiply-add would keep 2^-60.
# Coefficient 1: alpha*lhs = -0 gives 0 + (-0) = +0.
ROUNDING_CASE = dict(lhs=np.array([[-(1 + 2.0 ** -29), -0.0, 3.0]]),
                     rhs=np.array([[1 + 2.0 ** -30, 0.0, 0.0]]),
                     alpha=1.0, beta=1 + 2.0 ** -30)


def check_rounding_case(out):
    assert out[0, 0] == 0.0, out[0, 0]
    assert out[0, 1] == 0.0 and not np.signbit(out[0, 1])
    assert out[0, 2] == 3.0


def test_vector_add_rounds_each_product_as_the_reference():
    start = O.degree_begin(2, 1)
    c = ROUNDING_CASE
    out = np.zeros((1, 3))
    O.orc.vector_add(start, 1, out, c["lhs"], c["rhs"], alpha=c["alpha"], beta=c["beta"])
    check_rounding_case(out)
    if O.has_ref():
        ref = np.zeros((1, 3))
        O.refimpl.vector_add(start, 1, ref, c["lhs"], c["rhs"], alpha=c["alpha"], beta=c["beta"])
        assert same(out, ref) and np.array_equal(np.signbit(out), np.signbit(ref))


@needs_ref
@pytest.mark.parametrize("shape", [(2, 2), (3, 4), (2, 7), (4, 5), (3, 8), (5, 4), (8, 3), (1, 3), (6, 1)])
def test_lie_degree_offsets_match_the_reference_hall_basis(shape):
    w, N = shape
    ref = [int(v) for v in O.ref_lie(w, N)["degree_begin"]]
    assert O.lie_degree_begin(w, N) == ref
    lb = rpp_hip.LieBasis(w, N)
    assert [lb.start_of_degree(k) for k in range(N + 2)] == ref


def test_lie_degree_offsets_known_answers():
    # Hall w3 d4: degree ranges [0,1),[1,4),[4,7),[7,15),[15,33)
    # (tests/rpp/basis/lie_basis_tests.cpp:15-32)
    assert O.lie_degree_begin(3, 4) == [0, 1, 4, 7, 15, 33]
    assert rpp_hip.LieBasis(3, 4).true_size() == 33
    assert rpp_hip.LieBasis(3, 4).size() == 32
    # the library's own Hall builder agrees at (3,8)
    assert list(rpp_hip.lie_basis(3, 8)["degree_begin"]) == O.lie_degree_begin(3, 8)


# ------------------------------------------------------- C-ABI validation

def _call(name, desc):
    return getattr(rpp_hip.lib(), name)(C.byref(desc), None)


def test_vector_entry_points_validate_on_host():
    g = A.GradedBasis(A.TENSOR_BASIS, 3, 2)
    buf = (C.c_double * 64)()
    p = C.cast(buf, C.c_void_p).value
    om = A.TensorBatch(p, 13, 1, 0, 2)
    il = A.TensorBatch(p, 1, 4, 0, 2)
    # unknown basis kind
    d = A.VectorAssignDesc(A.GradedBasis(7, 3, 2), 2, om, om)
    assert _call("rpp_hip_vector_assign_f64", d) == 3
    # window beyond depth
    d = A.VectorAssignDesc(g, 2, A.TensorBatch(p, 13, 1, 0, 3), om)
    assert _call("rpp_hip_vector_assign_f64", d) == 3
    # transposing in place is not supported
    d = A.VectorAssignDesc(g, 4, om, il)
    assert _call("rpp_hip_vector_assign_f64", d) == 3
    assert b"share memory" in rpp_hip.lib().rpp_hip_last_error()
    # negative count
    d = A.VectorAddDesc(g, -1, om, om, om, 1.0, 1.0)
    assert _call("rpp_hip_vector_add_f64", d) == 3
    # a broadcast output
    d = A.VectorScalarDesc(g, 2, A.TensorBatch(p, 0, 1, 0, 2), 2.0)
    assert _call("rpp_hip_vector_set_constant_f64", d) == 3
    d = A.VectorInplaceAddDesc(A.GradedBasis(A.LIE_BASIS, 0, 2), 1, om, om, 1.0)
    assert _call("rpp_hip_vector_inplace_add_f64", d) == 3
    # empty work returns success without touching the device
    d = A.VectorScalarDesc(g, 0, om, 2.0)
    assert _call("rpp_hip_vector_scalar_multiply_f64", d) == 0
    d = A.VectorScalarDesc(g, 3, A.TensorBatch(p, 13, 1, 2, 1), 2.0)
    assert _call("rpp_hip_tensor_set_identity_f64", d) == 0
