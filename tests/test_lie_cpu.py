"""Hall basis and Lie matrices of librpp_hip (host code, no GPU needed):
known answers from the reference's own tests, the reference builders
(oracle/_ref) where built, and the identities the mathematics fixes."""
import itertools

import numpy as np
import pytest

import oracle as O
import rpp_hip
from helpers import degree_begin


def csc_dense(m, rows, cols):
    d = np.zeros((rows, cols))
    for c in range(cols):
        for e in range(m["offsets"][c], m["offsets"][c + 1]):
            d[m["indices"][e], c] += m["data"][e]
    return d


def test_hall_known_answers_w3_d4():
    """basis/lie_basis_tests.cpp:15-42."""
    lb = rpp_hip.lie_basis(3, 4)
    assert lb["true_size"] == 33
    assert list(lb["degree_begin"]) == [0, 1, 4, 7, 15, 33]
    pairs = [tuple(p) for p in lb["pairs"]]
    index = {p: i for i, p in enumerate(pairs)}
    assert index[(1, 2)] == 4 and index[(2, 5)] == 10 and index[(3, 4)] == 12


def witt(w, k):
    from sympy import divisors, mobius
    return sum(mobius(d) * w ** (k // d) for d in divisors(k)) // k


@pytest.mark.parametrize("shape", [(2, 6), (3, 5), (4, 4), (3, 8)])
def test_witt_dimensions(shape):
    w, N = shape
    lb = rpp_hip.lie_basis(w, N)
    db = lb["degree_begin"]
    for k in range(1, N + 1):
        assert db[k + 1] - db[k] == witt(w, k)


def tensor_expansion(pairs, w, N):
    """Independent dense expansion of every Hall element: letters -> e_l,
    [a, b] -> a b - b a (concatenation in numpy)."""
    db = degree_begin(w, N)
    dim = db[-1]
    deg = {0: 0}
    exp = {}
    for key in range(1, len(pairs)):
        l, r = pairs[key]
        v = np.zeros(dim)
        if l == 0:
            v[db[1] + r - 1] = 1.0
            deg[key] = 1
        else:
            da, dbb = deg[l], deg[r]
            deg[key] = da + dbb
            A = exp[l][db[da]:db[da + 1]]
            B = exp[r][db[dbb]:db[dbb + 1]]
            k = da + dbb
            v[db[k]:db[k + 1]] = np.outer(A, B).ravel() - np.outer(B, A).ravel()
        exp[key] = v
    return exp


@pytest.mark.parametrize("shape", [(2, 5), (3, 4), (4, 3), (3, 6)])
def test_l2t_columns_are_commutator_expansions(shape):
    w, N = shape
    lb = rpp_hip.lie_basis(w, N)
    ts, dim = lb["true_size"], lb["tensor_dim"]
    L2T = csc_dense(lb["l2t"], dim, ts - 1)
    exp = tensor_expansion([tuple(p) for p in lb["pairs"]], w, N)
    for key in range(1, ts):
        assert np.array_equal(L2T[:, key - 1], exp[key])


@pytest.mark.parametrize("shape", [(2, 4), (3, 4), (2, 6), (4, 3), (3, 5)])
def test_t2l_inverts_l2t(shape):
    """T2L . L2T = I (sparse/lie_tensor_matrix_tests.cpp:104-186), and T2L
    columns are the right bracketings divided by the degree."""
    w, N = shape
    lb = rpp_hip.lie_basis(w, N)
    ts, dim = lb["true_size"], lb["tensor_dim"]
    L2T = csc_dense(lb["l2t"], dim, ts - 1)
    T2L = csc_dense(lb["t2l"], ts - 1, dim)
    assert np.allclose(T2L @ L2T, np.eye(ts - 1), atol=1e-13)
    # right bracketing of a degree-2 word xy is [x, y]: (1/2) * key([x,y])
    db = degree_begin(w, N)
    pairs = [tuple(p) for p in lb["pairs"]]
    for x, y in itertools.product(range(w), repeat=2):
        col = T2L[:, db[2] + x * w + y]
        if x == y:
            assert np.all(col == 0)
        elif x < y:
            key = pairs.index((x + 1, y + 1))
            assert col[key - 1] == 0.5 and np.count_nonzero(col) == 1


@pytest.mark.skipif(not O.has_ref(), reason="reference driver not built here")
@pytest.mark.parametrize("shape", [(2, 2), (3, 4), (2, 7), (4, 5), (3, 8), (5, 4), (8, 3)])
def test_matrices_identical_to_reference(shape):
    w, N = shape
    mine = rpp_hip.lie_basis(w, N)
    ref = O.ref_lie(w, N)
    assert np.array_equal(mine["pairs"], ref["pairs"])
    assert np.array_equal(mine["degree_begin"], ref["degree_begin"])
    for m in ("l2t", "t2l"):
        for k in ("data", "indices", "offsets"):
            assert np.array_equal(mine[m][k], ref[m][k]), (m, k)


@pytest.mark.skipif(not O.has_ref(), reason="reference driver not built here")
def test_bracket_cache_known_answers():
    """basis/lie_multiplication_cache_tests.cpp:30-86 (through the reference
    cache, pinning the keys our T2L relies on)."""
    assert O.ref_bracket(3, 4, 1, 6) == {10: 1, 12: -1}
    assert O.ref_bracket(3, 4, 1, 9) == {17: 1}
    assert O.ref_bracket(3, 4, 1, 13) == {23: 1}
    assert O.ref_bracket(3, 4, 1, 11) == {20: 1, 24: -1, 31: 2}
    assert O.ref_bracket(3, 4, 1, 14) == {25: 1, 27: -1, 32: 1}
