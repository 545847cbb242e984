"""Pin the oracle against what the mathematics fixes, independently of the
reference: exact dyadic brute force of the concatenation product, unit and
associativity identities, closed forms of exp for a degree-one argument,
log(exp(x)) = x, fmexp(exp x, y) = exp(x) exp(y), the adjoint pairing
identities, Chen's identity and central finite differences of the signature
gradient (tests/rpp/cpu/single_thread/ops/ft_fmexp_tests.cpp:90-116,
ft_exp_log_tests.cpp:77-120, ft_adj_lmul_tests.cpp:244-259)."""
import itertools
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle as O
from helpers import degree_begin, rel_err


def words(w, k):
    return list(itertools.product(range(w), repeat=k))


def index_of(w, word):
    db = degree_begin(w, len(word))
    local = 0
    for letter in word:
        local = local * w + letter
    return db[len(word)] + local


@pytest.mark.parametrize("shape", [(2, 3), (3, 2), (1, 4), (4, 2)])
def test_ft_mul_exact_dyadic(shape):
    """Brute-force concatenation with Fractions; dyadic inputs make the
    double result exact."""
    w, N = shape
    rng = np.random.default_rng(0)
    dim = O.tensor_dimension(w, N)
    lhs = rng.integers(-8, 9, size=(1, dim)) / 8.0
    rhs = rng.integers(-8, 9, size=(1, dim)) / 16.0
    out = np.zeros((1, dim))
    O.orc.ft_mul(w, N, out, lhs, rhs)
    for k in range(N + 1):
        for word in words(w, k):
            acc = Fraction(0)
            for j in range(k + 1):
                acc += Fraction(lhs[0, index_of(w, word[:j])]) * Fraction(rhs[0, index_of(w, word[j:])])
            assert Fraction(out[0, index_of(w, word)]) == acc


def test_unit_and_associativity():
    w, N = 3, 4
    rng = np.random.default_rng(1)
    dim = O.tensor_dimension(w, N)
    a, b, c = (rng.uniform(-1, 1, size=(2, dim)) for _ in range(3))
    unit = np.zeros((2, dim)); unit[:, 0] = 1
    t = np.zeros_like(a)
    O.orc.ft_mul(w, N, t, unit, a)
    assert np.array_equal(t, a)
    O.orc.ft_mul(w, N, t, a, unit)
    assert np.array_equal(t, a)
    ab, abc1, bc, abc2 = (np.zeros_like(a) for _ in range(4))
    O.orc.ft_mul(w, N, ab, a, b)
    O.orc.ft_mul(w, N, abc1, ab, c)
    O.orc.ft_mul(w, N, bc, b, c)
    O.orc.ft_mul(w, N, abc2, a, bc)
    assert rel_err(abc1, abc2) < 1e-13


def test_mul_wrong_order_is_caught():
    """A plausible mistake (operands swapped) must fail the brute force."""
    w, N = 2, 3
    rng = np.random.default_rng(2)
    dim = O.tensor_dimension(w, N)
    lhs, rhs = rng.uniform(-1, 1, (1, dim)), rng.uniform(-1, 1, (1, dim))
    ab, ba = np.zeros((1, dim)), np.zeros((1, dim))
    O.orc.ft_mul(w, N, ab, lhs, rhs)
    O.orc.ft_mul(w, N, ba, rhs, lhs)
    assert rel_err(ab, ba) > 1e-3


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(2, 6), (3, 4), (5, 3)])
def test_exp_degree_one_closed_form(shape, dtype):
    w, N = shape
    rng = np.random.default_rng(3)
    dim = O.tensor_dimension(w, N)
    x = np.zeros((1, dim), dtype=dtype)
    v = rng.normal(size=w)
    x[0, 1:1 + w] = v
    out = np.zeros((1, dim), dtype=dtype)
    O.orc.ft_exp(w, N, out, x)
    for k in range(N + 1):
        for word in words(w, k):
            expected = np.prod([v[l] for l in word]) / math.factorial(k)
            tol = 1e-14 if dtype == np.float64 else 1e-6
            assert abs(out[0, index_of(w, word)] - expected) <= tol * max(1, abs(expected))


def test_log_exp_roundtrip_and_fmexp_identity():
    w, N = 3, 5
    rng = np.random.default_rng(4)
    dim = O.tensor_dimension(w, N)
    x = rng.uniform(-0.5, 0.5, (2, dim)); x[:, 0] = 0
    y = rng.uniform(-0.5, 0.5, (2, dim)); y[:, 0] = 0
    ex, lg, ey = np.zeros_like(x), np.zeros_like(x), np.zeros_like(x)
    O.orc.ft_exp(w, N, ex, x)
    O.orc.ft_log(w, N, lg, ex)
    assert rel_err(lg, x) < 1e-13
    O.orc.ft_exp(w, N, ey, y)
    prod, fm = np.zeros_like(x), np.zeros_like(x)
    O.orc.ft_mul(w, N, prod, ex, ey)
    O.orc.ft_fmexp(w, N, fm, ex, y)
    assert rel_err(fm, prod) < 1e-13
    unit = np.zeros_like(x); unit[:, 0] = 1
    O.orc.ft_fmexp(w, N, fm, unit, x)
    assert rel_err(fm, ex) < 1e-14
    zero = np.zeros_like(x)
    O.orc.ft_exp(w, N, fm, zero)
    assert np.array_equal(fm, unit)
    O.orc.ft_log(w, N, fm, unit)
    assert np.array_equal(fm, zero)


@pytest.mark.parametrize("side", ["lmul", "rmul"])
def test_adjoint_pairing_identity(side):
    """<L_a^* g, t> = <g, a t>  and  <R_a^* g, t> = <g, t a>."""
    w, N = 3, 4
    rng = np.random.default_rng(5)
    dim = O.tensor_dimension(w, N)
    a, g, t = (rng.uniform(-1, 1, (1, dim)) for _ in range(3))
    adj = np.zeros((1, dim))
    getattr(O.orc, "ft_adj_" + side)(w, N, adj, a, g)
    prod = np.zeros((1, dim))
    if side == "lmul":
        O.orc.ft_mul(w, N, prod, a, t)
    else:
        O.orc.ft_mul(w, N, prod, t, a)
    assert abs((adj @ t.T).item() - (g @ prod.T).item()) < 1e-12


def test_chen_identity_and_single_segment():
    w, N = 4, 4
    rng = np.random.default_rng(6)
    inc = rng.normal(0, 0.3, size=(7, w))
    s_all = O.signature(w, N, inc, np.array([0, 7]))
    s1 = O.signature(w, N, inc[:3], np.array([0, 3]))
    s2 = O.signature(w, N, inc[3:], np.array([0, 4]))
    prod = np.zeros((1, s1.shape[1]))
    O.orc.ft_mul(w, N, prod, s1, s2)
    assert rel_err(prod, s_all) < 1e-14
    # one segment: log of the signature is the increment in degree 1 only
    one = O.signature(w, N, inc[:1], np.array([0, 1]))
    lg = np.zeros_like(one)
    O.orc.ft_log(w, N, lg, one)
    expected = np.zeros_like(one)
    expected[0, 1:1 + w] = inc[0]
    assert rel_err(lg, expected) < 1e-14


def test_signature_vjp_finite_differences():
    w, N, L = 3, 4, 4
    rng = np.random.default_rng(7)
    inc = rng.normal(0, 0.4, size=(L, w))
    off = np.array([0, L])
    ell = rng.normal(size=O.tensor_dimension(w, N))
    g = O.signature_vjp(w, N, inc, off, ell)
    h = 1e-6
    fd = np.zeros_like(inc)
    for t in range(L):
        for a in range(w):
            p, m = inc.copy(), inc.copy()
            p[t, a] += h
            m[t, a] -= h
            fd[t, a] = (O.signature(w, N, p, off)[0] @ ell - O.signature(w, N, m, off)[0] @ ell) / (2 * h)
    assert rel_err(g, fd) < 1e-7


@pytest.mark.parametrize("shape", [(2, 5), (3, 4), (5, 3)])
def test_fmexp_vjp_pairing_and_finite_differences(shape):
    """Adjoint of one ft_fmexp step: out = S (x) exp(x) is linear in S, so
    <grad_S, V> = <G, fmexp(V, x)> for every V (degree 0 included); the x part
    against central differences."""
    w, N = shape
    dim = O.tensor_dimension(w, N)
    rng = np.random.default_rng(70 + w)
    n = 3
    S = rng.normal(size=(n, dim))
    x = rng.normal(0, 0.5, size=(n, w))
    G = rng.normal(size=(n, dim))
    gm, gx = O.fmexp_vjp(w, N, S, x, G)
    e = np.zeros((n, dim))
    e[:, 1:1 + w] = x

    def fmexp(mult, ex):
        out = np.zeros((n, dim))
        O.orc.ft_fmexp(w, N, out, mult, ex, ew=(1, 1))
        return out

    V = rng.normal(size=(n, dim))
    lhs = np.einsum("bi,bi->b", gm, V)
    rhs = np.einsum("bi,bi->b", G, fmexp(V, e))
    assert rel_err(lhs, rhs) < 1e-12
    h = 1e-5
    fd = np.zeros_like(x)
    for a in range(w):
        p, m = e.copy(), e.copy()
        p[:, 1 + a] += h
        m[:, 1 + a] -= h
        fd[:, a] = np.einsum("bi,bi->b", G, fmexp(S, p) - fmexp(S, m)) / (2 * h)
    assert rel_err(gx, fd) < 1e-7


@pytest.mark.parametrize("shape", [(2, 5), (3, 4), (4, 3)])
def test_log_vjp_finite_differences(shape):
    """Gradient of <G, log(X)> against central differences of the oracle's
    ft_log, and degree 0 of the argument (ignored by ft_log) gets 0."""
    w, N = shape
    dim = O.tensor_dimension(w, N)
    rng = np.random.default_rng(90 + w)
    n = 2
    X = rng.normal(0, 0.3, size=(n, dim))
    X[:, 0] = 1.0
    G = rng.normal(size=(n, dim))
    g = O.log_vjp(w, N, X, G)
    assert np.all(g[:, 0] == 0)

    def loss(Z):
        out = np.zeros((n, dim))
        O.orc.ft_log(w, N, out, Z)
        return np.einsum("bi,bi->b", G, out)

    h = 1e-6
    idx = rng.choice(np.arange(1, dim), size=min(dim - 1, 25), replace=False)
    fd = np.zeros((n, idx.size))
    for c, i in enumerate(idx):
        p, m = X.copy(), X.copy()
        p[:, i] += h
        m[:, i] -= h
        fd[:, c] = (loss(p) - loss(m)) / (2 * h)
    assert rel_err(g[:, idx], fd) < 1e-7


@pytest.mark.parametrize("shape", [(2, 5), (3, 4)])
def test_antipode_identities(shape):
    """The antipode reverses products, alpha(a b) = alpha(b) alpha(a), maps
    exp(x) to exp(-x) for a degree-one x, and reflection is an involution."""
    w, N = shape
    dim = O.tensor_dimension(w, N)
    rng = np.random.default_rng(40 + w)
    a = rng.normal(size=(2, dim))
    b = rng.normal(size=(2, dim))

    def mul(x, y):
        out = np.zeros_like(x)
        O.orc.ft_mul(w, N, out, x, y)
        return out

    def anti(x):
        out = np.zeros_like(x)
        O.orc.tensor_antipode(w, N, out, x)
        return out

    assert rel_err(anti(mul(a, b)), mul(anti(b), anti(a))) < 1e-13
    x = np.zeros((2, dim))
    x[:, 1:1 + w] = rng.normal(size=(2, w))
    ex = np.zeros_like(x)
    O.orc.ft_exp(w, N, ex, x)
    emx = np.zeros_like(x)
    O.orc.ft_exp(w, N, emx, -x)
    assert rel_err(anti(ex), emx) < 1e-14
    r1 = np.zeros_like(a)
    O.orc.tensor_reflect(w, N, r1, a)
    r2 = np.zeros_like(a)
    O.orc.tensor_reflect(w, N, r2, r1)
    assert np.array_equal(r2, a)
    # a non-palindromic word moves: reflect is not the identity
    assert not np.array_equal(r1, a)
