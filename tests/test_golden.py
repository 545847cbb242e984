"""Parity against the golden fixtures written from the real reference
(tests/golden/make_golden.py: the reference's own cpu::single_thread free
functions on seeded inputs).  The CPU tests pin the oracle to them; the GPU
tests hold the HIP path to them, so both run where the reference is absent."""
import os

import numpy as np
import pytest

import oracle as O
from helpers import assert_close, assert_exact, rel_err

HERE = os.path.dirname(os.path.abspath(__file__))
TOL = 1e-12


def fixture(name):
    with np.load(os.path.join(HERE, "golden", name + ".npz")) as z:
        return {k: z[k] for k in z.files}


def windows_of(name):
    return dict(ow=(1, 3), lw=(0, 2), rw=(1, 4), beta=0.5) if name == "mul_w4d4_windows" else {}


# ------------------------------------------------------------------- oracle

@pytest.mark.parametrize("name", ["mul_w4d4", "mul_w4d4_windows"])
def test_oracle_products(name):
    f = fixture(name)
    o = np.zeros_like(f["out"])
    O.orc.ft_mul(4, 4, o, f["lhs"], f["rhs"], **windows_of(name))
    assert rel_err(o, f["out"]) <= TOL


def test_oracle_fma_family():
    f = fixture("fma_w3d4")
    o = np.zeros_like(f["out"])
    O.orc.ft_fma(3, 4, o, f["a"], f["b"], f["c"], alpha=0.75, beta=1.25)
    assert rel_err(o, f["out"]) <= TOL
    for variant in (231, 123, 213):
        g = fixture(f"inplace_fma{variant}_w3d4")
        x = g["a"].copy()
        O.orc.ft_inplace_fma(variant, 3, 4, x, g["b"], g["c"], alpha=0.5, beta=2.0)
        assert rel_err(x, g["out"]) <= TOL, variant


@pytest.mark.parametrize("name", ["exp", "log", "antipode", "reflect"])
def test_oracle_unary(name):
    f = fixture(f"{name}_w3d4")
    o = np.zeros_like(f["out"])
    op = {"exp": O.orc.ft_exp, "log": O.orc.ft_log, "antipode": O.orc.tensor_antipode,
          "reflect": O.orc.tensor_reflect}[name]
    op(3, 4, o, f["arg"])
    assert rel_err(o, f["out"]) <= TOL


@pytest.mark.parametrize("side", ["lmul", "rmul"])
def test_oracle_adjoints(side):
    f = fixture(f"adj_{side}_w3d4")
    o = np.zeros_like(f["out"])
    getattr(O.orc, "ft_adj_" + side)(3, 4, o, f["op"], f["arg"])
    assert rel_err(o, f["out"]) <= TOL


def test_oracle_pairing_and_chains():
    f = fixture("pairing_w6d4")
    o = np.zeros(f["out"].shape)
    O.orc.tensor_pairing(6, 4, o, f["functional"], f["arg"])
    assert rel_err(o, f["out"]) <= TOL
    s = fixture("signature_w5d6")
    n, L, w = s["increments"].shape
    got = O.signature(5, 6, s["increments"].reshape(-1, w), np.arange(n + 1) * L)
    assert rel_err(got, s["out"]) <= TOL
    g = fixture("logsig_w3d8")
    n, L, w = g["increments"].shape
    sig = O.signature(3, 8, g["increments"].reshape(-1, w), np.arange(n + 1) * L)
    assert rel_err(sig, g["signature"]) <= TOL
    lg = np.zeros_like(sig)
    O.orc.ft_log(3, 8, lg, sig)
    assert rel_err(lg, g["log"]) <= TOL


def test_oracle_lie_conversions_exact():
    """tensor_to_lie and lie_to_tensor at (3,8): the oracle's CSC product with
    the library's Hall matrices (byte-identical to the reference builders')
    reproduces the reference's outputs bit for bit: both sum each row in
    ascending column order, one rounding per product and per addition."""
    import rpp_hip

    g = fixture("logsig_w3d8")
    lie = rpp_hip.lie_basis(3, 8)
    ts, dim = lie["true_size"], lie["tensor_dim"]
    m = lie["t2l"]
    got = np.zeros_like(g["lie"])
    O.spmv(m["data"], m["indices"], m["offsets"], dim, got, g["log"], csc=True, out_offset=1, out_size=ts - 1)
    assert np.array_equal(got[:, 1:], g["lie"][:, 1:])
    m = lie["l2t"]
    back = np.zeros_like(g["back"])
    O.spmv(m["data"], m["indices"], m["offsets"], ts - 1, back, g["lie"], csc=True, out_offset=0, out_size=dim,
           arg_offset=1)
    assert np.array_equal(back, g["back"])


LINEAR = [("tensor", 3, 4), ("lie", 3, 4)]


def linear_start(kind):
    return O.degree_begin(3, 4) if kind == "tensor" else O.lie_degree_begin(3, 4)


def linear_cases(f):
    """(name, initial, call) of each op in the linear fixtures; `call(impl,
    start, arrays)` runs it through an implementation with the oracle's
    argument order."""
    return [
        ("assign", np.full_like(f["y"], 7.0), lambda I, st, a: I.vector_assign(st, 4, a, f["y"], ow=(1, 3), aw=(2, 4))),
        ("add", np.full_like(f["x"], 7.0),
         lambda I, st, a: I.vector_add(st, 4, a, f["x"], f["y"], ow=(0, 3), lw=(1, 4), rw=(0, 2), alpha=0.75,
                                       beta=-1.5)),
        ("inplace", f["z"].copy(),
         lambda I, st, a: I.vector_inplace_add(st, 4, a, f["x"], lw=(1, 4), rw=(0, 2), alpha=-0.25)),
        ("scaled", f["z"].copy(), lambda I, st, a: I.vector_scalar_multiply(st, 4, a, 1.25, vw=(2, 4))),
        ("const", f["z"].copy(), lambda I, st, a: I.vector_set_constant(st, 4, a, -0.5, vw=(0, 1))),
        ("ident", f["z"].copy(), lambda I, st, a: I.tensor_set_identity(st, 4, a, 2.0, vw=(0, 2))),
        ("addid", f["z"].copy(), lambda I, st, a: I.tensor_add_identity(st, 4, a, 0.5)),
    ]


@pytest.mark.parametrize("kind", ["tensor", "lie"])
def test_oracle_linear_operations(kind):
    f = fixture(f"linear_{kind}_w3d4")
    start = linear_start(kind)
    for name, init, call in linear_cases(f):
        call(O.orc, start, init)
        assert np.array_equal(init, f[name]), name


SHUFFLE_CALLS = [
    ("st_mul", lambda I, o, f: I.st_mul(3, 4, o, f["x"], f["y"], ow=(0, 4), lw=(0, 3), rw=(1, 4), beta=-0.5)),
    ("st_fma", lambda I, o, f: I.st_fma(3, 4, o, f["z"], f["x"], f["y"], ow=(1, 4), aw=(0, 2), bw=(0, 4), cw=(0, 4),
                                        alpha=0.75, beta=1.25)),
    ("st_inplace_fma", lambda I, o, f: I.st_inplace_fma(3, 4, o, f["x"], f["y"], aw=(0, 3), bw=(1, 4), cw=(0, 4),
                                                        alpha=0.5, beta=2.0)),
    ("st_adj_mul", lambda I, o, f: I.st_adj_mul(3, 4, o, f["x"], f["y"], ow=(0, 4), pw=(0, 2), gw=(0, 4))),
    ("left_hs_mul", lambda I, o, f: I.left_hs_mul(3, 4, o, f["x"], f["y"], ow=(0, 4), xw=(1, 4), yw=(0, 3),
                                                  beta=1.5)),
    ("left_hs_adj_lmul", lambda I, o, f: I.left_hs_adj_lmul(3, 4, o, f["x"], f["y"], ow=(0, 3))),
    ("left_hs_adj_rmul", lambda I, o, f: I.left_hs_adj_rmul(3, 4, o, f["x"], f["y"], ow=(0, 3))),
]


def test_oracle_shuffle_family():
    f = fixture("shuffle_w3d4")
    for name, call in SHUFFLE_CALLS:
        o = f["z"].copy() if name == "st_inplace_fma" else np.full_like(f["x"], 7.0)
        call(O.orc, o, f)
        assert np.array_equal(o, f[name]), name


# ---------------------------------------------------------------------- GPU

@pytest.mark.gpu
def test_gpu_shuffle_family_against_golden(dev):
    import torch

    import rpp_hip

    f = fixture("shuffle_w3d4")
    B = rpp_hip.TensorBasis(3, 4)
    Bt = rpp_hip.Batch

    def t(x):
        return torch.tensor(np.ascontiguousarray(x), device=dev)

    x, y, z = t(f["x"]), t(f["y"]), t(f["z"])

    def check(name, o):
        torch.cuda.synchronize()
        assert_exact(o.cpu().numpy(), f[name], np.float64, name)

    o = t(np.full_like(f["x"], 7.0))
    rpp_hip.st_mul(Bt(o, 0, 4), Bt(x, 0, 3), Bt(y, 1, 4), B, beta=-0.5)
    check("st_mul", o)
    o = t(np.full_like(f["x"], 7.0))
    rpp_hip.st_fma(Bt(o, 1, 4), Bt(z, 0, 2), x, y, B, alpha=0.75, beta=1.25)
    check("st_fma", o)
    o = t(f["z"])
    rpp_hip.st_inplace_fma(Bt(o, 0, 3), Bt(x, 1, 4), y, B, alpha=0.5, beta=2.0)
    check("st_inplace_fma", o)
    o = t(np.full_like(f["x"], 7.0))
    rpp_hip.st_adj_mul(o, Bt(x, 0, 2), y, B)
    check("st_adj_mul", o)
    o = t(np.full_like(f["x"], 7.0))
    rpp_hip.left_hs_mul(o, Bt(x, 1, 4), Bt(y, 0, 3), B, beta=1.5)
    check("left_hs_mul", o)
    for name in ("left_hs_adj_lmul", "left_hs_adj_rmul"):
        o = t(np.full_like(f["x"], 7.0))
        getattr(rpp_hip, name)(Bt(o, 0, 3), x, y, B)
        check(name, o)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["tensor", "lie"])
def test_gpu_linear_operations_against_golden(dev, kind):
    import torch

    import rpp_hip

    f = fixture(f"linear_{kind}_w3d4")
    B = rpp_hip.TensorBasis(3, 4) if kind == "tensor" else rpp_hip.LieBasis(3, 4)

    def t(x):
        return torch.tensor(np.ascontiguousarray(x), device=dev)

    def run(name, init, fn):
        buf = t(init)
        fn(buf)
        torch.cuda.synchronize()
        assert np.array_equal(buf.cpu().numpy(), f[name]), name

    Bt = rpp_hip.Batch
    run("assign", np.full_like(f["y"], 7.0), lambda o: rpp_hip.vector_assign(Bt(o, 1, 3), Bt(t(f["y"]), 2, 4), B))
    run("add", np.full_like(f["x"], 7.0),
        lambda o: rpp_hip.vector_add(Bt(o, 0, 3), Bt(t(f["x"]), 1, 4), Bt(t(f["y"]), 0, 2), B, alpha=0.75,
                                     beta=-1.5))
    run("inplace", f["z"], lambda o: rpp_hip.vector_inplace_add(Bt(o, 1, 4), Bt(t(f["x"]), 0, 2), B, alpha=-0.25))
    run("scaled", f["z"], lambda o: rpp_hip.vector_scalar_multiply(Bt(o, 2, 4), B, scalar=1.25))
    run("const", f["z"], lambda o: rpp_hip.vector_set_constant(Bt(o, 0, 1), B, value=-0.5))
    run("ident", f["z"], lambda o: rpp_hip.tensor_set_identity(Bt(o, 0, 2), B, scalar=2.0))
    run("addid", f["z"], lambda o: rpp_hip.tensor_add_identity(o, B, scalar=0.5))


def c3_worst_signatures(f):
    """The reference's signatures of the fixture's paths, rebuilt by the
    oracle's chain (bit for bit: their SHA-256 is the fixture's)."""
    import hashlib
    inc = f["increments"]
    n, L, w = inc.shape
    sig = O.signature(w, 8, inc.reshape(-1, w), np.arange(n + 1, dtype=np.int64) * L)
    assert hashlib.sha256(np.ascontiguousarray(sig).tobytes()).digest() == f["signature_sha256"].tobytes(), \
        "the oracle's chain no longer reproduces the reference's signatures"
    return sig


def test_oracle_log_c3_worst():
    """The 32 worst-conditioned log-signatures of the width 3, depth 8
    workload (|log S| up to 30): the oracle's ft_log equals the reference's bit
    for bit."""
    f = fixture("log_c3_worst_w3d8")
    sig = c3_worst_signatures(f)
    got = np.zeros_like(sig)
    O.orc.ft_log(3, 8, got, sig)
    assert np.abs(f["log"]).max() > 20
    assert np.array_equal(got, f["log"]), f"oracle ft_log differs: {rel_err(got, f['log']):.3e}"


@pytest.mark.gpu
def test_gpu_log_c3_worst(dev):
    """The HIP ft_log (the LDS-resident kernel at width 3, depth 8) on the 32
    worst-conditioned reference signatures of the log-signature workload:
    within the reference's 1e-12 relative tolerance, and bit for bit."""
    import torch

    import rpp_hip
    f = fixture("log_c3_worst_w3d8")
    sig = c3_worst_signatures(f)
    out = torch.full(sig.shape, float("nan"), dtype=torch.float64, device=dev)
    rpp_hip.ft_log(out, torch.tensor(sig, device=dev), rpp_hip.TensorBasis(3, 8))
    got = out.cpu().numpy()
    err = rel_err(got, f["log"])
    assert err <= TOL, f"ft_log on the worst C3 paths: {err:.3e} > {TOL:.0e}"
    assert np.array_equal(got, f["log"]), f"ft_log on the worst C3 paths not bit-identical ({err:.3e})"


@pytest.mark.gpu
def test_gpu_against_golden(dev):
    import torch

    import rpp_hip

    def t(x):
        return torch.tensor(np.ascontiguousarray(x), device=dev)

    for name in ("mul_w4d4", "mul_w4d4_windows"):
        f = fixture(name)
        basis = rpp_hip.TensorBasis(4, 4)
        o = t(np.zeros_like(f["out"]))
        kw = windows_of(name)
        if kw:
            rpp_hip.ft_mul(rpp_hip.Batch(o, *kw["ow"]), rpp_hip.Batch(t(f["lhs"]), *kw["lw"]),
                           rpp_hip.Batch(t(f["rhs"]), *kw["rw"]), basis, beta=kw["beta"])
        else:
            rpp_hip.ft_mul(o, t(f["lhs"]), t(f["rhs"]), basis)
        assert_exact(o.cpu().numpy(), f["out"], np.float64, name)
    b34 = rpp_hip.TensorBasis(3, 4)
    f = fixture("fma_w3d4")
    o = t(np.zeros_like(f["out"]))
    rpp_hip.ft_fma(o, t(f["a"]), t(f["b"]), t(f["c"]), b34, alpha=0.75, beta=1.25)
    assert_exact(o.cpu().numpy(), f["out"], np.float64, "fma")
    for variant in (231, 123, 213):
        g = fixture(f"inplace_fma{variant}_w3d4")
        x = t(g["a"])
        rpp_hip.ft_inplace_fma(variant, x, t(g["b"]), t(g["c"]), b34, alpha=0.5, beta=2.0)
        assert_exact(x.cpu().numpy(), g["out"], np.float64, f"inplace_fma {variant}")
    for name, op in (("exp", rpp_hip.ft_exp), ("log", rpp_hip.ft_log), ("antipode", rpp_hip.tensor_antipode),
                     ("reflect", rpp_hip.tensor_reflect)):
        g = fixture(f"{name}_w3d4")
        o = t(np.zeros_like(g["out"]))
        op(o, t(g["arg"]), b34)
        assert_exact(o.cpu().numpy(), g["out"], np.float64, name)
    for side in ("lmul", "rmul"):
        g = fixture(f"adj_{side}_w3d4")
        o = t(np.zeros_like(g["out"]))
        getattr(rpp_hip, "ft_adj_" + side)(o, t(g["op"]), t(g["arg"]), b34)
        assert_exact(o.cpu().numpy(), g["out"], np.float64, "adj_" + side)
    g = fixture("pairing_w6d4")
    o = t(np.zeros(g["out"].shape))
    rpp_hip.tensor_pairing(o, t(g["functional"]), t(g["arg"]), rpp_hip.TensorBasis(6, 4))
    assert_close(o.cpu().numpy(), g["out"], np.float64, "pairing")
    # the plain chain and the fused kernel against the reference's chain
    s = fixture("signature_w5d6")
    n, L, w = s["increments"].shape
    b56 = rpp_hip.TensorBasis(5, 6)
    pad = np.zeros((n, 1 + L * w))
    pad[:, 1:] = s["increments"].reshape(n, L * w)
    pad_d = t(pad)
    cur = t(np.zeros((n, b56.size())))
    cur[:, 0] = 1
    nxt = torch.empty_like(cur)
    for k in range(L):
        rpp_hip.ft_fmexp(nxt, cur, rpp_hip.Batch(pad_d[:, k * w: k * w + 1 + w], 1, 1), b56)
        cur, nxt = nxt, cur
    # the plain chain is the reference's own ft_fmexp recursion: bit for bit
    assert_exact(cur.cpu().numpy(), s["out"], np.float64, "plain chain")
    fused = rpp_hip.signature(t(s["increments"].reshape(-1, w)), t(np.arange(n + 1) * L), b56)
    assert_close(fused.cpu().numpy(), s["out"], np.float64, "fused chain")
    # the log-signature pipeline
    g = fixture("logsig_w3d8")
    n, L, w = g["increments"].shape
    b38 = rpp_hip.TensorBasis(3, 8)
    # the plain chained ft_fmexp, then ft_log, tensor_to_lie and lie_to_tensor
    # on its output: the reference's whole log-signature pipeline, bit for bit
    pad = np.zeros((n, 1 + L * w))
    pad[:, 1:] = g["increments"].reshape(n, L * w)
    pad_d = t(pad)
    cur = t(np.zeros((n, b38.size())))
    cur[:, 0] = 1
    nxt = torch.empty_like(cur)
    for k in range(L):
        rpp_hip.ft_fmexp(nxt, cur, rpp_hip.Batch(pad_d[:, k * w: k * w + 1 + w], 1, 1), b38)
        cur, nxt = nxt, cur
    assert_exact(cur.cpu().numpy(), g["signature"], np.float64, "plain chain (3,8)")
    lg0 = torch.empty_like(cur)
    rpp_hip.ft_log(lg0, cur, b38)
    assert_exact(lg0.cpu().numpy(), g["log"], np.float64, "log of the plain chain (3,8)")
    lie0 = t(np.zeros(g["lie"].shape))
    rpp_hip.tensor_to_lie(lie0, lg0, b38)
    assert_exact(lie0.cpu().numpy()[:, 1:], g["lie"][:, 1:], np.float64, "tensor_to_lie of the plain chain (3,8)")
    # the fused signature sums in its own order: within the tolerance
    sig = rpp_hip.signature(t(g["increments"].reshape(-1, w)), t(np.arange(n + 1) * L), b38)
    assert_close(sig.cpu().numpy(), g["signature"], np.float64, "signature (3,8)")
    lg = torch.empty_like(sig)
    rpp_hip.ft_log(lg, sig, b38)
    assert_close(lg.cpu().numpy(), g["log"], np.float64, "log (3,8)")
    lie = t(np.zeros(g["lie"].shape))
    rpp_hip.tensor_to_lie(lie, lg, b38)
    assert_close(lie.cpu().numpy()[:, 1:], g["lie"][:, 1:], np.float64, "tensor_to_lie (3,8)")
    # on the reference's own log-signatures the conversions agree bit for bit
    rpp_hip.tensor_to_lie(lie, t(g["log"]), b38)
    assert np.array_equal(lie.cpu().numpy()[:, 1:], g["lie"][:, 1:]), "tensor_to_lie (3,8) not bit-identical"
    back = torch.empty_like(lg)
    rpp_hip.lie_to_tensor(back, t(g["lie"]), b38)
    assert_close(back.cpu().numpy(), g["back"], np.float64, "lie_to_tensor (3,8)")
    assert np.array_equal(back.cpu().numpy(), g["back"]), "lie_to_tensor (3,8) not bit-identical"
