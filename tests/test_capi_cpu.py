"""The C-ABI library loads and exports every symbol include/rpp_hip.h
declares; host-side validation rejects bad descriptors before any HIP call
(so these run without a GPU)."""
import ctypes as C
import os
import re

import rpp_hip
from rpp_hip import _capi as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "rpp_hip.h")).read()
    return sorted(set(re.findall(r"\b(rpp_hip_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = rpp_hip.lib()
    names = declared_symbols()
    assert len(names) >= 35
    for name in names:
        assert hasattr(lib, name), name
    assert set(A.exported_symbols()) <= set(names)


def test_version_string():
    assert b"gfx950" in rpp_hip.lib().rpp_hip_version()


def _call(name, desc):
    lib = rpp_hip.lib()
    return getattr(lib, name)(C.byref(desc), None)


def test_invalid_width_depth_and_windows_rejected_on_host():
    tb = A.TensorBatch(None, 0, 1, 0, 2)
    d = A.FtMulDesc(A.Basis(0, 2), 1, tb, tb, tb, 1.0)
    assert _call("rpp_hip_ft_mul_f64", d) == 3
    d = A.FtMulDesc(A.Basis(3, 2), 1, A.TensorBatch(None, 10, 1, 0, 3), tb, tb, 1.0)
    assert _call("rpp_hip_ft_mul_f64", d) == 3
    assert b"window" in rpp_hip.lib().rpp_hip_last_error()
    # null data with a non-empty window
    d = A.FtMulDesc(A.Basis(3, 2), 1, A.TensorBatch(None, 13, 1, 0, 2), tb, tb, 1.0)
    assert _call("rpp_hip_ft_mul_f64", d) == 3
    # level size beyond 2^31
    d = A.FtMulDesc(A.Basis(255, 5), 1, tb, tb, tb, 1.0)
    assert _call("rpp_hip_ft_mul_f64", d) == 7


def test_fmexp_alias_and_variant_rejected():
    buf = C.c_double * 64
    a = buf()
    p = C.cast(a, C.c_void_p).value
    t = A.TensorBatch(p, 13, 1, 0, 2)
    d = A.FtFmexpDesc(A.Basis(3, 2), 1, t, t, A.TensorBatch(p + 8 * 20, 13, 1, 1, 1))
    assert _call("rpp_hip_ft_fmexp_f64", d) == 3
    d = A.FtInplaceFmaDesc(A.Basis(3, 2), 1, 999, t, t, t, 1.0, 1.0)
    assert _call("rpp_hip_ft_inplace_fma_f64", d) == 3


def test_signature_rejects_partial_windows():
    t = A.TensorBatch(8, 13, 1, 1, 2)
    d = A.SignatureDesc(A.Basis(3, 2), 2, 8, 8, t, 0, None, None, 0, 0)
    assert _call("rpp_hip_signature_f64", d) == 3


def test_fmexp_adjoint_validation_on_host():
    """Windows and aliasing of the ft_fmexp adjoint are checked before any HIP
    call (fake non-null pointers never reach a kernel)."""
    full = A.TensorBatch(0x1000, 31, 1, 0, 2)
    other = A.TensorBatch(0x9000, 31, 1, 0, 2)
    cot = A.TensorBatch(0x5000, 31, 1, 0, 2)
    x = A.TensorBatch(0x2000, 6, 1, 1, 1)
    gx = A.TensorBatch(0x3000, 6, 1, 1, 1)
    basis = A.Basis(5, 2)
    bad_x = A.TensorBatch(0x2000, 6, 1, 1, 2)
    assert _call("rpp_hip_ft_fmexp_adjoint_f64", A.FtFmexpAdjointDesc(basis, 2, full, bad_x, cot, other, gx)) == 3
    partial = A.TensorBatch(0x9000, 31, 1, 1, 2)
    assert _call("rpp_hip_ft_fmexp_adjoint_f64", A.FtFmexpAdjointDesc(basis, 2, full, x, cot, partial, gx)) == 3
    assert _call("rpp_hip_ft_fmexp_adjoint_f64", A.FtFmexpAdjointDesc(basis, 2, full, x, cot, cot, gx)) == 3
    assert _call("rpp_hip_ft_fmexp_adjoint_f64", A.FtFmexpAdjointDesc(A.Basis(5, 0), 2, full, x, cot, other,
                                                                      gx)) == 3
    assert _call("rpp_hip_ft_fmexp_adjoint_f64", A.FtFmexpAdjointDesc(basis, 0, full, x, cot, other, gx)) == 0


def test_checkpoint_policy():
    """The library's checkpoint spacing for a backward pass: none in FP64,
    none for FP32 paths under 32 increments, else ceil(sqrt(L)); bad
    arguments give a negative error code."""
    import rpp_hip
    lib = rpp_hip.lib()
    assert [lib.rpp_hip_signature_checkpoint_every(4, L) for L in (0, 1, 31, 32, 33, 256, 1000, 1024, 1025)] == \
        [0, 0, 0, 6, 6, 16, 32, 32, 33]
    assert lib.rpp_hip_signature_checkpoint_every(8, 4096) == 0
    assert lib.rpp_hip_signature_checkpoint_every(2, 64) == -3
    assert lib.rpp_hip_signature_checkpoint_every(4, -1) == -3


def _shared(name, desc):
    """The call is refused on the host as sharing memory (no HIP call)."""
    rc = _call(name, desc)
    msg = rpp_hip.lib().rpp_hip_last_error()
    assert rc == 3 and b"share memory" in msg, (name, rc, msg)


def test_partial_overlaps_rejected_on_host():
    """Views whose windows share elements without being the same batch are
    refused before any launch: a base pointer shifted by one object, one
    coefficient, or a column slice that reaches into the other's window.
    Fake pointers: a refusal never dereferences them."""
    B = A.Basis(3, 2)  # dim 13
    base = 0x100000
    n = 4
    om = lambda off, mn=0, mx=2, bs=13: A.TensorBatch(base + 8 * off, bs, 1, mn, mx)
    il = lambda off, mn=0, mx=2: A.TensorBatch(base + 8 * off, 1, n, mn, mx)
    # unary ops: out one object past arg (object-major and interleaved)
    for name in ("rpp_hip_ft_log_f64", "rpp_hip_ft_exp_f64", "rpp_hip_tensor_antipode_f64",
                 "rpp_hip_tensor_reflect_f64"):
        _shared(name, A.FtUnaryDesc(B, n, om(13), om(0)))
        _shared(name, A.FtUnaryDesc(B, n, il(1), il(0)))
    # column slices of a [n, 26] buffer: arg columns [0, 13), out columns
    # [12, 25) reach into arg's last coefficient
    _shared("rpp_hip_ft_log_f64", A.FtUnaryDesc(B, n, om(12, bs=26), om(0, bs=26)))
    # out's window degree 2 (coefficients 4..12) over arg's degrees 0-1 of the
    # next object: out = arg + 9 elements, bs 13
    _shared("rpp_hip_ft_exp_f64", A.FtUnaryDesc(B, n, om(9, 2, 2), om(0, 0, 1)))
    # products: the same data with other strides, or a shifted view
    _shared("rpp_hip_ft_mul_f64", A.FtMulDesc(B, n, om(13), om(0), om(100), 1.0))
    _shared("rpp_hip_ft_mul_f64", A.FtMulDesc(B, n, om(0, bs=14), om(200), om(0), 1.0))
    _shared("rpp_hip_ft_fma_f64", A.FtFmaDesc(B, n, om(0), om(1), om(100), om(200), 1.0, 1.0))
    _shared("rpp_hip_ft_inplace_mul_f64", A.FtInplaceMulDesc(B, n, om(0), il(5), 1.0))
    # fmexp: out over the exponent's degree-one slots of another object
    _shared("rpp_hip_ft_fmexp_f64", A.FtFmexpDesc(B, n, om(0), om(500), om(12, 1, 1)))
    # adjoints, the log adjoint and the fmexp adjoint
    _shared("rpp_hip_ft_adj_lmul_f64", A.AdjointDesc(B, n, om(0), om(13), om(500)))
    _shared("rpp_hip_ft_log_adjoint_f64", A.FtLogAdjointDesc(B, n, om(0), om(500), om(7)))
    _shared("rpp_hip_ft_fmexp_adjoint_f64",
            A.FtFmexpAdjointDesc(B, n, om(0), om(100, 1, 1), om(200), om(300), om(113, 1, 1)))
    # shuffle products: the output over a factor; st_fma's addend shifted
    _shared("rpp_hip_st_mul_f64", A.FtMulDesc(B, n, om(26), om(0), om(200), 1.0))
    _shared("rpp_hip_st_fma_f64", A.FtFmaDesc(B, n, om(1), om(0), om(200), om(300), 1.0, 1.0))
    # pairing: the scalar outputs land inside arg
    _shared("rpp_hip_tensor_pairing_f64", A.PairingDesc(B, n, A.ScalarBatch(base + 8 * 20, 1), om(200), om(0)))
    # vector ops over a graded tensor basis
    G = A.GradedBasis(0, 3, 2)
    _shared("rpp_hip_vector_assign_f64", A.VectorAssignDesc(G, n, om(2), om(0)))
    _shared("rpp_hip_vector_add_f64", A.VectorAddDesc(G, n, il(0), om(200), om(0), 1.0, 1.0))
    _shared("rpp_hip_vector_inplace_add_f64", A.VectorInplaceAddDesc(G, n, om(0), om(13), 1.0))
    # sparse product: out rows inside arg
    M = A.SparseMatrix(0, base + 0x10000, base + 0x20000, base + 0x30000, 1, 2, 2)
    _shared("rpp_hip_sparse_matrix_vector_product_f64",
            A.SpmvDesc(n, A.VectorBatch(base + 8, 13, 1, 0, 2), A.VectorBatch(base, 13, 1, 0, 2), M, 1.0))
    # Lie conversions: the Lie batch inside the tensor batch
    _shared("rpp_hip_tensor_to_lie_f64", A.LieConvertDesc(B, n, om(0), A.TensorBatch(base + 8 * 14, 6, 1, 1, 2)))


def test_alias_predicate_matches_brute_force(tmp_path):
    """The C-ABI's aliasing test (csrc/alias.hpp) against enumeration of the
    byte sets of 200000 random view pairs: it never misses a shared byte and
    is exact for equal-stride object-major and interleaved views."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        import pytest
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path / "alias_check")
    src = os.path.join(ROOT, "tests", "native", "alias_check.cpp")
    inc = os.path.join(ROOT, "rough-path-primitives_amd", "csrc")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I" + inc, "-o", exe, src], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "ALIAS OK" in r.stdout, r.stdout + r.stderr


def test_oversized_shapes_and_views_refused_on_host():
    """width^depth is bounded without wrapping (256^8 = 2^64 wraps to 0 in
    64-bit arithmetic), view extents must fit below 2^59 elements, and a
    pairing output cannot be broadcast."""
    tb = A.TensorBatch(0x100000, 13, 1, 0, 2)
    for w, d in ((256, 8), (65535, 30), (2, 31), (65536, 2)):
        rc = _call("rpp_hip_ft_mul_f64", A.FtMulDesc(A.Basis(w, d), 1, tb, tb, tb, 1.0))
        assert rc in (3, 7), (w, d, rc)
    assert _call("rpp_hip_ft_mul_f64", A.FtMulDesc(A.Basis(256, 8), 1, tb, tb, tb, 1.0)) == 7
    huge = A.TensorBatch(0x100000, 1 << 40, 1, 0, 2)
    rc = _call("rpp_hip_ft_log_f64", A.FtUnaryDesc(A.Basis(3, 2), 1 << 20, huge, A.TensorBatch(0x200000, 13, 1, 0, 2)))
    assert rc == 7 and b"2^59" in rpp_hip.lib().rpp_hip_last_error()
    v = A.VectorBatch(0x100000, 1 << 62, 1, 0, 4)
    M = A.SparseMatrix(0, 0x10, 0x20, 0x30, 0, 4, 4)
    assert _call("rpp_hip_sparse_matrix_vector_product_f64", A.SpmvDesc(4, v, A.VectorBatch(0x900000, 4, 1, 0, 4),
                                                                         M, 1.0)) == 7
    d = A.PairingDesc(A.Basis(3, 2), 4, A.ScalarBatch(0x100000, 0), tb, A.TensorBatch(0x200000, 13, 1, 0, 2))
    assert _call("rpp_hip_tensor_pairing_f64", d) == 3


def test_checkpoint_policy_bounds():
    lib = rpp_hip.lib()
    assert lib.rpp_hip_signature_checkpoint_every(4, 1 << 60) == 1 << 30
    assert lib.rpp_hip_signature_checkpoint_every(4, (1 << 60) + 1) < 0
    assert lib.rpp_hip_signature_checkpoint_every(4, 1025) == 33
    assert lib.rpp_hip_signature_checkpoint_every(4, 1024) == 32
