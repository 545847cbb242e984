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
