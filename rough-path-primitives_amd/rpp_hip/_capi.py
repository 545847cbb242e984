"""ctypes declarations of librpp_hip.so (include/rpp_hip.h).

The library is the product path: loading fails loudly when it is missing, and
nothing here falls back to a CPU implementation.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librpp_hip.so")

_i32 = C.c_int32
_i64 = C.c_int64
_f64 = C.c_double
_vp = C.c_void_p


class Basis(C.Structure):
    _fields_ = [("width", _i32), ("depth", _i32)]


class TensorBatch(C.Structure):
    _fields_ = [("data", _vp), ("batch_stride", _i64), ("coeff_stride", _i64),
                ("min_degree", _i32), ("max_degree", _i32)]


class ScalarBatch(C.Structure):
    _fields_ = [("data", _vp), ("stride", _i64)]


class FtMulDesc(C.Structure):
    _fields_ = [("basis", Basis), ("num_batches", _i64), ("out", TensorBatch), ("lhs", TensorBatch),
                ("rhs", TensorBatch), ("beta", _f64)]


class FtFmaDesc(C.Structure):
    _fields_ = [("basis", Basis), ("num_batches", _i64), ("out", TensorBatch), ("a", TensorBatch),
                ("b", TensorBatch), ("c", TensorBatch), ("alpha", _f64), ("beta", _f64)]


class FtInplaceFmaDesc(C.Structure):
    _fields_ = [("basis", Basis), ("num_batches", _i64), ("variant", _i32), ("a", TensorBatch),
                ("b", TensorBatch), ("c", TensorBatch), ("alpha", _f64), ("beta", _f64)]


class FtInplaceMulDesc(C.Structure):
    _fields_ = [("basis", Basis), ("num_batches", _i64), ("lhs", TensorBatch), ("rhs", TensorBatch),
                ("beta", _f64)]


class FtUnaryDesc(C.Structure):
    _fields_ = [("basis", Basis), ("num_batches", _i64), ("out", TensorBatch), ("arg", TensorBatch)]


class FtFmexpDesc(C.Structure):
    _fields_ = [("basis", Basis), ("num_batches", _i64), ("out", TensorBatch),
                ("multiplier", TensorBatch), ("exponent", TensorBatch)]


class FtFmexpAdjointDesc(C.Structure):
    _fields_ = [("basis", Basis), ("num_batches", _i64), ("multiplier", TensorBatch),
                ("exponent", TensorBatch), ("cotangent", TensorBatch), ("grad_multiplier", TensorBatch),
                ("grad_exponent", TensorBatch)]


class FtLogAdjointDesc(C.Structure):
    _fields_ = [("basis", Basis), ("num_batches", _i64), ("arg", TensorBatch), ("cotangent", TensorBatch),
                ("grad_arg", TensorBatch)]


class PairingDesc(C.Structure):
    _fields_ = [("basis", Basis), ("num_batches", _i64), ("out", ScalarBatch),
                ("functional", TensorBatch), ("arg", TensorBatch)]


class AdjointDesc(C.Structure):
    _fields_ = [("basis", Basis), ("num_batches", _i64), ("out", TensorBatch), ("op", TensorBatch),
                ("arg", TensorBatch)]


class SparseMatrix(C.Structure):
    _fields_ = [("format", _i32), ("values", _vp), ("indices", _vp), ("offsets", _vp), ("nnz", _i64),
                ("rows", _i64), ("cols", _i64)]


class VectorBatch(C.Structure):
    _fields_ = [("data", _vp), ("batch_stride", _i64), ("coeff_stride", _i64), ("begin", _i64),
                ("size", _i64)]


class SpmvDesc(C.Structure):
    _fields_ = [("num_batches", _i64), ("out", VectorBatch), ("arg", VectorBatch),
                ("matrix", SparseMatrix), ("alpha", _f64)]


class LieConvertDesc(C.Structure):
    _fields_ = [("basis", Basis), ("num_batches", _i64), ("tensor", TensorBatch), ("lie", TensorBatch)]


class LieMatrixConvertDesc(C.Structure):
    _fields_ = [("basis", Basis), ("num_batches", _i64), ("tensor", TensorBatch), ("lie", TensorBatch),
                ("matrix", SparseMatrix)]


class StInplaceFmaDesc(C.Structure):
    _fields_ = [("basis", Basis), ("num_batches", _i64), ("a", TensorBatch), ("b", TensorBatch),
                ("c", TensorBatch), ("alpha", _f64), ("beta", _f64)]


class HalfShuffleDesc(C.Structure):
    _fields_ = [("basis", Basis), ("num_batches", _i64), ("out", TensorBatch), ("integrator", TensorBatch),
                ("integrand", TensorBatch), ("beta", _f64)]


class GradedBasis(C.Structure):
    _fields_ = [("kind", _i32), ("width", _i32), ("depth", _i32)]


TENSOR_BASIS = 0
LIE_BASIS = 1


class VectorAssignDesc(C.Structure):
    _fields_ = [("basis", GradedBasis), ("num_batches", _i64), ("out", TensorBatch), ("arg", TensorBatch)]


class VectorAddDesc(C.Structure):
    _fields_ = [("basis", GradedBasis), ("num_batches", _i64), ("out", TensorBatch), ("lhs", TensorBatch),
                ("rhs", TensorBatch), ("alpha", _f64), ("beta", _f64)]


class VectorInplaceAddDesc(C.Structure):
    _fields_ = [("basis", GradedBasis), ("num_batches", _i64), ("lhs", TensorBatch), ("rhs", TensorBatch),
                ("alpha", _f64)]


class VectorScalarDesc(C.Structure):
    _fields_ = [("basis", GradedBasis), ("num_batches", _i64), ("vec", TensorBatch), ("scalar", _f64)]


class SignatureDesc(C.Structure):
    _fields_ = [("basis", Basis), ("num_paths", _i64), ("increments", _vp), ("offsets", _vp),
                ("out", TensorBatch), ("checkpoint_every", _i32), ("checkpoints", _vp),
                ("ckpt_offsets", _vp), ("tile_threads", _i32), ("tile_split", _i32),
                ("tile_kind", _i32), ("tile_expand", _i32), ("order", _vp), ("num_increments", _i64)]


class SignatureBackwardDesc(C.Structure):
    _fields_ = [("basis", Basis), ("num_paths", _i64), ("increments", _vp), ("offsets", _vp),
                ("signatures", TensorBatch), ("cotangent", TensorBatch), ("grad_increments", _vp),
                ("checkpoint_every", _i32), ("checkpoints", _vp), ("ckpt_offsets", _vp),
                ("scratch", _vp), ("tile_threads", _i32), ("tile_split", _i32), ("order", _vp),
                ("num_increments", _i64), ("tile_expand", _i32), ("tile_kind", _i32)]


# entry point -> descriptor type; every one also takes a stream pointer.
ENTRY_POINTS = {
    "rpp_hip_ft_mul": FtMulDesc,
    "rpp_hip_ft_fma": FtFmaDesc,
    "rpp_hip_ft_inplace_fma": FtInplaceFmaDesc,
    "rpp_hip_ft_inplace_mul": FtInplaceMulDesc,
    "rpp_hip_ft_exp": FtUnaryDesc,
    "rpp_hip_ft_log": FtUnaryDesc,
    "rpp_hip_tensor_antipode": FtUnaryDesc,
    "rpp_hip_tensor_reflect": FtUnaryDesc,
    "rpp_hip_ft_log_adjoint": FtLogAdjointDesc,
    "rpp_hip_ft_fmexp": FtFmexpDesc,
    "rpp_hip_ft_fmexp_adjoint": FtFmexpAdjointDesc,
    "rpp_hip_tensor_pairing": PairingDesc,
    "rpp_hip_ft_adj_lmul": AdjointDesc,
    "rpp_hip_ft_adj_rmul": AdjointDesc,
    "rpp_hip_sparse_matrix_vector_product": SpmvDesc,
    "rpp_hip_tensor_to_lie": LieConvertDesc,
    "rpp_hip_lie_to_tensor": LieConvertDesc,
    "rpp_hip_tensor_to_lie_matrix": LieMatrixConvertDesc,
    "rpp_hip_lie_to_tensor_matrix": LieMatrixConvertDesc,
    "rpp_hip_st_mul": FtMulDesc,
    "rpp_hip_st_fma": FtFmaDesc,
    "rpp_hip_st_inplace_fma": StInplaceFmaDesc,
    "rpp_hip_st_adj_mul": AdjointDesc,
    "rpp_hip_left_hs_mul": HalfShuffleDesc,
    "rpp_hip_left_hs_adj_lmul": HalfShuffleDesc,
    "rpp_hip_left_hs_adj_rmul": HalfShuffleDesc,
    "rpp_hip_vector_assign": VectorAssignDesc,
    "rpp_hip_vector_add": VectorAddDesc,
    "rpp_hip_vector_inplace_add": VectorInplaceAddDesc,
    "rpp_hip_vector_scalar_multiply": VectorScalarDesc,
    "rpp_hip_vector_set_constant": VectorScalarDesc,
    "rpp_hip_tensor_set_identity": VectorScalarDesc,
    "rpp_hip_tensor_add_identity": VectorScalarDesc,
    "rpp_hip_signature": SignatureDesc,
    "rpp_hip_signature_backward": SignatureBackwardDesc,
    "rpp_hip_signature_backward_noalloc": SignatureBackwardDesc,
}

OTHER_SYMBOLS = ["rpp_hip_lie_basis_sizes", "rpp_hip_lie_basis_copy", "rpp_hip_lie_basis_save",
                 "rpp_hip_lie_basis_load", "rpp_hip_last_error",
                 "rpp_hip_version", "rpp_hip_signature_backward_scratch_bytes", "rpp_hip_signature_tiles",
                 "rpp_hip_fmexp_tile", "rpp_hip_sparse_matrix_forget", "rpp_hip_signature_checkpoint_every"]


def exported_symbols() -> list[str]:
    names = []
    for base in ENTRY_POINTS:
        names += [base + "_f32", base + "_f64"]
    return names + OTHER_SYMBOLS


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"librpp_hip.so is not built ({LIB_PATH}); run __graft_entry__.build() "
                "or make -C rough-path-primitives_amd/csrc")
        lib_ = C.CDLL(LIB_PATH)
        for base, desc in ENTRY_POINTS.items():
            for sfx in ("_f32", "_f64"):
                fn = getattr(lib_, base + sfx)
                fn.argtypes = [C.POINTER(desc), _vp]
                fn.restype = C.c_int
        lib_.rpp_hip_last_error.restype = C.c_char_p
        lib_.rpp_hip_version.restype = C.c_char_p
        lib_.rpp_hip_lie_basis_sizes.argtypes = [_i32, _i32, _vp]
        lib_.rpp_hip_lie_basis_copy.argtypes = [_i32, _i32] + [_vp] * 8
        lib_.rpp_hip_lie_basis_save.argtypes = [_i32, _i32, C.c_char_p]
        lib_.rpp_hip_lie_basis_load.argtypes = [C.c_char_p, _vp, _vp]
        lib_.rpp_hip_signature_backward_scratch_bytes.argtypes = [C.POINTER(SignatureBackwardDesc), _i32]
        lib_.rpp_hip_signature_backward_scratch_bytes.restype = _i64
        lib_.rpp_hip_signature_checkpoint_every.argtypes = [_i32, _i64]
        lib_.rpp_hip_signature_checkpoint_every.restype = _i32
        lib_.rpp_hip_sparse_matrix_forget.argtypes = [_vp]
        lib_.rpp_hip_sparse_matrix_forget.restype = C.c_int
        lib_.rpp_hip_signature_tiles.argtypes = [_i32, _vp, _i32]
        lib_.rpp_hip_signature_tiles.restype = C.c_int
        _lib = lib_
    return _lib
