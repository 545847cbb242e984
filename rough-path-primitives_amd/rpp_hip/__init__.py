"""rpp_hip -- MI355X (gfx950) strategy for the truncated tensor algebra hot
path of RoughPathPrimitives, over librpp_hip.so (include/rpp_hip.h)."""
from ._capi import LIB_PATH, exported_symbols, lib
from .ops import (Batch, LieBasis, RppHipError, SparseMatrix, TensorBasis, checkpoint_every_for, fmexp_tile, ft_adj_lmul, ft_adj_rmul,
                  ft_exp, ft_fma, ft_fmexp, ft_fmexp_adjoint, ft_inplace_fma, ft_inplace_mul, ft_log, ft_log_adjoint,
                  ft_mul, interleaved, lie_basis, lie_basis_load, lie_basis_save, lie_to_tensor, make_tensor_batch, path_order, signature,
                  signature_backward, signature_tiles, sparse_matrix_forget, sparse_matrix_vector_product, st_adj_mul, st_fma,
                  st_inplace_fma, st_mul, left_hs_adj_lmul, left_hs_adj_rmul, left_hs_mul, tensor_add_identity,
                  tensor_antipode, tensor_dimension, tensor_pairing, tensor_reflect, tensor_set_identity,
                  tensor_to_lie, to_interleaved, to_object_major, vector_add, vector_assign, vector_inplace_add,
                  vector_scalar_multiply, vector_set_constant)
from .graph import SignatureGraph

__all__ = [
    "Batch", "LieBasis", "RppHipError", "SignatureGraph", "SparseMatrix", "TensorBasis", "checkpoint_every_for",
    "fmexp_tile",
    "ft_adj_lmul", "ft_adj_rmul", "ft_exp", "ft_fma", "ft_fmexp", "ft_fmexp_adjoint", "ft_inplace_fma",
    "ft_inplace_mul", "ft_log", "ft_log_adjoint", "ft_mul", "interleaved", "lie_basis", "lie_basis_load",
    "lie_basis_save", "lie_to_tensor",
    "make_tensor_batch", "path_order", "signature", "signature_backward", "signature_tiles",
    "sparse_matrix_forget", "sparse_matrix_vector_product", "st_adj_mul", "st_fma", "st_inplace_fma", "st_mul", "left_hs_adj_lmul",
    "left_hs_adj_rmul", "left_hs_mul", "tensor_add_identity", "tensor_antipode", "tensor_dimension", "tensor_pairing",
    "tensor_reflect", "tensor_set_identity", "tensor_to_lie", "to_interleaved", "to_object_major", "vector_add",
    "vector_assign", "vector_inplace_add", "vector_scalar_multiply", "vector_set_constant", "LIB_PATH",
    "exported_symbols", "lib",
]
