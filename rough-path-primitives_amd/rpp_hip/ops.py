"""Operation API over torch device tensors, mirroring the reference's free
functions (rpp::ops::ft_mul etc., operations/**.hpp) with the hip::block
strategy implicit: ``strategy`` and ``LaunchConfig`` collapse into an
optional ``stream`` (the reference's GPU LaunchConfig carries exactly a
stream, gpu/device.hpp:21-24).

Batches mirror ``make_tensor_batch`` (views/dense_tensor_view.hpp:38-91):

* a 2-D tensor ``[n, stride]`` is object-major (``layouts::StrideLayout``);
* a 1-D tensor is one object broadcast to the batch (stride 0);
* ``interleaved(t)`` marks a ``[dim, n]`` tensor as batch-interleaved
  (coefficient i of object b at ``t[i, b]``).

Errors raise :class:`RppHipError` carrying the rpp::ErrorCode value
(support/error.hpp:28-39); nothing falls back to the CPU.
"""
from __future__ import annotations

import ctypes as C
import weakref
from dataclasses import dataclass
from typing import Optional

import torch

from . import _capi as A

ERROR_NAMES = {0: "Success", 1: "Cancelled", 2: "Unknown", 3: "InvalidArgument", 4: "Timeout",
               5: "OutOfResources", 6: "ContractViolation", 7: "OutOfBounds", 8: "NotImplemented",
               9: "Internal"}


class RppHipError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{ERROR_NAMES.get(code, code)}: {message}")
        self.code = code


@dataclass(frozen=True)
class TensorBasis:
    """TensorBasis(width, depth) (basis/tensor_basis.hpp:40-118)."""
    width: int
    depth: int

    def start_of_degree(self, k: int) -> int:
        s, p = 0, 1
        for _ in range(k):
            s += p
            p *= self.width
        return s

    def end_of_degree(self, k: int) -> int:
        return self.start_of_degree(k + 1)

    def size_of_degree(self, k: int) -> int:
        return self.width ** k

    def size(self) -> int:
        return self.start_of_degree(self.depth + 1)

    def _c(self):
        return A.Basis(self.width, self.depth)


def tensor_dimension(width: int, depth: int) -> int:
    return TensorBasis(width, depth).size()


def _witt(width: int, k: int) -> int:
    """Number of Lie elements of degree k (Witt's formula)."""
    def mobius(m):
        mu, p = 1, 2
        while p * p <= m:
            if m % p == 0:
                m //= p
                if m % p == 0:
                    return 0
                mu = -mu
            p += 1
        return -mu if m > 1 else mu
    return sum(mobius(d) * width ** (k // d) for d in range(1, k + 1) if k % d == 0) // k


@dataclass(frozen=True)
class LieBasis:
    """The degree grading of the Hall-set LieBasis(width, depth)
    (basis/lie_basis.hpp:42-97): degree 0 holds one slot, degree k the Witt
    number of Lie elements; storage is true_size() per object."""
    width: int
    depth: int

    def start_of_degree(self, k: int) -> int:
        return 0 if k == 0 else 1 + sum(_witt(self.width, j) for j in range(1, k))

    def end_of_degree(self, k: int) -> int:
        return self.start_of_degree(k + 1)

    def size_of_degree(self, k: int) -> int:
        return 1 if k == 0 else _witt(self.width, k)

    def true_size(self) -> int:
        return self.start_of_degree(self.depth + 1)

    def size(self) -> int:
        return 0 if self.width <= 1 or self.depth == 0 else self.true_size() - 1


def _graded(basis) -> A.GradedBasis:
    kind = A.LIE_BASIS if isinstance(basis, LieBasis) else A.TENSOR_BASIS
    return A.GradedBasis(kind, basis.width, basis.depth)


@dataclass
class Batch:
    """A batch of graded views: data + window [min_degree, max_degree]."""
    data: torch.Tensor
    min_degree: int
    max_degree: int
    interleaved: bool = False

    def truncate(self, mn: int, mx: int) -> "Batch":
        return Batch(self.data, max(mn, self.min_degree), min(mx, self.max_degree), self.interleaved)

    def _c(self) -> A.TensorBatch:
        t = self.data
        if not t.is_cuda:
            raise RppHipError(3, "batch data must be a device tensor")
        if self.interleaved:
            if t.dim() != 2 or t.stride(1) != 1:
                raise RppHipError(3, "interleaved batches are [dim, n] with unit batch stride")
            return A.TensorBatch(t.data_ptr(), 1, t.stride(0), self.min_degree, self.max_degree)
        if t.dim() == 1:
            if t.stride(0) != 1:
                raise RppHipError(3, "a broadcast object must be contiguous")
            return A.TensorBatch(t.data_ptr(), 0, 1, self.min_degree, self.max_degree)
        if t.dim() != 2 or t.stride(1) != 1:
            raise RppHipError(3, "object-major batches are [n, stride] with unit coefficient stride")
        return A.TensorBatch(t.data_ptr(), t.stride(0), 1, self.min_degree, self.max_degree)

    @property
    def num_objects(self) -> int:
        t = self.data
        if t.dim() == 1:
            return 1
        return t.shape[1] if self.interleaved else t.shape[0]


def make_tensor_batch(data: torch.Tensor, basis: TensorBasis, min_degree: int = 0,
                      max_degree: Optional[int] = None) -> Batch:
    return Batch(data, min_degree, basis.depth if max_degree is None else max_degree)


def interleaved(data: torch.Tensor, basis: TensorBasis, min_degree: int = 0,
                max_degree: Optional[int] = None) -> Batch:
    return Batch(data, min_degree, basis.depth if max_degree is None else max_degree, True)


def _as_batch(x, basis) -> Batch:
    if isinstance(x, Batch):
        return x
    return make_tensor_batch(x, basis)


def _sfx(*batches) -> str:
    dt = None
    for b in batches:
        d = b.data.dtype if isinstance(b, Batch) else b.dtype
        if dt is None:
            dt = d
        elif d != dt:
            raise RppHipError(3, f"mixed dtypes {dt} and {d}")
    if dt == torch.float64:
        return "_f64"
    if dt == torch.float32:
        return "_f32"
    raise RppHipError(8, f"unsupported dtype {dt}")


# The current stream's raw handle without building a torch.cuda.Stream object
# (2.7 us per call on the host, against 0.1 us): every operation asks for it.
_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def _stream(stream) -> int:
    if stream is None:
        if _raw_stream is not None:
            return _raw_stream(torch.cuda.current_device())
        return torch.cuda.current_stream().cuda_stream
    if isinstance(stream, torch.cuda.Stream):
        return stream.cuda_stream
    return int(stream)


def _call(name: str, desc, stream) -> None:
    lib = A.lib()
    rc = getattr(lib, name)(C.byref(desc), C.c_void_p(_stream(stream)))
    if rc != 0:
        msg = lib.rpp_hip_last_error()
        raise RppHipError(rc, msg.decode() if msg else name)


def _n(num_batches, *batches) -> int:
    if num_batches is not None:
        return int(num_batches)
    return max(b.num_objects for b in batches)


# ------------------------------------------------------------------ products

def ft_mul(out, lhs, rhs, basis: TensorBasis, num_batches=None, beta: float = 1.0, stream=None):
    """rpp::ops::ft_mul (operations/basic/ft_mul.hpp:53-79)."""
    out, lhs, rhs = (_as_batch(x, basis) for x in (out, lhs, rhs))
    d = A.FtMulDesc(basis._c(), _n(num_batches, out), out._c(), lhs._c(), rhs._c(), beta)
    _call("rpp_hip_ft_mul" + _sfx(out, lhs, rhs), d, stream)


def ft_fma(out, a, b, c, basis: TensorBasis, num_batches=None, alpha: float = 1.0, beta: float = 1.0,
           stream=None):
    """rpp::ops::ft_fma (operations/basic/ft_fma.hpp:54-83)."""
    out, a, b, c = (_as_batch(x, basis) for x in (out, a, b, c))
    d = A.FtFmaDesc(basis._c(), _n(num_batches, out), out._c(), a._c(), b._c(), c._c(), alpha, beta)
    _call("rpp_hip_ft_fma" + _sfx(out, a, b, c), d, stream)


def ft_inplace_fma(variant: int, a, b, c, basis: TensorBasis, num_batches=None, alpha: float = 1.0,
                   beta: float = 1.0, stream=None):
    """rpp::ops::ft_inplace_fma<FMAType> (operations/basic/ft_inplace_fma.hpp:64-100);
    variant 231 (AEqualsBCPlusA), 123 (AEqualsABPlusC) or 213 (AEqualsBAPlusC)."""
    a, b, c = (_as_batch(x, basis) for x in (a, b, c))
    d = A.FtInplaceFmaDesc(basis._c(), _n(num_batches, a), variant, a._c(), b._c(), c._c(), alpha, beta)
    _call("rpp_hip_ft_inplace_fma" + _sfx(a, b, c), d, stream)


def ft_inplace_mul(lhs, rhs, basis: TensorBasis, num_batches=None, beta: float = 1.0, stream=None):
    """rpp::ops::ft_inplace_mul (operations/basic/ft_inplace_mul.hpp:38-69)."""
    lhs, rhs = (_as_batch(x, basis) for x in (lhs, rhs))
    d = A.FtInplaceMulDesc(basis._c(), _n(num_batches, lhs), lhs._c(), rhs._c(), beta)
    _call("rpp_hip_ft_inplace_mul" + _sfx(lhs, rhs), d, stream)


# ---------------------------------------------------------------- composites

def ft_exp(out, arg, basis: TensorBasis, num_batches=None, stream=None):
    """rpp::ops::ft_exp (operations/intermediate/ft_exp.hpp:89-112)."""
    out, arg = (_as_batch(x, basis) for x in (out, arg))
    d = A.FtUnaryDesc(basis._c(), _n(num_batches, out), out._c(), arg._c())
    _call("rpp_hip_ft_exp" + _sfx(out, arg), d, stream)


def ft_log(out, arg, basis: TensorBasis, num_batches=None, stream=None):
    """rpp::ops::ft_log (operations/intermediate/ft_log.hpp:91-114)."""
    out, arg = (_as_batch(x, basis) for x in (out, arg))
    d = A.FtUnaryDesc(basis._c(), _n(num_batches, out), out._c(), arg._c())
    _call("rpp_hip_ft_log" + _sfx(out, arg), d, stream)


def tensor_antipode(out, arg, basis: TensorBasis, num_batches=None, stream=None):
    """rpp::ops::tensor_antipode (operations/basic/tensor_antipode.hpp:15-39):
    out[reverse(w)] = (-1)^|w| arg[w] on out's window."""
    out, arg = (_as_batch(x, basis) for x in (out, arg))
    d = A.FtUnaryDesc(basis._c(), _n(num_batches, out), out._c(), arg._c())
    _call("rpp_hip_tensor_antipode" + _sfx(out, arg), d, stream)


def tensor_reflect(out, arg, basis: TensorBasis, num_batches=None, stream=None):
    """rpp::ops::tensor_reflect (operations/basic/tensor_reflect.hpp:15-39):
    out[reverse(w)] = arg[w] on out's window."""
    out, arg = (_as_batch(x, basis) for x in (out, arg))
    d = A.FtUnaryDesc(basis._c(), _n(num_batches, out), out._c(), arg._c())
    _call("rpp_hip_tensor_reflect" + _sfx(out, arg), d, stream)


def ft_log_adjoint(grad_arg, arg, cotangent, basis: TensorBasis, num_batches=None, stream=None):
    """Adjoint of ft_log (operations/intermediate/ft_log.hpp:66-85): grad_arg =
    d loss / d arg on arg's window, given cotangent = d loss / d log(arg)."""
    ga, a, g = (_as_batch(t, basis) for t in (grad_arg, arg, cotangent))
    d = A.FtLogAdjointDesc(basis._c(), _n(num_batches, ga), a._c(), g._c(), ga._c())
    _call("rpp_hip_ft_log_adjoint" + _sfx(ga, a, g), d, stream)


def ft_fmexp(out, multiplier, exponent, basis: TensorBasis, num_batches=None, stream=None):
    """rpp::ops::ft_fmexp (operations/intermediate/ft_fmexp.hpp:96-121)."""
    out, multiplier, exponent = (_as_batch(x, basis) for x in (out, multiplier, exponent))
    d = A.FtFmexpDesc(basis._c(), _n(num_batches, out), out._c(), multiplier._c(), exponent._c())
    _call("rpp_hip_ft_fmexp" + _sfx(out, multiplier, exponent), d, stream)


def ft_fmexp_adjoint(grad_multiplier, grad_exponent, multiplier, exponent, cotangent, basis: TensorBasis,
                     num_batches=None, stream=None):
    """Adjoint of ft_fmexp for a degree-one exponent (one Chen step
    out = multiplier (x) exp(x)).  Writes grad_multiplier = R_{exp(x)}^* cotangent
    (ft_adj_rmul with op = exp(x), operations/basic/ft_adj_rmul.hpp:47-71) and
    the exponent gradient into degree 1 of grad_exponent.  The exponent windows
    end at degree 1; grad_multiplier holds the full window."""
    gm, gx, m, x, g = (_as_batch(t, basis) for t in (grad_multiplier, grad_exponent, multiplier, exponent,
                                                      cotangent))
    d = A.FtFmexpAdjointDesc(basis._c(), _n(num_batches, gm), m._c(), x._c(), g._c(), gm._c(), gx._c())
    _call("rpp_hip_ft_fmexp_adjoint" + _sfx(gm, gx, m, x, g), d, stream)


def tensor_pairing(out: torch.Tensor, functional, arg, basis: TensorBasis, num_batches=None, stream=None):
    """rpp::ops::tensor_pairing (operations/basic/tensor_pairing.hpp:47-73);
    ``out`` is a 1-D device tensor (make_scalar_batch)."""
    functional, arg = (_as_batch(x, basis) for x in (functional, arg))
    if not out.is_cuda or out.dim() != 1:
        raise RppHipError(3, "out must be a 1-D device tensor")
    n = _n(num_batches, functional, arg) if num_batches is None else num_batches
    d = A.PairingDesc(basis._c(), n, A.ScalarBatch(out.data_ptr(), out.stride(0)), functional._c(), arg._c())
    _call("rpp_hip_tensor_pairing" + _sfx(out, functional, arg), d, stream)


def ft_adj_lmul(out, op, arg, basis: TensorBasis, num_batches=None, stream=None):
    """rpp::ops::ft_adj_lmul (operations/basic/ft_adj_lmul.hpp:49-73)."""
    out, op, arg = (_as_batch(x, basis) for x in (out, op, arg))
    d = A.AdjointDesc(basis._c(), _n(num_batches, out), out._c(), op._c(), arg._c())
    _call("rpp_hip_ft_adj_lmul" + _sfx(out, op, arg), d, stream)


def ft_adj_rmul(out, op, arg, basis: TensorBasis, num_batches=None, stream=None):
    """rpp::ops::ft_adj_rmul (operations/basic/ft_adj_rmul.hpp:47-71)."""
    out, op, arg = (_as_batch(x, basis) for x in (out, op, arg))
    d = A.AdjointDesc(basis._c(), _n(num_batches, out), out._c(), op._c(), arg._c())
    _call("rpp_hip_ft_adj_rmul" + _sfx(out, op, arg), d, stream)


# --------------------------------------------------------- shuffle family

def st_mul(out, lhs, rhs, basis: TensorBasis, num_batches=None, beta: float = 1.0, stream=None):
    """rpp::ops::st_mul (operations/basic/st_mul.hpp:48-83): out = beta lhs . rhs
    (shuffle product)."""
    out, lhs, rhs = (_as_batch(x, basis) for x in (out, lhs, rhs))
    d = A.FtMulDesc(basis._c(), _n(num_batches, out), out._c(), lhs._c(), rhs._c(), beta)
    _call("rpp_hip_st_mul" + _sfx(out, lhs, rhs), d, stream)


def st_fma(out, a, b, c, basis: TensorBasis, num_batches=None, alpha: float = 1.0, beta: float = 1.0, stream=None):
    """rpp::ops::st_fma (operations/basic/st_fma.hpp:50-89): out = alpha a + beta b . c."""
    out, a, b, c = (_as_batch(x, basis) for x in (out, a, b, c))
    d = A.FtFmaDesc(basis._c(), _n(num_batches, out), out._c(), a._c(), b._c(), c._c(), alpha, beta)
    _call("rpp_hip_st_fma" + _sfx(out, a, b, c), d, stream)


def st_inplace_fma(a, b, c, basis: TensorBasis, num_batches=None, alpha: float = 1.0, beta: float = 1.0,
                   stream=None):
    """rpp::ops::st_inplace_fma (operations/basic/st_inplace_fma.hpp:45-85):
    a <- alpha a + beta b . c."""
    a, b, c = (_as_batch(x, basis) for x in (a, b, c))
    d = A.StInplaceFmaDesc(basis._c(), _n(num_batches, a), a._c(), b._c(), c._c(), alpha, beta)
    _call("rpp_hip_st_inplace_fma" + _sfx(a, b, c), d, stream)


def st_adj_mul(out, op, arg, basis: TensorBasis, num_batches=None, stream=None):
    """rpp::ops::st_adj_mul (operations/basic/st_adj_mul.hpp:40-73):
    <out, x> = <arg, op . x>."""
    out, op, arg = (_as_batch(x, basis) for x in (out, op, arg))
    d = A.AdjointDesc(basis._c(), _n(num_batches, out), out._c(), op._c(), arg._c())
    _call("rpp_hip_st_adj_mul" + _sfx(out, op, arg), d, stream)


def _half(name, out, integrator, integrand, basis, num_batches, beta, stream):
    out, x, y = (_as_batch(t, basis) for t in (out, integrator, integrand))
    d = A.HalfShuffleDesc(basis._c(), _n(num_batches, out), out._c(), x._c(), y._c(), beta)
    _call(name + _sfx(out, x, y), d, stream)


def left_hs_mul(out, integrator, integrand, basis: TensorBasis, num_batches=None, beta: float = 1.0, stream=None):
    """rpp::ops::left_hs_mul (operations/basic/left_hs_mul.hpp:53-93):
    out = beta integrator < integrand (left half-shuffle)."""
    _half("rpp_hip_left_hs_mul", out, integrator, integrand, basis, num_batches, beta, stream)


def left_hs_adj_lmul(out, integrator, integrand, basis: TensorBasis, num_batches=None, beta: float = 1.0,
                     stream=None):
    """rpp::ops::left_hs_adj_lmul (operations/basic/left_hs_adj_lmul.hpp:33-61):
    <out, z> = beta <integrand, integrator < z>."""
    _half("rpp_hip_left_hs_adj_lmul", out, integrator, integrand, basis, num_batches, beta, stream)


def left_hs_adj_rmul(out, integrator, integrand, basis: TensorBasis, num_batches=None, beta: float = 1.0,
                     stream=None):
    """rpp::ops::left_hs_adj_rmul (operations/basic/left_hs_adj_rmul.hpp:33-61):
    <out, z> = beta <integrator, z < integrand>."""
    _half("rpp_hip_left_hs_adj_rmul", out, integrator, integrand, basis, num_batches, beta, stream)


# -------------------------------------------------------- linear operations

def vector_assign(out, arg, basis, num_batches=None, stream=None):
    """rpp::ops::vector_assign (operations/linalg/vector_assign.hpp:36-64): out =
    arg on the overlap of the windows, 0 on the rest of out's window.  ``basis``
    is a TensorBasis or a LieBasis.  Between an object-major and an
    interleaved batch this is the layout transposition."""
    out, arg = (_as_batch(x, basis) for x in (out, arg))
    d = A.VectorAssignDesc(_graded(basis), _n(num_batches, out), out._c(), arg._c())
    _call("rpp_hip_vector_assign" + _sfx(out, arg), d, stream)


def vector_add(out, lhs, rhs, basis, num_batches=None, alpha: float = 1.0, beta: float = 1.0, stream=None):
    """rpp::ops::vector_add (operations/linalg/vector_add.hpp:40-78): out =
    alpha*lhs + beta*rhs on out's window, each operand 0 outside its own."""
    out, lhs, rhs = (_as_batch(x, basis) for x in (out, lhs, rhs))
    d = A.VectorAddDesc(_graded(basis), _n(num_batches, out), out._c(), lhs._c(), rhs._c(), alpha, beta)
    _call("rpp_hip_vector_add" + _sfx(out, lhs, rhs), d, stream)


def vector_inplace_add(lhs, rhs, basis, num_batches=None, alpha: float = 1.0, stream=None):
    """rpp::ops::vector_inplace_add (operations/linalg/vector_inplace_add.hpp:40-77):
    lhs += alpha*rhs on the overlap of the windows."""
    lhs, rhs = (_as_batch(x, basis) for x in (lhs, rhs))
    d = A.VectorInplaceAddDesc(_graded(basis), _n(num_batches, lhs), lhs._c(), rhs._c(), alpha)
    _call("rpp_hip_vector_inplace_add" + _sfx(lhs, rhs), d, stream)


def _scalar_op(name, vec, basis, num_batches, scalar, stream):
    vec = _as_batch(vec, basis)
    d = A.VectorScalarDesc(_graded(basis), _n(num_batches, vec), vec._c(), float(scalar))
    _call(name + _sfx(vec), d, stream)


def vector_scalar_multiply(vec, basis, num_batches=None, *, scalar: float, stream=None):
    """rpp::ops::vector_scalar_multiply (operations/linalg/vector_scalar_multiply.hpp:28-71):
    vec *= scalar on its window."""
    _scalar_op("rpp_hip_vector_scalar_multiply", vec, basis, num_batches, scalar, stream)


def vector_set_constant(vec, basis, num_batches=None, value: float = 0.0, stream=None):
    """rpp::ops::vector_set_constant (operations/linalg/vector_set_constant.hpp:27-72)."""
    _scalar_op("rpp_hip_vector_set_constant", vec, basis, num_batches, value, stream)


def tensor_set_identity(tensor, basis, num_batches=None, scalar: float = 1.0, stream=None):
    """rpp::ops::tensor_set_identity (operations/basic/tensor_set_identity.hpp:27-73):
    the window is zeroed and degree 0, if held, set to scalar."""
    _scalar_op("rpp_hip_tensor_set_identity", tensor, basis, num_batches, scalar, stream)


def tensor_add_identity(tensor, basis, num_batches=None, scalar: float = 1.0, stream=None):
    """rpp::ops::tensor_add_identity (operations/basic/tensor_add_identity.hpp:26-71)."""
    _scalar_op("rpp_hip_tensor_add_identity", tensor, basis, num_batches, scalar, stream)


def to_interleaved(x: torch.Tensor, basis, stream=None) -> torch.Tensor:
    """[n, dim] object-major -> a new [dim, n] batch-interleaved tensor."""
    out = torch.empty((x.shape[1], x.shape[0]), dtype=x.dtype, device=x.device)
    vector_assign(interleaved(out, basis), make_tensor_batch(x, basis), basis, x.shape[0], stream)
    return out


def to_object_major(x: torch.Tensor, basis, stream=None) -> torch.Tensor:
    """[dim, n] batch-interleaved -> a new [n, dim] object-major tensor."""
    out = torch.empty((x.shape[1], x.shape[0]), dtype=x.dtype, device=x.device)
    vector_assign(make_tensor_batch(out, basis), interleaved(x, basis), basis, x.shape[1], stream)
    return out


# ------------------------------------------------------------- sparse, Lie

def _forget_values(ptr: int) -> None:
    A.lib().rpp_hip_sparse_matrix_forget(ptr)


@dataclass
class SparseMatrix:
    """CompressedMatrix with device arrays (sparse/detail/compressed_matrix.hpp:38-126).

    The library caches what it derives from a matrix (the CSR form of a CSC
    matrix, whether it is a Hall T2L/L2T matrix) by its device arrays, which
    must not change while this object is in use; the cache entry is dropped
    when this object is collected, so its memory can be reused safely."""
    format: str  # "csr" or "csc"
    values: torch.Tensor
    indices: torch.Tensor  # int64
    offsets: torch.Tensor  # int64
    rows: int
    cols: int

    def __post_init__(self):
        fin = weakref.finalize(self, _forget_values, self.values.data_ptr())
        fin.atexit = False

    def _c(self):
        return A.SparseMatrix(0 if self.format == "csr" else 1, self.values.data_ptr(),
                              self.indices.data_ptr(), self.offsets.data_ptr(), self.values.numel(),
                              self.rows, self.cols)


def _vec(t: torch.Tensor, begin: int, size: int) -> A.VectorBatch:
    if t.dim() == 1:
        return A.VectorBatch(t.data_ptr(), 0, t.stride(0), begin, size)
    return A.VectorBatch(t.data_ptr(), t.stride(0), t.stride(1), begin, size)


def sparse_matrix_vector_product(out: torch.Tensor, arg: torch.Tensor, matrix: SparseMatrix,
                                 out_begin: int = 0, arg_begin: int = 0, out_size: Optional[int] = None,
                                 alpha: float = 1.0, num_batches=None, stream=None):
    """rpp::ops::sparse_matrix_vector_product (operations/linalg/sparse_matrix_vector.hpp:57-96)
    with view-relative indexing: row 0 <-> out_begin, column 0 <-> arg_begin."""
    n = out.shape[0] if num_batches is None else num_batches
    if out_size is None:
        out_size = matrix.rows
    d = A.SpmvDesc(n, _vec(out, out_begin, out_size), _vec(arg, arg_begin, matrix.cols), matrix._c(), alpha)
    _call("rpp_hip_sparse_matrix_vector_product" + _sfx(out, arg, matrix.values), d, stream)


def sparse_matrix_forget(matrix: SparseMatrix) -> None:
    """Drop what the library cached for this matrix (its CSR form when it is
    CSC, its Hall-matrix recognition); call before rewriting or freeing its
    arrays."""
    rc = A.lib().rpp_hip_sparse_matrix_forget(matrix.values.data_ptr())
    if rc:
        raise RppHipError(rc, A.lib().rpp_hip_last_error().decode())


def tensor_to_lie(lie_out, tensor_in, basis: TensorBasis, num_batches=None, stream=None,
                  matrix: Optional[SparseMatrix] = None):
    """rpp::ops::tensor_to_lie (operations/basic/tensor_to_lie.hpp:65-103).
    Without ``matrix``: the library's Hall-basis T2L.  With ``matrix`` (CSR or
    CSC device arrays, e.g. the reference's make_t2l_matrix output): that
    matrix, view-relative, as the reference's operation takes it.
    ``lie_out`` holds true_size slots."""
    lie_out = lie_out if isinstance(lie_out, Batch) else Batch(lie_out, 1, basis.depth)
    tensor_in = _as_batch(tensor_in, basis)
    n = _n(num_batches, tensor_in)
    if matrix is None:
        d = A.LieConvertDesc(basis._c(), n, tensor_in._c(), lie_out._c())
        _call("rpp_hip_tensor_to_lie" + _sfx(lie_out, tensor_in), d, stream)
    else:
        d = A.LieMatrixConvertDesc(basis._c(), n, tensor_in._c(), lie_out._c(), matrix._c())
        _call("rpp_hip_tensor_to_lie_matrix" + _sfx(lie_out, tensor_in, matrix.values), d, stream)


def lie_to_tensor(tensor_out, lie_in, basis: TensorBasis, num_batches=None, stream=None,
                  matrix: Optional[SparseMatrix] = None):
    """rpp::ops::lie_to_tensor (operations/basic/lie_to_tensor.hpp:70-108);
    ``matrix`` as for tensor_to_lie (an L2T matrix)."""
    lie_in = lie_in if isinstance(lie_in, Batch) else Batch(lie_in, 1, basis.depth)
    tensor_out = _as_batch(tensor_out, basis)
    n = _n(num_batches, tensor_out)
    if matrix is None:
        d = A.LieConvertDesc(basis._c(), n, tensor_out._c(), lie_in._c())
        _call("rpp_hip_lie_to_tensor" + _sfx(tensor_out, lie_in), d, stream)
    else:
        d = A.LieMatrixConvertDesc(basis._c(), n, tensor_out._c(), lie_in._c(), matrix._c())
        _call("rpp_hip_lie_to_tensor_matrix" + _sfx(tensor_out, lie_in, matrix.values), d, stream)


def lie_basis(width: int, depth: int) -> dict:
    """Hall basis pairs, degree_begin and the CSC L2T/T2L matrices (host numpy)."""
    import numpy as np
    lib = A.lib()
    sizes = np.zeros(4, dtype=np.int64)
    rc = lib.rpp_hip_lie_basis_sizes(width, depth, sizes.ctypes.data)
    if rc != 0:
        raise RppHipError(rc, lib.rpp_hip_lie_basis_sizes.__name__)
    ts, l2t_nnz, t2l_nnz, dim = (int(v) for v in sizes)
    pairs = np.zeros(2 * ts, dtype=np.int64)
    db = np.zeros(depth + 2, dtype=np.int64)
    l2t = (np.zeros(l2t_nnz), np.zeros(l2t_nnz, dtype=np.int64), np.zeros(ts, dtype=np.int64))
    t2l = (np.zeros(t2l_nnz), np.zeros(t2l_nnz, dtype=np.int64), np.zeros(dim + 1, dtype=np.int64))
    rc = lib.rpp_hip_lie_basis_copy(width, depth, pairs.ctypes.data, db.ctypes.data,
                                    *(a.ctypes.data for a in l2t), *(a.ctypes.data for a in t2l))
    if rc != 0:
        raise RppHipError(rc, "rpp_hip_lie_basis_copy")
    return dict(true_size=ts, tensor_dim=dim, pairs=pairs.reshape(-1, 2), degree_begin=db,
                l2t=dict(data=l2t[0], indices=l2t[1], offsets=l2t[2]),
                t2l=dict(data=t2l[0], indices=t2l[1], offsets=t2l[2]))


def lie_basis_save(width: int, depth: int, path) -> None:
    """Write the Hall basis metadata of (width, depth) to ``path`` (the
    library's checksummed binary format)."""
    lib = A.lib()
    rc = lib.rpp_hip_lie_basis_save(width, depth, str(path).encode())
    if rc != 0:
        raise RppHipError(rc, lib.rpp_hip_last_error().decode())


def lie_basis_load(path) -> tuple:
    """Read a file written by :func:`lie_basis_save` and make it this process's
    metadata for its (width, depth); returns (width, depth)."""
    lib = A.lib()
    w, d = C.c_int32(0), C.c_int32(0)
    rc = lib.rpp_hip_lie_basis_load(str(path).encode(), C.byref(w), C.byref(d))
    if rc != 0:
        raise RppHipError(rc, lib.rpp_hip_last_error().decode())
    return w.value, d.value


# -------------------------------------------------- batched Chen recursion

def fmexp_tile(width: int, depth: int, dtype):
    """(kernel, threads, split) of the tile serving plain ft_fmexp at
    (width, depth): kernel "k_fmexp_coal" (subtree items, split leading
    letters) or "k_fmexp_wave" (sibling-group wavefronts, split 0); None when
    a generic kernel serves it."""
    nbytes = torch.tensor([], dtype=dtype).element_size()
    th, sp = C.c_int32(0), C.c_int32(0)
    kind = A.lib().rpp_hip_fmexp_tile(nbytes, width, depth, C.byref(th), C.byref(sp))
    if kind == 2:
        return "k_fmexp_coal", th.value, sp.value
    if kind == 3:
        return "k_fmexp_wave", th.value, sp.value
    return None


def signature_tiles(dtype) -> list:
    """Compiled tiles (width, depth, threads, split, kind, expand) for a dtype:
    kind 1 = two-barrier forward tile, kind 0 = Horner register tile (forward
    and, for split <= 1, backward), kind 2 = thread-owned subtree tile
    (forward, and backward where expand = 1)."""
    import numpy as np
    nbytes = torch.tensor([], dtype=dtype).element_size()
    lib = A.lib()
    n = lib.rpp_hip_signature_tiles(nbytes, None, 0)
    buf = np.zeros(6 * max(n, 1), dtype=np.int32)
    lib.rpp_hip_signature_tiles(nbytes, buf.ctypes.data, n)
    return [tuple(int(v) for v in buf[6 * i: 6 * i + 6]) for i in range(n)]


def thread_backward_tiles(width: int, depth: int, dtype) -> list:
    """(threads, split) of the thread-owned subtree tiles of (width, depth)
    that have a backward."""
    return [(t, s) for (w, d, t, s, k, e) in signature_tiles(dtype) if (w, d) == (width, depth) and k == 2 and e == 1]


def checkpoint_every_for(dtype, max_length: int) -> int:
    """The library's checkpoint spacing for a signature whose backward pass
    follows (rpp_hip_signature_checkpoint_every): 0 in FP64 and for FP32
    paths under 32 increments, else ceil(sqrt(max_length))."""
    nbytes = torch.tensor([], dtype=dtype).element_size()
    k = A.lib().rpp_hip_signature_checkpoint_every(nbytes, int(max_length))
    if k < 0:
        raise RppHipError(-k, A.lib().rpp_hip_last_error().decode())
    return k


def _offsets_of(lengths, device) -> torch.Tensor:
    lengths = torch.as_tensor(lengths, dtype=torch.int64)
    off = torch.zeros(lengths.numel() + 1, dtype=torch.int64)
    off[1:] = torch.cumsum(lengths, 0)
    return off.to(device)


def path_order(offsets: torch.Tensor) -> torch.Tensor:
    """Queue order of the batched Chen recursion computed on the device: path
    indices by decreasing length.  Passing it to signature/signature_backward
    removes their host synchronisation (graph-capturable calls)."""
    lengths = offsets[1:] - offsets[:-1]
    return torch.argsort(lengths, descending=True, stable=True).to(torch.int64)


def signature(increments: torch.Tensor, offsets: torch.Tensor, basis: TensorBasis,
              out: Optional[torch.Tensor] = None, checkpoint_every: int = 0,
              checkpoints: Optional[torch.Tensor] = None, ckpt_offsets: Optional[torch.Tensor] = None,
              tile_threads: int = 0, tile_split: int = 0, tile_kind: int = 0, tile_expand: int = 0,
              order: Optional[torch.Tensor] = None, stream=None) -> torch.Tensor:
    """Signatures of many paths of uneven length in one launch.

    increments: [total, width] device tensor; offsets: int64 device tensor [n+1]
    with offsets[n] = total; order: optional device queue order (path_order).
    """
    n = offsets.numel() - 1
    dim = basis.size()
    if tile_threads == 0:
        from . import autotune
        tile_threads, tile_split, tile_kind, tile_expand = (
            autotune.best(basis.width, basis.depth, increments.dtype, "forward") or (0, 0, 0, 0))
    if out is None:
        out = torch.empty((n, dim), dtype=increments.dtype, device=increments.device)
    inc = increments.contiguous()
    d = A.SignatureDesc(basis._c(), n, inc.data_ptr(), offsets.data_ptr(),
                        Batch(out, 0, basis.depth)._c(), checkpoint_every,
                        checkpoints.data_ptr() if checkpoints is not None else None,
                        ckpt_offsets.data_ptr() if ckpt_offsets is not None else None,
                        tile_threads, tile_split, tile_kind, tile_expand,
                        order.data_ptr() if order is not None else None, inc.shape[0])
    _call("rpp_hip_signature" + _sfx(inc, out), d, stream)
    return out


def signature_backward(increments: torch.Tensor, offsets: torch.Tensor, basis: TensorBasis,
                       signatures: torch.Tensor, cotangent: torch.Tensor,
                       grad: Optional[torch.Tensor] = None, checkpoint_every: int = 0,
                       checkpoints: Optional[torch.Tensor] = None,
                       ckpt_offsets: Optional[torch.Tensor] = None, tile_threads: int = 0,
                       tile_split: int = 0, order: Optional[torch.Tensor] = None, tile_expand: int = 0,
                       stream=None, tile_kind: int = 0) -> torch.Tensor:
    """d loss / d increments given d loss / d signatures (cotangent may be 1-D,
    one functional for every path).  tile_kind 0 selects a Horner register
    tile, 2 a thread-owned subtree tile (rpp_hip_signature_tiles)."""
    n = offsets.numel() - 1
    inc = increments.contiguous()
    if tile_threads == 0:
        from . import autotune
        best = autotune.best(basis.width, basis.depth, increments.dtype, "backward") or (0, 0, 0, 0)
        tile_threads, tile_split, tile_kind, tile_expand = best[0], best[1], best[2], best[3]
    if grad is None:
        # rows outside [offsets[0], offsets[n]) belong to no path: the library
        # leaves them untouched, so they start (and stay) at zero
        grad = torch.zeros_like(inc)
    d = A.SignatureBackwardDesc(basis._c(), n, inc.data_ptr(), offsets.data_ptr(),
                                Batch(signatures, 0, basis.depth)._c(),
                                Batch(cotangent, 0, basis.depth)._c(), grad.data_ptr(), checkpoint_every,
                                checkpoints.data_ptr() if checkpoints is not None else None,
                                ckpt_offsets.data_ptr() if ckpt_offsets is not None else None, None,
                                tile_threads, tile_split, order.data_ptr() if order is not None else None,
                                inc.shape[0], tile_expand, tile_kind)
    _call("rpp_hip_signature_backward" + _sfx(inc, signatures, cotangent), d, stream)
    return grad
