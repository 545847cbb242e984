"""ctypes wrappers for the oracle (liboracle.so) and, where it was built, the
real reference driver (_ref/librpp_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of bench.py, always as the checker or the CPU timing,
never as the product path.

Objects are numpy arrays; a batch is a 2-D array ``[n, stride]`` (row b is
object b, coefficient 0 first), or a 1-D array broadcast to every object
(stride 0), mirroring ``layouts::StrideLayout`` (views/batch.hpp:49-68).
Windows are ``(min_degree, max_degree)`` pairs as in
``DenseGradedVectorView`` (views/dense_graded_vector_view.hpp:17-128).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "librpp_ref.so")

_ll = C.c_longlong
_i = C.c_int
_p = C.c_void_p


def _load(path):
    if not os.path.exists(path):
        return None
    return C.CDLL(path)


_orc = None
_ref = None


def lib():
    global _orc
    if _orc is None:
        _orc = _load(ORACLE_SO)
        if _orc is None:
            raise RuntimeError(f"oracle library missing: {ORACLE_SO} (run oracle/Makefile)")
        _orc.orc_tensor_dimension.restype = _ll
    return _orc


def ref():
    """The reference driver, or None where it was never built."""
    global _ref
    if _ref is None:
        _ref = _load(REF_SO)
    return _ref


def has_ref() -> bool:
    return ref() is not None


def tensor_dimension(width: int, depth: int) -> int:
    return int(lib().orc_tensor_dimension(width, depth))


def degree_begin(width: int, depth: int) -> list[int]:
    out = [0]
    p = 1
    for _ in range(depth + 1):
        out.append(out[-1] + p)
        p *= width
    return out


def lie_degree_begin(width: int, depth: int) -> list[int]:
    """Degree offsets of the Hall-set Lie basis: one slot for degree 0, then
    the Witt numbers (1/k) sum_{d|k} mu(d) width^(k/d)."""
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
    out = [0, 1]
    for k in range(1, depth + 1):
        out.append(out[-1] + sum(mobius(d) * width ** (k // d) for d in range(1, k + 1) if k % d == 0) // k)
    return out[: depth + 2]


def _sfx(dtype):
    dt = np.dtype(dtype)
    if dt == np.float64:
        return "f64", C.c_double
    if dt == np.float32:
        return "f32", C.c_float
    raise TypeError(f"unsupported dtype {dt}")


def _arg(a):
    """(pointer, stride) of a batch array."""
    assert a.flags.c_contiguous
    stride = 0 if a.ndim == 1 else a.shape[1]
    return a.ctypes.data_as(_p), _ll(stride)


def _win(w):
    return _i(int(w[0])), _i(int(w[1]))


def _call(libobj, name, *args):
    fn = getattr(libobj, name)
    rc = fn(*args)
    if rc != 0:
        raise RuntimeError(f"{name} returned error code {rc}")


def _full(depth):
    return (0, depth)


# --------------------------------------------------------------------------
# Generic dispatch used by both the oracle ("orc_batch_*") and the reference
# driver ("ref_*"): same argument order, see oracle/rpp_oracle_impl.h.


class _Impl:
    def __init__(self, which: str):
        self.which = which

    def _lib(self):
        return lib() if self.which == "orc" else ref()

    def _name(self, op, sfx):
        return f"orc_batch_{op}_{sfx}" if self.which == "orc" else f"ref_{op}_{sfx}"

    def ft_mul(self, width, depth, out, lhs, rhs, ow=None, lw=None, rw=None, beta=1.0, n=None, threads=1):
        sfx, ct = _sfx(out.dtype)
        n = out.shape[0] if n is None else n
        args = [_i(width), _i(depth), _ll(n), *_arg(out), *_win(ow or _full(depth)),
                *_arg(lhs), *_win(lw or _full(depth)), *_arg(rhs), *_win(rw or _full(depth)), ct(beta)]
        if self.which == "ref":
            args.append(_i(threads))
        _call(self._lib(), self._name("ft_mul", sfx), *args)
        return out

    def ft_fma(self, width, depth, out, a, b, c, ow=None, aw=None, bw=None, cw=None, alpha=1.0, beta=1.0, n=None):
        sfx, ct = _sfx(out.dtype)
        n = out.shape[0] if n is None else n
        f = _full(depth)
        _call(self._lib(), self._name("ft_fma", sfx), _i(width), _i(depth), _ll(n),
              *_arg(out), *_win(ow or f), *_arg(a), *_win(aw or f), *_arg(b), *_win(bw or f),
              *_arg(c), *_win(cw or f), ct(alpha), ct(beta))
        return out

    def ft_inplace_fma(self, variant, width, depth, a, b, c, aw=None, bw=None, cw=None, alpha=1.0, beta=1.0, n=None):
        sfx, ct = _sfx(a.dtype)
        n = a.shape[0] if n is None else n
        f = _full(depth)
        _call(self._lib(), self._name("ft_inplace_fma", sfx), _i(variant), _i(width), _i(depth), _ll(n),
              *_arg(a), *_win(aw or f), *_arg(b), *_win(bw or f), *_arg(c), *_win(cw or f),
              ct(alpha), ct(beta))
        return a

    def ft_inplace_mul(self, width, depth, lhs, rhs, lw=None, rw=None, beta=1.0, n=None):
        sfx, ct = _sfx(lhs.dtype)
        n = lhs.shape[0] if n is None else n
        f = _full(depth)
        _call(self._lib(), self._name("ft_inplace_mul", sfx), _i(width), _i(depth), _ll(n),
              *_arg(lhs), *_win(lw or f), *_arg(rhs), *_win(rw or f), ct(beta))
        return lhs

    def _unary(self, op, width, depth, out, arg, ow, aw, n):
        sfx, _ = _sfx(out.dtype)
        n = out.shape[0] if n is None else n
        f = _full(depth)
        _call(self._lib(), self._name(op, sfx), _i(width), _i(depth), _ll(n),
              *_arg(out), *_win(ow or f), *_arg(arg), *_win(aw or f))
        return out

    def ft_exp(self, width, depth, out, arg, ow=None, aw=None, n=None):
        return self._unary("ft_exp", width, depth, out, arg, ow, aw, n)

    def ft_log(self, width, depth, out, arg, ow=None, aw=None, n=None):
        return self._unary("ft_log", width, depth, out, arg, ow, aw, n)

    def ft_fmexp(self, width, depth, out, mult, ex, ow=None, mw=None, ew=None, n=None):
        sfx, _ = _sfx(out.dtype)
        n = out.shape[0] if n is None else n
        f = _full(depth)
        _call(self._lib(), self._name("ft_fmexp", sfx), _i(width), _i(depth), _ll(n),
              *_arg(out), *_win(ow or f), *_arg(mult), *_win(mw or f), *_arg(ex), *_win(ew or f))
        return out

    def tensor_pairing(self, width, depth, out, f, x, fw=None, xw=None, n=None):
        sfx, _ = _sfx(out.dtype)
        n = out.shape[0] if n is None else n
        full = _full(depth)
        if self.which == "orc":
            _call(self._lib(), self._name("tensor_pairing", sfx), _i(width), _i(depth), _ll(n),
                  out.ctypes.data_as(_p), _ll(1), *_arg(f), *_win(fw or full), *_arg(x), *_win(xw or full))
        else:
            _call(self._lib(), self._name("tensor_pairing", sfx), _i(width), _i(depth), _ll(n),
                  out.ctypes.data_as(_p), *_arg(f), *_win(fw or full), *_arg(x), *_win(xw or full))
        return out

    def _adj(self, op_name, width, depth, out, op, arg, ow, pw, gw, n):
        sfx, _ = _sfx(out.dtype)
        n = out.shape[0] if n is None else n
        f = _full(depth)
        _call(self._lib(), self._name(op_name, sfx), _i(width), _i(depth), _ll(n),
              *_arg(out), *_win(ow or f), *_arg(op), *_win(pw or f), *_arg(arg), *_win(gw or f))
        return out

    def tensor_antipode(self, width, depth, out, arg, ow=None, aw=None, n=None, sign_by_degree=True):
        sfx, _ = _sfx(out.dtype)
        n = out.shape[0] if n is None else n
        f = _full(depth)
        name = ("orc_batch_tensor_antipode_" if self.which == "orc" else "ref_tensor_antipode_") + sfx
        _call(self._lib(), name, _i(1 if sign_by_degree else 0), _i(width), _i(depth), _ll(n),
              *_arg(out), *_win(ow or f), *_arg(arg), *_win(aw or f))
        return out

    def tensor_reflect(self, width, depth, out, arg, ow=None, aw=None, n=None):
        return self.tensor_antipode(width, depth, out, arg, ow, aw, n, sign_by_degree=False)

    def ft_adj_lmul(self, width, depth, out, op, arg, ow=None, pw=None, gw=None, n=None):
        return self._adj("ft_adj_lmul", width, depth, out, op, arg, ow, pw, gw, n)

    def ft_adj_rmul(self, width, depth, out, op, arg, ow=None, pw=None, gw=None, n=None):
        return self._adj("ft_adj_rmul", width, depth, out, op, arg, ow, pw, gw, n)

    # Shuffle family.  kind 0 st_mul(out, b, c), 1 st_fma(out, a, b, c),
    # 2 st_inplace_fma(a, b, c) with out == a.
    def _st(self, kind, width, depth, out, a, b, c, ow, aw, bw, cw, alpha, beta, n):
        sfx, ct = _sfx(out.dtype)
        n = out.shape[0] if n is None else n
        f = _full(depth)
        a = out if a is None else a
        _call(self._lib(), self._name("st_fma", sfx), _i(kind), _i(width), _i(depth), _ll(n),
              *_arg(out), *_win(ow or f), *_arg(a), *_win(aw or f), *_arg(b), *_win(bw or f),
              *_arg(c), *_win(cw or f), ct(alpha), ct(beta))
        return out

    def st_mul(self, width, depth, out, lhs, rhs, ow=None, lw=None, rw=None, beta=1.0, n=None):
        return self._st(0, width, depth, out, None, lhs, rhs, ow, None, lw, rw, 0.0, beta, n)

    def st_fma(self, width, depth, out, a, b, c, ow=None, aw=None, bw=None, cw=None, alpha=1.0, beta=1.0, n=None):
        return self._st(1, width, depth, out, a, b, c, ow, aw, bw, cw, alpha, beta, n)

    def st_inplace_fma(self, width, depth, a, b, c, aw=None, bw=None, cw=None, alpha=1.0, beta=1.0, n=None):
        return self._st(2, width, depth, a, None, b, c, aw, aw, bw, cw, alpha, beta, n)

    # opcode 0 st_adj_mul(out, op, arg), 1 left_hs_mul(out, integrator,
    # integrand, beta), 2 left_hs_adj_lmul(out, integrator, integrand),
    # 3 left_hs_adj_rmul(out, integrator, integrand)
    def _sh2(self, opcode, width, depth, out, p, q, ow, pw, qw, beta, n):
        sfx, ct = _sfx(out.dtype)
        n = out.shape[0] if n is None else n
        f = _full(depth)
        _call(self._lib(), self._name("shuffle2", sfx), _i(opcode), _i(width), _i(depth), _ll(n),
              *_arg(out), *_win(ow or f), *_arg(p), *_win(pw or f), *_arg(q), *_win(qw or f), ct(beta))
        return out

    def st_adj_mul(self, width, depth, out, op, arg, ow=None, pw=None, gw=None, n=None):
        return self._sh2(0, width, depth, out, op, arg, ow, pw, gw, 1.0, n)

    def left_hs_mul(self, width, depth, out, integrator, integrand, ow=None, xw=None, yw=None, beta=1.0, n=None):
        return self._sh2(1, width, depth, out, integrator, integrand, ow, xw, yw, beta, n)

    def left_hs_adj_lmul(self, width, depth, out, integrator, integrand, ow=None, xw=None, yw=None, beta=1.0,
                         n=None):
        return self._sh2(2, width, depth, out, integrator, integrand, ow, xw, yw, beta, n)

    def left_hs_adj_rmul(self, width, depth, out, integrator, integrand, ow=None, xw=None, yw=None, beta=1.0,
                         n=None):
        return self._sh2(3, width, depth, out, integrator, integrand, ow, xw, yw, beta, n)

    # Linear operations over any grading: `start` = degree offsets
    # start_of_degree(k), k = 0..depth+1 (degree_begin or lie_degree_begin).

    @staticmethod
    def _start(start):
        s = np.ascontiguousarray(start, dtype=np.int64)
        return s, s.ctypes.data_as(_p)

    def vector_assign(self, start, depth, out, arg, ow=None, aw=None, n=None):
        sfx, _ = _sfx(out.dtype)
        n = out.shape[0] if n is None else n
        f = _full(depth)
        keep, sp = self._start(start)
        _call(self._lib(), self._name("vector_assign", sfx), sp, _i(depth), _ll(n),
              *_arg(out), *_win(ow or f), *_arg(arg), *_win(aw or f))
        return out

    def vector_add(self, start, depth, out, lhs, rhs, ow=None, lw=None, rw=None, alpha=1.0, beta=1.0, n=None):
        sfx, ct = _sfx(out.dtype)
        n = out.shape[0] if n is None else n
        f = _full(depth)
        keep, sp = self._start(start)
        _call(self._lib(), self._name("vector_add", sfx), sp, _i(depth), _ll(n), *_arg(out), *_win(ow or f),
              *_arg(lhs), *_win(lw or f), *_arg(rhs), *_win(rw or f), ct(alpha), ct(beta))
        return out

    def vector_inplace_add(self, start, depth, lhs, rhs, lw=None, rw=None, alpha=1.0, n=None):
        sfx, ct = _sfx(lhs.dtype)
        n = lhs.shape[0] if n is None else n
        f = _full(depth)
        keep, sp = self._start(start)
        _call(self._lib(), self._name("vector_inplace_add", sfx), sp, _i(depth), _ll(n), *_arg(lhs),
              *_win(lw or f), *_arg(rhs), *_win(rw or f), ct(alpha))
        return lhs

    def _vector_scalar(self, op, start, depth, vec, vw, scalar, n):
        sfx, ct = _sfx(vec.dtype)
        n = vec.shape[0] if n is None else n
        keep, sp = self._start(start)
        _call(self._lib(), self._name("vector_scalar", sfx), _i(op), sp, _i(depth), _ll(n), *_arg(vec),
              *_win(vw or _full(depth)), ct(scalar))
        return vec

    def vector_scalar_multiply(self, start, depth, vec, scalar, vw=None, n=None):
        return self._vector_scalar(0, start, depth, vec, vw, scalar, n)

    def vector_set_constant(self, start, depth, vec, value, vw=None, n=None):
        return self._vector_scalar(1, start, depth, vec, vw, value, n)

    def tensor_set_identity(self, start, depth, vec, scalar=1.0, vw=None, n=None):
        return self._vector_scalar(2, start, depth, vec, vw, scalar, n)

    def tensor_add_identity(self, start, depth, vec, scalar=1.0, vw=None, n=None):
        return self._vector_scalar(3, start, depth, vec, vw, scalar, n)


orc = _Impl("orc")
refimpl = _Impl("ref")


# --------------------------------------------------------------------------
# Oracle-only helpers


def spmv(values, inner, offsets, outer, out, arg, *, csc, out_offset=0, out_size=None,
         arg_offset=0, alpha=1.0):
    """Batched SpMV with view-relative indexing (row 0 <-> out.begin())."""
    sfx, ct = _sfx(out.dtype)
    values = np.ascontiguousarray(values, dtype=out.dtype)
    inner = np.ascontiguousarray(inner, dtype=np.int64)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    n = out.shape[0]
    if out_size is None:
        out_size = out.shape[1] - out_offset
    _call(lib(), f"orc_batch_spmv_{sfx}", _i(1 if csc else 0), _ll(n), *_arg(out), _ll(out_offset),
          _ll(out_size), *_arg(arg), _ll(arg_offset), values.ctypes.data_as(_p),
          inner.ctypes.data_as(_p), offsets.ctypes.data_as(_p), _ll(outer), ct(alpha))
    return out


def signature(width, depth, increments, offsets, dtype=np.float64):
    """Chen recursion by chained oracle ft_fmexp; increments [sum(len), width]."""
    sfx, _ = _sfx(dtype)
    inc = np.ascontiguousarray(increments, dtype=dtype)
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    n = off.shape[0] - 1
    dim = tensor_dimension(width, depth)
    out = np.zeros((n, dim), dtype=dtype)
    _call(lib(), f"orc_signature_{sfx}", _i(width), _i(depth), _ll(n), off.ctypes.data_as(_p),
          inc.ctypes.data_as(_p), out.ctypes.data_as(_p), _ll(dim))
    return out


def signature_threads(width, depth, increments, offsets, threads, dtype=np.float64):
    """signature() with the paths split into `threads` contiguous slices, each
    computed on its own Python thread (the C call releases the GIL): the
    restated single-thread strategy run on several cores, as the reference
    strategy is by splitting its batch."""
    from concurrent.futures import ThreadPoolExecutor
    inc = np.ascontiguousarray(increments, dtype=dtype)
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    n = off.shape[0] - 1
    dim = tensor_dimension(width, depth)
    out = np.zeros((n, dim), dtype=dtype)
    bounds = np.linspace(0, n, max(1, threads) + 1).astype(np.int64)
    sfx, _ = _sfx(dtype)

    def work(a, b):
        if b <= a:
            return
        sub = np.ascontiguousarray(off[a:b + 1] - off[a])
        part = inc[off[a]:off[b]]
        view = out[a:b]
        _call(lib(), f"orc_signature_{sfx}", _i(width), _i(depth), _ll(b - a), sub.ctypes.data_as(_p),
              part.ctypes.data_as(_p), view.ctypes.data_as(_p), _ll(dim))

    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        list(ex.map(lambda i: work(int(bounds[i]), int(bounds[i + 1])), range(len(bounds) - 1)))
    return out


def signature_vjp(width, depth, increments, offsets, ell):
    """Gradient of <ell, S_p> wrt increments, composed from the reference
    adjoints (FP64)."""
    inc = np.ascontiguousarray(increments, dtype=np.float64)
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    ell = np.ascontiguousarray(ell, dtype=np.float64)
    grad = np.zeros_like(inc)
    _call(lib(), "orc_signature_vjp_f64", _i(width), _i(depth), _ll(off.shape[0] - 1),
          off.ctypes.data_as(_p), inc.ctypes.data_as(_p), ell.ctypes.data_as(_p),
          grad.ctypes.data_as(_p))
    return grad


def fmexp_vjp(width, depth, mult, x, cot):
    """Adjoint of one ft_fmexp with a degree-one exponent, composed from the
    reference adjoints (FP64): mult, cot [n, dim]; x [n, width].  Returns
    (grad_mult [n, dim], grad_x [n, width])."""
    mult = np.ascontiguousarray(mult, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    cot = np.ascontiguousarray(cot, dtype=np.float64)
    n = mult.shape[0]
    gm = np.zeros_like(mult)
    gx = np.zeros_like(x)
    _call(lib(), "orc_fmexp_vjp_f64", _i(width), _i(depth), _ll(n), mult.ctypes.data_as(_p),
          x.ctypes.data_as(_p), cot.ctypes.data_as(_p), gm.ctypes.data_as(_p), gx.ctypes.data_as(_p))
    return gm, gx


def log_vjp(width, depth, arg, cot):
    """Gradient of ft_log (full output window) with respect to arg, summed term
    by term from the log series with the reference adjoints (FP64):
    arg, cot [n, dim] -> grad [n, dim] (degree 0 is 0)."""
    arg = np.ascontiguousarray(arg, dtype=np.float64)
    cot = np.ascontiguousarray(cot, dtype=np.float64)
    grad = np.zeros_like(arg)
    _call(lib(), "orc_log_vjp_f64", _i(width), _i(depth), _ll(arg.shape[0]), arg.ctypes.data_as(_p),
          cot.ctypes.data_as(_p), grad.ctypes.data_as(_p))
    return grad


def ref_signature(width, depth, increments, threads=1):
    """Reference Chen recursion (one ft_fmexp per increment over the batch);
    increments [n, L, width]."""
    r = ref()
    if r is None:
        raise RuntimeError("reference driver not built")
    inc = np.ascontiguousarray(increments)
    sfx, _ = _sfx(inc.dtype)
    n, L, w = inc.shape
    assert w == width
    dim = tensor_dimension(width, depth)
    out = np.zeros((n, dim), dtype=inc.dtype)
    _call(r, f"ref_signature_{sfx}", _i(width), _i(depth), _ll(n), _ll(L), inc.ctypes.data_as(_p),
          out.ctypes.data_as(_p), _i(threads))
    return out


def ref_lie(width, depth):
    """Hall basis pairs, degree_begin and the CSC L2T/T2L matrices as built by
    the reference (dict of numpy arrays)."""
    r = ref()
    if r is None:
        raise RuntimeError("reference driver not built")
    r.ref_lie_build.restype = _p
    h = r.ref_lie_build(_i(width), _i(depth))
    try:
        sz = np.zeros(7, dtype=np.int64)
        r.ref_lie_sizes(_p(h), sz.ctypes.data_as(_p))
        ts, l2t_nnz, l2t_outer, l2t_inner, t2l_nnz, t2l_outer, t2l_inner = (int(v) for v in sz)
        pairs = np.zeros(2 * ts, dtype=np.int64)
        db = np.zeros(depth + 2, dtype=np.int64)
        l2t_d = np.zeros(l2t_nnz)
        l2t_i = np.zeros(l2t_nnz, dtype=np.int64)
        l2t_o = np.zeros(l2t_outer + 1, dtype=np.int64)
        t2l_d = np.zeros(t2l_nnz)
        t2l_i = np.zeros(t2l_nnz, dtype=np.int64)
        t2l_o = np.zeros(t2l_outer + 1, dtype=np.int64)
        r.ref_lie_copy(_p(h), *(a.ctypes.data_as(_p) for a in
                                (pairs, db, l2t_d, l2t_i, l2t_o, t2l_d, t2l_i, t2l_o)))
    finally:
        r.ref_lie_free(_p(h))
    return dict(pairs=pairs.reshape(-1, 2), degree_begin=db,
                l2t=dict(data=l2t_d, indices=l2t_i, offsets=l2t_o, outer=l2t_outer, inner=l2t_inner),
                t2l=dict(data=t2l_d, indices=t2l_i, offsets=t2l_o, outer=t2l_outer, inner=t2l_inner))


def ref_bracket(width, depth, left, right):
    r = ref()
    keys = np.zeros(4096, dtype=np.int64)
    coeffs = np.zeros(4096, dtype=np.int64)
    cnt = r.ref_lie_bracket(_i(width), _i(depth), _ll(left), _ll(right), keys.ctypes.data_as(_p),
                            coeffs.ctypes.data_as(_p), _i(4096))
    return {int(k): int(c) for k, c in zip(keys[:cnt], coeffs[:cnt])}


def ref_lie_convert(to_lie, csr, width, depth, inp):
    """tensor_to_lie (to_lie=True) or lie_to_tensor through the reference ops;
    Lie objects hold true_size slots (slot 0 unused)."""
    r = ref()
    inp = np.ascontiguousarray(inp, dtype=np.float64)
    n = inp.shape[0]
    dim = tensor_dimension(width, depth)
    lie = ref_lie(width, depth)
    lsize = lie["pairs"].shape[0]
    out = np.full((n, lsize if to_lie else dim), np.nan)
    if to_lie:
        out[:, 0] = 0.0
    _call(r, "ref_lie_convert_f64", _i(1 if to_lie else 0), _i(1 if csr else 0), _i(width),
          _i(depth), _ll(n), out.ctypes.data_as(_p), _ll(out.shape[1]), inp.ctypes.data_as(_p),
          _ll(inp.shape[1]))
    return out
