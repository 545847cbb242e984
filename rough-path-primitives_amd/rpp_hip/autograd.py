"""torch.autograd bindings over the C-ABI: the batched signature (forward and
backward kernels), ft_mul (backward through ft_adj_rmul / ft_adj_lmul,
operations/basic/ft_adj_rmul.hpp:47-71, ft_adj_lmul.hpp:49-73), ft_fmexp with
a degree-one exponent (the ft_fmexp adjoint kernel) and ft_log (the ft_log
adjoint kernel)."""
from __future__ import annotations

from typing import Optional

import torch

from . import ops


def _checkpoint_layout(offsets: torch.Tensor, every: int):
    lengths = (offsets[1:] - offsets[:-1]).cpu()
    counts = torch.div(lengths, every, rounding_mode="floor")
    ck_off = torch.zeros(counts.numel() + 1, dtype=torch.int64)
    ck_off[1:] = torch.cumsum(counts, 0)
    return ck_off.to(offsets.device), int(ck_off[-1])


class _Signature(torch.autograd.Function):
    @staticmethod
    def forward(ctx, increments, offsets, width: int, depth: int, checkpoint_every: int, tile):
        basis = ops.TensorBasis(width, depth)
        ckpts = ck_off = None
        if checkpoint_every > 0:
            ck_off, total = _checkpoint_layout(offsets, checkpoint_every)
            ckpts = torch.empty((max(total, 1), basis.size()), dtype=increments.dtype, device=increments.device)
        tile = tuple(tile) + (0, 0, 0, 0)
        sig = ops.signature(increments, offsets, basis, checkpoint_every=checkpoint_every, checkpoints=ckpts,
                            ckpt_offsets=ck_off, tile_threads=tile[0], tile_split=tile[1], tile_kind=tile[2],
                            tile_expand=tile[3])
        ctx.basis = basis
        ctx.every = checkpoint_every
        ctx.tile = tile
        ctx.save_for_backward(increments, offsets, sig, ckpts if ckpts is not None else sig,
                              ck_off if ck_off is not None else offsets)
        return sig

    @staticmethod
    def backward(ctx, grad_sig):
        increments, offsets, sig, ckpts, ck_off = ctx.saved_tensors
        kw = {}
        if ctx.every > 0:
            kw = dict(checkpoint_every=ctx.every, checkpoints=ckpts, ckpt_offsets=ck_off)
        t = ctx.tile
        th, sp, ex = (t[0], t[1], t[3]) if t[0] > 0 and t[1] <= 1 and t[2] == 0 else (0, 0, 0)
        grad = ops.signature_backward(increments, offsets, ctx.basis, sig, grad_sig.contiguous(),
                                      tile_threads=th, tile_split=sp, tile_expand=ex, **kw)
        return grad, None, None, None, None, None


def signature(increments: torch.Tensor, offsets: torch.Tensor, depth: int, checkpoint_every="auto",
              tile: Optional[tuple] = None) -> torch.Tensor:
    """Differentiable signatures of paths of uneven length.

    increments: [total, width]; offsets: int64 [n+1] on the same device.
    checkpoint_every > 0 stores the state every that many increments so the
    backward pass re-anchors its exp(-x) reconstruction there; 0 reconstructs
    every state; "auto" (the default) takes the library's choice for the
    dtype and the longest path (ops.checkpoint_every_for: sqrt(L) checkpoints
    in FP32 from 32 increments, where reconstruction drift doubles the
    gradient error; none in FP64).
    """
    if checkpoint_every == "auto":
        longest = int((offsets[1:] - offsets[:-1]).max().item()) if offsets.numel() > 1 else 0
        checkpoint_every = ops.checkpoint_every_for(increments.dtype, longest)
    return _Signature.apply(increments, offsets, increments.shape[1], depth, int(checkpoint_every), tile or (0, 0))


class _FTMul(torch.autograd.Function):
    @staticmethod
    def forward(ctx, lhs, rhs, width: int, depth: int):
        basis = ops.TensorBasis(width, depth)
        out = torch.empty_like(lhs)
        ops.ft_mul(out, lhs, rhs, basis)
        ctx.basis = basis
        ctx.save_for_backward(lhs, rhs)
        return out

    @staticmethod
    def backward(ctx, g):
        lhs, rhs = ctx.saved_tensors
        g = g.contiguous()
        gl = torch.empty_like(lhs)
        gr = torch.empty_like(rhs)
        # <g, lhs rhs>:  d/dlhs = R_rhs^* g,  d/drhs = L_lhs^* g
        ops.ft_adj_rmul(gl, rhs, g, ctx.basis)
        ops.ft_adj_lmul(gr, lhs, g, ctx.basis)
        return gl, gr, None, None


def ft_mul(lhs: torch.Tensor, rhs: torch.Tensor, width: int, depth: int) -> torch.Tensor:
    """Differentiable batched concatenation product (full windows)."""
    return _FTMul.apply(lhs.contiguous(), rhs.contiguous(), width, depth)


class _FTFmexp(torch.autograd.Function):
    @staticmethod
    def forward(ctx, mult, x, width: int, depth: int):
        basis = ops.TensorBasis(width, depth)
        n = mult.shape[0]
        ex = torch.zeros((n, 1 + width), dtype=mult.dtype, device=mult.device)
        ex[:, 1:] = x
        out = torch.empty_like(mult)
        ops.ft_fmexp(out, mult.contiguous(), ops.Batch(ex, 1, 1), basis)
        ctx.basis = basis
        ctx.save_for_backward(mult, ex)
        return out

    @staticmethod
    def backward(ctx, g):
        mult, ex = ctx.saved_tensors
        gm = torch.empty_like(mult)
        gx = torch.empty_like(ex)
        ops.ft_fmexp_adjoint(gm, ops.Batch(gx, 1, 1), mult.contiguous(), ops.Batch(ex, 1, 1), g.contiguous(),
                             ctx.basis)
        return gm, gx[:, 1:], None, None


def ft_fmexp(mult: torch.Tensor, x: torch.Tensor, depth: int) -> torch.Tensor:
    """Differentiable out = mult (x) exp(x) for a batch of tensors mult [n, dim]
    and degree-one exponents x [n, width]."""
    return _FTFmexp.apply(mult, x, x.shape[1], depth)


class _FTLog(torch.autograd.Function):
    @staticmethod
    def forward(ctx, arg, width: int, depth: int):
        basis = ops.TensorBasis(width, depth)
        out = torch.empty_like(arg)
        ops.ft_log(out, arg.contiguous(), basis)
        ctx.basis = basis
        ctx.save_for_backward(arg)
        return out

    @staticmethod
    def backward(ctx, g):
        (arg,) = ctx.saved_tensors
        ga = torch.empty_like(arg)
        ops.ft_log_adjoint(ga, arg.contiguous(), g.contiguous(), ctx.basis)
        return ga, None, None


def ft_log(arg: torch.Tensor, width: int, depth: int) -> torch.Tensor:
    """Differentiable ft_log over a batch [n, dim] (full windows)."""
    return _FTLog.apply(arg, width, depth)
