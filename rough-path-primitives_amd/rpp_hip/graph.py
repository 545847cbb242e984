"""The batched Chen recursion captured once in a hipGraph.

The queue order is computed on the device (ops.path_order) and handed to the
library, so a signature call issues kernel launches only: a counter reset
kernel and the persistent kernel (the backward adds the partial-sum kernel).
One uncaptured call first sizes the library workspace; the capture then
records those launches over static buffers, and replay() re-runs them on what
the static increment and cotangent buffers hold at that time.  The buffers'
shapes (and so the path lengths) are fixed at construction.
"""
from __future__ import annotations

from typing import Optional

import torch

from . import ops


class SignatureGraph:
    """Forward, and with ``backward=True`` also the backward pass, of
    ops.signature over fixed offsets, replayed from one hipGraph.

    increments: [total, width] device tensor, copied into ``self.increments``
    (write new data there between replays); cotangent: [n, dim] or one [dim]
    functional, copied into ``self.cotangent``.  checkpoint_every: as for
    autograd.signature; "auto" (the default) stores sqrt(L) checkpoints only
    when a backward pass follows in FP32 (ops.checkpoint_every_for).
    """

    def __init__(self, increments: torch.Tensor, offsets: torch.Tensor, basis: ops.TensorBasis,
                 backward: bool = False, cotangent: Optional[torch.Tensor] = None, checkpoint_every="auto",
                 tile: Optional[tuple] = None):
        dev = increments.device
        self.basis = basis
        self.offsets = offsets.to(dev, torch.int64).contiguous()
        if checkpoint_every == "auto":
            longest = int((self.offsets[1:] - self.offsets[:-1]).max().item()) if self.offsets.numel() > 1 else 0
            checkpoint_every = ops.checkpoint_every_for(increments.dtype, longest) if backward else 0
        self.increments = increments.detach().clone().contiguous()
        n = self.offsets.numel() - 1
        self.out = torch.empty((n, basis.size()), dtype=increments.dtype, device=dev)
        self.order = ops.path_order(self.offsets)
        self.tile = tuple(tile) if tile else (0, 0, 0, 0)
        self.every = checkpoint_every
        self.ckpts = self.ck_off = None
        if checkpoint_every > 0:
            lengths = (self.offsets[1:] - self.offsets[:-1]).cpu()
            counts = torch.div(lengths, checkpoint_every, rounding_mode="floor")
            ck_off = torch.zeros(n + 1, dtype=torch.int64)
            ck_off[1:] = torch.cumsum(counts, 0)
            self.ck_off = ck_off.to(dev)
            self.ckpts = torch.empty((max(int(ck_off[-1]), 1), basis.size()), dtype=increments.dtype, device=dev)
        self.backward = backward
        self.grad = None
        self.cotangent = None
        if backward:
            if cotangent is None:
                raise ValueError("backward=True needs a cotangent")
            self.cotangent = cotangent.detach().clone().contiguous()
            self.grad = torch.empty_like(self.increments)
        # uncaptured call: sizes the library workspace (no allocation may
        # happen under capture)
        self._launch()
        torch.cuda.synchronize(dev)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._launch()
        torch.cuda.synchronize(dev)

    def _launch(self):
        kw = {}
        if self.every > 0:
            kw = dict(checkpoint_every=self.every, checkpoints=self.ckpts, ckpt_offsets=self.ck_off)
        th, sp, kind, ex = (self.tile + (0, 0, 0, 0))[:4]
        ops.signature(self.increments, self.offsets, self.basis, out=self.out, tile_threads=th, tile_split=sp,
                      tile_kind=kind, tile_expand=ex, order=self.order, **kw)
        if self.backward:
            bth, bsp, bex = (th, sp, ex) if kind == 0 and sp <= 1 and th > 0 else (0, 0, 0)
            ops.signature_backward(self.increments, self.offsets, self.basis, self.out, self.cotangent,
                                   grad=self.grad, tile_threads=bth, tile_split=bsp, tile_expand=bex,
                                   order=self.order, **kw)

    def replay(self):
        """Re-run the captured launches; returns out (and grad with backward)."""
        self.graph.replay()
        return (self.out, self.grad) if self.backward else self.out
