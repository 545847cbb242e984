"""Tile autotuner of the batched Chen recursion.

For a (width, depth, dtype) it times every compiled register tile
(rpp_hip_signature_tiles) and the storage-generic kernel on a synthetic batch
of the caller's shape, forward and backward, and records the fastest in a JSON
table (autotune.json beside this file, or the path in $RPP_HIP_AUTOTUNE).
ops.signature / ops.signature_backward consult the table when no tile is
named; without an entry the library's default tile is used.

    python -m rpp_hip.autotune --width 5 --depth 6 --dtype f64 --paths 4096 --length 128
"""
from __future__ import annotations

import argparse
import json
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.environ.get("RPP_HIP_AUTOTUNE", os.path.join(_HERE, "autotune.json"))

_cache = None


def _key(width, depth, dtype) -> str:
    return f"{width},{depth},{'f64' if dtype == torch.float64 else 'f32'}"


def load() -> dict:
    global _cache
    if _cache is None:
        try:
            with open(TABLE) as f:
                _cache = json.load(f)
        except (OSError, ValueError):
            _cache = {}
    return _cache


def best(width: int, depth: int, dtype, kind: str):
    """(threads, split, kind, expand) recorded for 'forward' or 'backward', or None."""
    e = load().get(_key(width, depth, dtype), {}).get(kind)
    return tuple(e) if e else None


def _time(fn, reps):
    """Seconds per call: events on the current stream around reps calls
    after one warm-up call, so host work outside the launches is excluded."""
    fn()
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(reps):
        fn()
    end.record()
    end.synchronize()
    return start.elapsed_time(end) * 1e-3 / reps


def tune(width: int, depth: int, dtype=torch.float64, paths: int = 1024, length: int = 64, reps: int = 5,
         backward: bool = True, write: bool = True) -> dict:
    from . import ops
    dev = torch.device("cuda:0")
    basis = ops.TensorBasis(width, depth)
    g = torch.Generator(device="cpu").manual_seed(1234)
    inc = (torch.randn(paths * length, width, generator=g, dtype=torch.float64) / length ** 0.5).to(dev, dtype)
    off = torch.arange(paths + 1, dtype=torch.int64, device=dev) * length
    # the queue order on the device, so no call synchronises with the host
    order = ops.path_order(off)
    cands = [(t, s, k, e) for (w, n, t, s, k, e) in ops.signature_tiles(dtype) if (w, n) == (width, depth)]
    cands.append((-1, 0, 0, 0))  # the storage-generic kernel
    thread_bwd = ops.thread_backward_tiles(width, depth, dtype)
    result = {"forward": {}, "backward": {}}
    sig = ops.signature(inc, off, basis)
    ell = torch.randn(basis.size(), generator=g, dtype=torch.float64).to(dev, dtype)
    for t, s, k, e in cands:
        key = f"{t},{s},{k},{e}"
        result["forward"][key] = _time(
            lambda: ops.signature(inc, off, basis, out=sig, tile_threads=t, tile_split=s, tile_kind=k,
                                  tile_expand=e, order=order), reps)
        if backward and ((k == 0 and s <= 1) or (k == 2 and (t, s) in thread_bwd)):
            result["backward"][key] = _time(
                lambda: ops.signature_backward(inc, off, basis, sig, ell, tile_threads=t, tile_split=s,
                                               tile_expand=e, tile_kind=k, order=order), reps)
    entry = {}
    for kind in ("forward", "backward"):
        if result[kind]:
            k = min(result[kind], key=result[kind].get)
            entry[kind] = [int(v) for v in k.split(",")]
    entry["seconds"] = result
    entry["shape"] = {"paths": paths, "length": length}
    if write:
        table = dict(load())
        table[_key(width, depth, dtype)] = entry
        with open(TABLE, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        global _cache
        _cache = table
    return entry


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--width", type=int, required=True)
    ap.add_argument("--depth", type=int, required=True)
    ap.add_argument("--dtype", choices=["f32", "f64"], default="f64")
    ap.add_argument("--paths", type=int, default=1024)
    ap.add_argument("--length", type=int, default=64)
    ap.add_argument("--no-backward", action="store_true")
    a = ap.parse_args()
    e = tune(a.width, a.depth, torch.float64 if a.dtype == "f64" else torch.float32, a.paths, a.length,
             backward=not a.no_backward)
    print(json.dumps(e))


if __name__ == "__main__":
    main()
