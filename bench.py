"""Benchmark of the truncated tensor algebra hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): signatures of
4096 piecewise-linear paths, width 5, depth 6 (19531 coefficients), 128
increments per path drawn iid N(0, 1/128) per channel, by the chained
ft_fmexp of the reference's usage pattern (one call per increment,
ping-pong buffers, exponent view (1,1)); FP64.  One step is the whole chain
(128 launches) over the batch, inputs resident in HBM.

value = path-increments per second of that plain per-operation path.  The
same line carries the fused path (all increments in one persistent launch),
FP32 variants, the error of each against the FP64 oracle, the roofline of
the dominant kernel (k_fmexp_tile, HBM-bound) and the reference CPU path
(oracle/_ref, the real cpu::single_thread) timed on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rough-path-primitives_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

WIDTH, DEPTH, PATHS, LENGTH = 5, 6, 4096, 128
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3


def dim_of(w, n):
    return (w ** (n + 1) - 1) // (w - 1)


def make_increments(seed, paths, length, width):
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, np.sqrt(1.0 / length), size=(paths, length, width))


class Plain:
    """Chained ft_fmexp, one launch per increment (the reference usage)."""

    def __init__(self, rpp_hip, torch, inc, dtype, dev):
        self.rpp, self.torch = rpp_hip, torch
        n, L, w = inc.shape
        self.n, self.L, self.w = n, L, w
        self.basis = rpp_hip.TensorBasis(w, DEPTH)
        dim = self.basis.size()
        padded = np.zeros((n, 1 + L * w))
        padded[:, 1:] = inc.reshape(n, L * w)
        self.pad = torch.tensor(padded, dtype=dtype, device=dev)
        self.a = torch.zeros((n, dim), dtype=dtype, device=dev)
        self.b = torch.empty_like(self.a)
        self.ex = [rpp_hip.Batch(self.pad[:, t * w: t * w + 1 + w], 1, 1) for t in range(L)]

    def step(self):
        self.a.zero_()
        self.a[:, 0] = 1.0
        cur, nxt = self.a, self.b
        for t in range(self.L):
            self.rpp.ft_fmexp(nxt, cur, self.ex[t], self.basis)
            cur, nxt = nxt, cur
        return cur


class Fused:
    """All increments of every path in one persistent launch."""

    def __init__(self, rpp_hip, torch, inc, dtype, dev):
        self.rpp = rpp_hip
        n, L, w = inc.shape
        self.basis = rpp_hip.TensorBasis(w, DEPTH)
        self.inc = torch.tensor(inc.reshape(n * L, w), dtype=dtype, device=dev)
        self.off = torch.arange(n + 1, dtype=torch.int64, device=dev) * L
        self.out = torch.empty((n, self.basis.size()), dtype=dtype, device=dev)

    def step(self):
        return self.rpp.signature(self.inc, self.off, self.basis, out=self.out)


def time_steps(torch, runner, steps, warmup):
    for _ in range(warmup):
        runner.step()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record()
    for _ in range(steps):
        out = runner.step()
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return start.elapsed_time(stop) / steps, wall / steps, out


def rel_err(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.maximum(np.abs(a), np.abs(b)))))


def cpu_baseline(inc):
    """The reference cpu::single_thread chain (oracle/_ref), its batch split
    over std::threads, on a bounded sample of the same workload."""
    import oracle as O
    threads = max(1, min(16, os.cpu_count() or 1))
    if O.has_ref():
        sample = inc[: 48 * threads]
        kind = "reference"
        run = lambda: O.ref_signature(WIDTH, DEPTH, sample, threads=threads)  # noqa: E731
    else:  # the plain-C restatement, one thread
        sample = inc[:4]
        threads = 1
        kind = "port"
        flat = sample.reshape(-1, WIDTH)
        off = np.arange(sample.shape[0] + 1, dtype=np.int64) * LENGTH
        run = lambda: O.signature(WIDTH, DEPTH, flat, off)  # noqa: E731
    t0 = time.perf_counter()
    run()
    dt = time.perf_counter() - t0
    units = sample.shape[0] * sample.shape[1]
    return {"value": units / dt, "unit": "path-increments/s", "cores": threads, "kind": kind,
            "sample": f"{sample.shape[0]} paths x {sample.shape[1]} increments of the same workload "
                      f"(width 5, depth 6, FP64), {dt:.2f} s"}


def load_traffic():
    p = os.path.join(ROOT, "profiles", "traffic_k_fmexp_tile_c2_fp64.json")
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", default=None,
                    help="run one variant (plain_fp64, fused_fp64, plain_fp32, fused_fp32) and print its timing")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(paths are independent; see DESIGN.md for how the batch shards)", file=sys.stderr)
        sys.exit(2)

    import torch

    import oracle as O
    import rpp_hip

    dev = torch.device("cuda:0")
    inc = make_increments(0xC2, PATHS, LENGTH, WIDTH)
    dim = dim_of(WIDTH, DEPTH)
    units = PATHS * LENGTH

    # FP64 oracle on a few paths for the error statements
    n_chk = 8
    ref64 = O.signature(WIDTH, DEPTH, inc[:n_chk].reshape(-1, WIDTH),
                        np.arange(n_chk + 1, dtype=np.int64) * LENGTH)

    if args.copy_mib:
        src = torch.ones(args.copy_mib * (1 << 20) // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    variants = (("plain_fp64", Plain, torch.float64), ("fused_fp64", Fused, torch.float64),
                ("plain_fp32", Plain, torch.float32), ("fused_fp32", Fused, torch.float32))
    if args.only:
        variants = [v for v in variants if v[0] == args.only]
        name, cls, dt = variants[0]
        runner = cls(rpp_hip, torch, inc, dt, dev)
        ms_ev, _, _ = time_steps(torch, runner, args.steps, args.warmup)
        print(json.dumps({"variant": name, "ms_per_step": ms_ev, "path_increments_per_s": units / (ms_ev * 1e-3)}))
        return

    results = {}
    for name, cls, dt in variants:
        runner = cls(rpp_hip, torch, inc, dt, dev)
        ms_ev, s_wall, out = time_steps(torch, runner, args.steps, args.warmup)
        err = rel_err(out[:n_chk].double().cpu().numpy(), ref64)
        results[name] = {"ms_per_step": ms_ev, "wall_ms_per_step": s_wall * 1e3,
                         "path_increments_per_s": units / (ms_ev * 1e-3), "max_rel_err_vs_fp64_oracle": err}
        del runner
        torch.cuda.empty_cache()

    plain = results["plain_fp64"]
    # roofline of the dominant kernel: one k_fmexp_deg1 launch per increment
    launch_ms = plain["ms_per_step"] / LENGTH
    alg_bytes = PATHS * (2 * dim + WIDTH) * 8
    achieved = alg_bytes / (launch_ms * 1e-3) / 1e9
    fused = results["fused_fp64"]
    fused_flops = PATHS * LENGTH * 2 * sum((WIDTH ** (k + 1) - WIDTH) // (WIDTH - 1) for k in range(1, DEPTH + 1))
    line = {
        "metric": "ft_fmexp increments/s + fwd+bwd paths/s, fused vs plain GPU, err vs FP64",
        "value": plain["path_increments_per_s"],
        "unit": "path-increments/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": plain["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: increments iid N(0, 1/128) per channel, seed 0xC2",
        "config": {"workload": "plain signature via chained ft_fmexp (one launch per increment)",
                   "paths": PATHS, "width": WIDTH, "depth": DEPTH, "coefficients": dim,
                   "increments_per_path": LENGTH},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": load_traffic(),
                     "kernel": "k_fmexp_tile<double,5,6,640,1>", "avg_launch_ms": launch_ms,
                     "algorithmic_bytes_per_launch": alg_bytes},
        "max_rel_err_vs_fp64_oracle": plain["max_rel_err_vs_fp64_oracle"],
        "fused_vs_plain": {
            name: r for name, r in results.items()
        },
        "fused_speedup_fp64": fused["path_increments_per_s"] / plain["path_increments_per_s"],
        "fused_fp64_vector_tflops": fused_flops / (fused["ms_per_step"] * 1e-3) / 1e12,
        "fused_fp64_vector_frac": fused_flops / (fused["ms_per_step"] * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(inc)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
