"""Benchmark of the truncated tensor algebra hot path on one MI355X.

Headline workload (the configuration the headline metric is quoted on):
signatures of 4096 piecewise-linear paths, width 5, depth 6 (19531
coefficients), 128 increments per path drawn iid N(0, 1/128) per channel, by
the chained ft_fmexp of the reference's usage pattern (one call per increment,
ping-pong buffers, exponent view (1,1)); FP64.  One step is the whole chain
(128 launches) over the batch, inputs resident in HBM.

value = path-increments per second of that plain per-operation path.  The
same line carries, each under its own key:
  plain_vs_fused  the headline shape through the fused persistent kernel, and
                  FP32, with the error of each against the FP64 oracle;
  fwd_bwd         forward + backward of <ell, S(path)>: 2048 paths, width 6,
                  depth 6, 256 increments N(0, 1/256): the plain per-op path
                  (ft_fmexp chain, then per increment ft_fmexp(S, -x) to
                  reconstruct and the ft_fmexp adjoint) against the fused
                  kernels (exp(-x) reconstruction, and sqrt(L) checkpoints);
  ft_mul          1024 pairs of width 4 depth 4 tensors (GPU and reference
                  CPU) and 2^17 pairs for the HBM fraction;
  log_signature   16384 paths, width 3, depth 8, 64 increments: fused
                  signature, ft_log, tensor_to_lie, lie_to_tensor;
  uneven          8192 paths, width 8, depth 5, lengths uniform in [16, 1024]:
                  fused forward replayed from a hipGraph, FP32 and FP64;
  layout          the headline batch moved between the object-major and the
                  batch-interleaved layouts (vector_assign's tiled
                  transposition) and vector_add over it, against HBM peak;
  roofline        the dominant kernel of the headline path (the ft_fmexp
                  register tile, HBM-bound) from HIP events around its
                  launches;
  cpu_baseline    the oracle's bit-identical restatement of the reference
                  cpu::single_thread chain on the host's cores, on a bounded
                  sample of the headline workload, with the reference's
                  throughput ratio to it measured where the reference is
                  built (profiles/cpu_baseline_w5d6_chain.json).
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
SECTIONS = ("fwd_bwd", "ft_mul", "log_signature", "uneven", "layout")


def dim_of(w, n):
    return (w ** (n + 1) - 1) // (w - 1)


def chen_flops(w, n):
    """Floating-point operations of one Chen step S <- S (x) exp(x) in the
    Horner form: one multiply-add per (target k, level i <= k, word of level i),
    plus the scaled increments: 2 * sum_k sum_{i<=k} w^i."""
    return 2 * sum(sum(w ** i for i in range(1, k + 1)) for k in range(1, n + 1))


def make_increments(seed, paths, length, width):
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, np.sqrt(1.0 / length), size=(paths, length, width))


def rel_err(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.maximum(np.abs(a), np.abs(b)))))


class EventTimer:
    """HIP events on torch's current stream (where every rpp_hip call of this
    script is launched), summed over the bracketed regions."""

    def __init__(self, torch):
        self.torch = torch
        self.pairs = []

    def start(self):
        e = self.torch.cuda.Event(enable_timing=True)
        e.record()
        self.pairs.append([e, None])

    def stop(self):
        e = self.torch.cuda.Event(enable_timing=True)
        e.record()
        self.pairs[-1][1] = e

    def total_ms(self):
        self.torch.cuda.synchronize()
        return sum(a.elapsed_time(b) for a, b in self.pairs)


def timed(torch, step, steps, warmup):
    """Run `step` warmup times, then time exactly `steps` calls (device events
    and host wall clock, synchronised on both sides)."""
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record()
    out = None
    for _ in range(steps):
        out = step()
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return start.elapsed_time(stop) / steps, wall * 1e3 / steps, out


# ------------------------------------------------------------ headline (C2)

class Plain:
    """Chained ft_fmexp, one launch per increment (the reference usage)."""

    def __init__(self, rpp, torch, inc, dtype, dev, depth=DEPTH):
        self.rpp, self.torch = rpp, torch
        n, L, w = inc.shape
        self.n, self.L, self.w = n, L, w
        self.basis = rpp.TensorBasis(w, depth)
        dim = self.basis.size()
        padded = np.zeros((n, 1 + L * w))
        padded[:, 1:] = inc.reshape(n, L * w)
        self.pad = torch.tensor(padded, dtype=dtype, device=dev)
        self.a = torch.zeros((n, dim), dtype=dtype, device=dev)
        self.b = torch.empty_like(self.a)
        self.ex = [rpp.Batch(self.pad[:, t * w: t * w + 1 + w], 1, 1) for t in range(L)]
        self.timer = None

    def chain(self):
        self.a.zero_()
        self.a[:, 0] = 1.0
        cur, nxt = self.a, self.b
        if self.timer:
            self.timer.start()
        for t in range(self.L):
            self.rpp.ft_fmexp(nxt, cur, self.ex[t], self.basis)
            cur, nxt = nxt, cur
        if self.timer:
            self.timer.stop()
        return cur

    step = chain


class Fused:
    """All increments of every path in one persistent launch."""

    def __init__(self, rpp, torch, inc, dtype, dev, depth=DEPTH):
        self.rpp = rpp
        n, L, w = inc.shape
        self.basis = rpp.TensorBasis(w, depth)
        self.inc = torch.tensor(inc.reshape(n * L, w), dtype=dtype, device=dev)
        self.off = torch.arange(n + 1, dtype=torch.int64, device=dev) * L
        self.order = rpp.path_order(self.off)
        self.out = torch.empty((n, self.basis.size()), dtype=dtype, device=dev)

    def step(self):
        return self.rpp.signature(self.inc, self.off, self.basis, out=self.out, order=self.order)


def cpu_baseline(inc, runs=7):
    """The oracle's restatement of the reference cpu::single_thread chain
    (bit-identical to it), its batch split over threads on the GPU host's
    cores (at most 16, the host share of one GPU), on a bounded sample of the
    headline workload: median of `runs` timed runs after one warm-up.  The
    reference itself never runs here; its throughput relative to the
    restatement, measured on identical inputs where the reference is built
    (oracle/time_cpu_baseline.py), is read from the committed profile."""
    import statistics

    import oracle as O
    threads = max(1, min(16, os.cpu_count() or 1))
    sample = inc[:1024]
    flat = np.ascontiguousarray(sample.reshape(-1, WIDTH))
    off = np.arange(sample.shape[0] + 1, dtype=np.int64) * LENGTH
    O.signature_threads(WIDTH, DEPTH, flat, off, threads)  # warm-up
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        O.signature_threads(WIDTH, DEPTH, flat, off, threads)
        times.append(time.perf_counter() - t0)
    dt = statistics.median(times)
    units = sample.shape[0] * sample.shape[1]
    res = {"value": units / dt, "unit": "path-increments/s", "cores": threads, "kind": "port",
           "sample": f"{sample.shape[0]} paths x {sample.shape[1]} increments of the headline workload (width 5, "
                     f"depth 6, FP64); median of {runs} runs after one warm-up, {dt:.3f} s each, "
                     f"{threads} threads over contiguous path slices",
           "runs_s": times}
    prof = os.path.join(ROOT, "profiles", "cpu_baseline_w5d6_chain.json")
    if os.path.exists(prof):
        with open(prof) as f:
            ref = json.load(f)
        res["reference_over_port"] = {c: v["reference_over_restatement"] for c, v in ref["cores"].items()}
        res["reference_over_port_measured_on"] = f"{ref['host']['cpu']} ({ref['host']['logical_cpus']} CPUs)"
    return res


def load_traffic(name):
    p = os.path.join(ROOT, "profiles", name)
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f).get("hbm_bytes_per_launch")
    return None


def headline(rpp, torch, O, dev, steps, warmup):
    inc = make_increments(0xC2, PATHS, LENGTH, WIDTH)
    n_chk = 8
    chk_off = np.arange(n_chk + 1, dtype=np.int64) * LENGTH
    ref64 = O.signature(WIDTH, DEPTH, inc[:n_chk].reshape(-1, WIDTH), chk_off)
    # the oracle's FP32 chain, for the bitwise check of the plain FP32 path
    ref32 = O.signature(WIDTH, DEPTH, inc[:n_chk].reshape(-1, WIDTH), chk_off, dtype=np.float32)
    results = {}
    roof = None
    for name, cls, dt in (("plain_fp64", Plain, torch.float64), ("fused_fp64", Fused, torch.float64),
                          ("plain_fp32", Plain, torch.float32), ("fused_fp32", Fused, torch.float32)):
        runner = cls(rpp, torch, inc, dt, dev)
        if name == "plain_fp64":
            for _ in range(warmup):
                runner.step()
            runner.timer = EventTimer(torch)
            ms, wall, out = timed(torch, runner.step, steps, 0)
            chain_ms = runner.timer.total_ms() / steps
            roof = chain_ms / LENGTH
        else:
            ms, wall, out = timed(torch, runner.step, steps, warmup)
        got = out[:n_chk].cpu().numpy()
        results[name] = {"ms_per_step": ms, "wall_ms_per_step": wall,
                         "path_increments_per_s": PATHS * LENGTH / (ms * 1e-3),
                         "max_rel_err_vs_fp64_oracle": rel_err(got.astype(np.float64), ref64)}
        if name.startswith("plain"):
            # the plain chain is the reference's own ft_fmexp recursion
            results[name]["bit_identical_to_oracle_chain"] = bool(
                np.array_equal(got, ref64 if dt == torch.float64 else ref32))
        del runner
        torch.cuda.empty_cache()
    return inc, results, roof


# ----------------------------------------------------------- fwd+bwd (C5)

def fwd_bwd(rpp, torch, O, dev, steps=2, warmup=1):
    w, N, n, L = 6, 6, 2048, 256
    inc = make_increments(0xC5, n, L, w)
    basis = rpp.TensorBasis(w, N)
    dim = basis.size()
    rng = np.random.default_rng(0xC5 + 1)
    ell = torch.tensor(rng.normal(size=dim), device=dev)
    flat = torch.tensor(inc.reshape(n * L, w), device=dev)
    off = torch.arange(n + 1, dtype=torch.int64, device=dev) * L
    order = rpp.path_order(off)
    loss = torch.empty(n, dtype=torch.float64, device=dev)

    # plain per-op path: buffers for two states and two cotangents
    padded = torch.zeros((n, 1 + L * w), dtype=torch.float64, device=dev)
    padded[:, 1:] = flat.view(n, L * w)
    gpad = torch.zeros_like(padded)
    S0 = torch.zeros((n, dim), dtype=torch.float64, device=dev)
    S1 = torch.empty_like(S0)
    G0 = torch.empty_like(S0)
    G1 = torch.empty_like(S0)
    ex = [rpp.Batch(padded[:, t * w: t * w + 1 + w], 1, 1) for t in range(L)]
    gx = [rpp.Batch(gpad[:, t * w: t * w + 1 + w], 1, 1) for t in range(L)]

    def plain():
        S0.zero_()
        S0[:, 0] = 1.0
        cur, nxt = S0, S1
        for t in range(L):
            rpp.ft_fmexp(nxt, cur, ex[t], basis)
            cur, nxt = nxt, cur
        rpp.tensor_pairing(loss, rpp.Batch(ell, 0, N), cur, basis, num_batches=n)
        neg = -padded
        nex = [rpp.Batch(neg[:, t * w: t * w + 1 + w], 1, 1) for t in range(L)]
        Gc, Gn = G0, G1
        Gc.copy_(ell.expand(n, dim))
        for t in range(L - 1, -1, -1):
            rpp.ft_fmexp(nxt, cur, nex[t], basis)  # S_{t-1} = S_t exp(-x_t)
            rpp.ft_fmexp_adjoint(Gn, gx[t], nxt, ex[t], Gc, basis)
            cur, nxt = nxt, cur
            Gc, Gn = Gn, Gc
        return gpad

    def fused(dtype, every=0):
        x = flat.to(dtype)
        f = ell.to(dtype)
        out_sig = torch.empty((n, dim), dtype=dtype, device=dev)
        g = torch.empty_like(x)
        lss = torch.empty(n, dtype=dtype, device=dev)
        kw = {}
        if every:
            kw = dict(checkpoint_every=every,
                      checkpoints=torch.empty((n * (L // every), dim), dtype=dtype, device=dev),
                      ckpt_offsets=torch.arange(n + 1, dtype=torch.int64, device=dev) * (L // every))

        def step():
            rpp.signature(x, off, basis, out=out_sig, order=order, **kw)
            rpp.tensor_pairing(lss, rpp.Batch(f, 0, N), out_sig, basis, num_batches=n)
            rpp.signature_backward(x, off, basis, out_sig, f, grad=g, order=order, **kw)
            return g
        return step

    # FP64 oracle gradients of four paths spread over the batch
    chk = [0, n // 3, 2 * n // 3, n - 1]
    ell_np = ell.cpu().numpy()
    g_ref = np.concatenate([O.signature_vjp(w, N, inc[p], np.array([0, L], dtype=np.int64), ell_np) for p in chk])
    every32 = rpp.checkpoint_every_for(torch.float32, L)
    out = {"config": {"paths": n, "width": w, "depth": N, "coefficients": dim, "increments_per_path": L,
                      "loss": "tensor_pairing with a seeded N(0,1) functional", "dtype": "f64 (fp32_* lines: f32)",
                      "steps": steps, "warmup": warmup, "grad_err_paths": chk,
                      "fp32_default_checkpoint_every": every32}}
    runs = (("plain", plain), ("fused_reconstruct", fused(torch.float64)),
            ("fused_ckpt16", fused(torch.float64, 16)), ("fp32_fused_reconstruct", fused(torch.float32)),
            (f"fp32_fused_ckpt{every32}", fused(torch.float32, every32)))
    for name, fn in runs:
        ms, wall, g = timed(torch, fn, steps, warmup)
        if name == "plain":
            got = torch.cat([g[p, 1:].view(L, w) for p in chk])
        else:
            got = torch.cat([g[p * L:(p + 1) * L] for p in chk])
        out[name] = {"ms_per_step": ms, "wall_ms_per_step": wall, "paths_per_s": n / (ms * 1e-3),
                     "grad_max_rel_err_vs_fp64_oracle": rel_err(got.double().cpu().numpy(), g_ref)}
        del fn
        torch.cuda.empty_cache()
    out["fused_speedup"] = out["plain"]["ms_per_step"] / out["fused_reconstruct"]["ms_per_step"]
    out["fused_ckpt_speedup"] = out["plain"]["ms_per_step"] / out["fused_ckpt16"]["ms_per_step"]
    del S0, S1, G0, G1, padded, gpad
    torch.cuda.empty_cache()
    return out


# ------------------------------------------------------------- ft_mul (C1)

def capi_call_cost(pairs):
    """The same ft_mul called from C++ through the C-ABI, as the drop-in layer
    calls it (rpp_hip/capi_call_bench, a child process): host time per call of
    the enqueueing loop, wall time per call until the stream drains, device
    time per call."""
    exe = os.path.join(ROOT, "rough-path-primitives_amd", "rpp_hip", "capi_call_bench")
    if not os.path.exists(exe):
        return None
    import subprocess
    r = subprocess.run([exe, str(pairs), "2000"], capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        return {"error": (r.stderr or r.stdout).strip()[-300:]}
    res = json.loads(r.stdout.strip().splitlines()[-1])
    res["achieved_gbs"] = 3 * pairs * 341 * 8 / (res["wall_us_per_call"] * 1e-6) / 1e9
    return res


def ft_mul_section(rpp, torch, O, dev):
    w, N = 4, 4
    basis = rpp.TensorBasis(w, N)
    dim = basis.size()
    res = {}
    for label, n, reps in (("c1_1024_pairs", 1024, 200), ("hbm_131072_pairs", 1 << 17, 50)):
        rng = np.random.default_rng(0xC1)
        lhs = rng.uniform(-1, 1, size=(n, dim))
        rhs = rng.uniform(-1, 1, size=(n, dim))
        lhs[:, 0] = 1.0
        rhs[:, 0] = 1.0
        L = torch.tensor(lhs, device=dev)
        R = torch.tensor(rhs, device=dev)
        O_ = torch.empty_like(L)
        ms, _, _ = timed(torch, lambda: rpp.ft_mul(O_, L, R, basis), reps, 5)
        k = min(n, 256)
        expected = np.zeros((k, dim))
        O.orc.ft_mul(w, N, expected, lhs[:k], rhs[:k])
        entry = {"pairs": n, "us_per_call": ms * 1e3,
                 "max_rel_err_vs_fp64_oracle": rel_err(O_[:k].cpu().numpy(), expected)}
        alg = 3 * n * dim * 8
        entry["algorithmic_bytes"] = alg
        entry["achieved_gbs"] = alg / (ms * 1e-3) / 1e9
        entry["hbm_frac"] = entry["achieved_gbs"] / HBM_PEAK_GBS
        if n == 1 << 17:
            entry["traffic_bytes_per_call"] = load_traffic("k_product_small_w4d4_131072pairs_fp64_traffic.json")
        if n == 1024:
            # the same calls replayed from one captured graph: the kernel and
            # its launch without the Python wrapper's per-call work
            calls = 50
            graph = torch.cuda.CUDAGraph()
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                rpp.ft_mul(O_, L, R, basis)
            torch.cuda.current_stream().wait_stream(side)
            torch.cuda.synchronize()
            with torch.cuda.graph(graph):
                for _ in range(calls):
                    rpp.ft_mul(O_, L, R, basis)
            gms, _, _ = timed(torch, graph.replay, 20, 3)
            entry["graph_us_per_call"] = gms * 1e3 / calls
            entry["graph_achieved_gbs"] = 3 * n * dim * 8 / (gms / calls * 1e-3) / 1e9
            entry["cpp_capi"] = capi_call_cost(n)
        if n == 1024:
            # the oracle's restatement of the reference product (the same
            # summation order, bit for bit) on one core: median of 7 runs
            out_cpu = np.zeros((n, dim))
            O.orc.ft_mul(w, N, out_cpu, lhs, rhs)
            runs = []
            for _ in range(7):
                t0 = time.perf_counter()
                O.orc.ft_mul(w, N, out_cpu, lhs, rhs)
                runs.append(time.perf_counter() - t0)
            entry["cpu_port_us_per_call"] = float(np.median(runs)) * 1e6
            entry["cpu_port_cores"] = 1
            entry["cpu_port_bit_identical_to_gpu"] = bool(np.array_equal(O_.cpu().numpy(), out_cpu))
        res[label] = entry
        del L, R, O_
        torch.cuda.empty_cache()
    res["config"] = {"width": w, "depth": N, "coefficients": dim, "dtype": "f64",
                     "inputs": "Uniform[-1,1], degree 0 set to 1"}
    return res


# ------------------------------------------------- layouts and linear operations

def layout_section(rpp, torch, O, dev, reps=20):
    """The wire-format transposition (vector_assign between object-major and
    batch-interleaved batches) and vector_add, at the headline shape: 4096
    objects of width 5, depth 6 (19531 coefficients), FP64.  Algorithmic bytes:
    one read and one write of every coefficient (transposition), two reads
    and one write (vector_add).  torch's own transpose copy is timed beside
    it."""
    B = rpp.TensorBasis(WIDTH, DEPTH)
    n, dim = PATHS, B.size()
    g = torch.Generator(device=dev)
    g.manual_seed(0xC2)
    x = torch.randn((n, dim), dtype=torch.float64, device=dev, generator=g)
    y = torch.randn((n, dim), dtype=torch.float64, device=dev, generator=g)
    il = torch.empty((dim, n), dtype=torch.float64, device=dev)
    om = torch.empty_like(x)
    out = torch.empty_like(x)
    nbytes = n * dim * 8
    res = {"config": {"objects": n, "width": WIDTH, "depth": DEPTH, "coefficients": dim, "dtype": "f64"}}

    def entry(ms, traffic):
        gbs = traffic / (ms * 1e-3) / 1e9
        return {"us_per_call": ms * 1e3, "algorithmic_bytes": traffic, "achieved_gbs": gbs,
                "hbm_frac": gbs / HBM_PEAK_GBS}

    to_il = lambda: rpp.vector_assign(rpp.interleaved(il, B), x, B, num_batches=n)
    to_om = lambda: rpp.vector_assign(om, rpp.interleaved(il, B), B, num_batches=n)
    add = lambda: rpp.vector_add(out, x, y, B, alpha=0.5, beta=-2.0)
    ms, _, _ = timed(torch, to_il, reps, 3)
    res["object_major_to_interleaved"] = entry(ms, 2 * nbytes)
    res["object_major_to_interleaved"]["exact"] = bool(torch.equal(il, x.t()))
    ms, _, _ = timed(torch, to_om, reps, 3)
    res["interleaved_to_object_major"] = entry(ms, 2 * nbytes)
    res["interleaved_to_object_major"]["exact"] = bool(torch.equal(om, x))
    ms, _, _ = timed(torch, add, reps, 3)
    res["vector_add"] = entry(ms, 3 * nbytes)
    res["vector_add"]["exact"] = bool(torch.equal(out, 0.5 * x + (-2.0) * y))
    ms, _, _ = timed(torch, lambda: il.copy_(x.t()), reps, 3)
    res["torch_transpose_copy"] = entry(ms, 2 * nbytes)
    del x, y, il, om, out
    torch.cuda.empty_cache()
    return res


# ---------------------------------------------------------- log-signature (C3)

def lie_l2t(lie, dim, ts):
    import scipy.sparse as sp
    l2t = lie["l2t"]
    return sp.csc_matrix((l2t["data"], l2t["indices"], l2t["offsets"]), shape=(dim, ts - 1))


def log_signature(rpp, torch, O, dev, steps=3, warmup=1):
    """Signature -> ft_log -> tensor_to_lie -> lie_to_tensor, each stage timed
    with HIP events; the round trip lie_to_tensor(tensor_to_lie(log S)) returns
    log S (a Lie element), a size-independent check on all 16384 paths."""
    import scipy.sparse as sp
    w, N, n, L = 3, 8, 16384, 64
    inc = make_increments(0xC3, n, L, w)
    basis = rpp.TensorBasis(w, N)
    dim = basis.size()
    lie = rpp.lie_basis(w, N)
    ts = lie["true_size"]
    flat = torch.tensor(inc.reshape(n * L, w), device=dev)
    off = torch.arange(n + 1, dtype=torch.int64, device=dev) * L
    order = rpp.path_order(off)
    sig = torch.empty((n, dim), dtype=torch.float64, device=dev)
    lg = torch.empty_like(sig)
    back = torch.empty_like(sig)
    lie_out = torch.empty((n, ts), dtype=torch.float64, device=dev)
    rpp.tensor_to_lie(lie_out, lg, basis)  # uploads the Hall metadata once
    stages = ("signature", "ft_log", "tensor_to_lie", "lie_to_tensor")
    timers = {k: EventTimer(torch) for k in stages}

    def step(timed_stages=False):
        calls = ((lambda: rpp.signature(flat, off, basis, out=sig, order=order)),
                 (lambda: rpp.ft_log(lg, sig, basis)),
                 (lambda: rpp.tensor_to_lie(lie_out, lg, basis)),
                 (lambda: rpp.lie_to_tensor(back, lie_out, basis)))
        for name, call in zip(stages, calls):
            if timed_stages:
                timers[name].start()
            call()
            if timed_stages:
                timers[name].stop()
        return lie_out

    ms, wall, got = timed(torch, step, steps, warmup)
    for _ in range(steps):
        step(True)
    stage_ms = {k: timers[k].total_ms() / steps for k in stages}
    k = 4
    s_ref = O.signature(w, N, inc[:k].reshape(-1, w), np.arange(k + 1, dtype=np.int64) * L)
    log_ref = np.zeros_like(s_ref)
    O.orc.ft_log(w, N, log_ref, s_ref)
    t2l = lie["t2l"]
    T2L = sp.csc_matrix((t2l["data"], t2l["indices"], t2l["offsets"]), shape=(ts - 1, dim))
    lie_ref = (T2L @ log_ref.T).T
    lg_all = lg.cpu().numpy()
    back_all = back.cpu().numpy()
    rt = np.max(np.abs(back_all - lg_all) / np.maximum(1.0, np.maximum(np.abs(back_all), np.abs(lg_all))), axis=1)
    worst = int(np.argmax(rt))
    # the worst round-trip path against the oracle: a large round-trip error
    # with a small error against the oracle is rounding in the log itself
    s_w = O.signature(w, N, inc[worst], np.array([0, L], dtype=np.int64))
    log_w = np.zeros_like(s_w)
    O.orc.ft_log(w, N, log_w, s_w)
    back_w = np.asarray((lie_l2t(lie, dim, ts) @ (T2L @ log_w.T)).T)
    # the same path through the drop-in operations: the plain chained
    # ft_fmexp (the reference's recursion), ft_log and tensor_to_lie, each
    # compared bit for bit with the oracle's (the reference's) results
    plain = Plain(rpp, torch, inc[worst:worst + 1], torch.float64, dev, depth=N).chain()
    lg_p = torch.empty_like(plain)
    rpp.ft_log(lg_p, plain, basis)
    lie_p = torch.zeros((1, ts), dtype=torch.float64, device=dev)
    rpp.tensor_to_lie(lie_p, lg_p, basis)
    lie_w = np.zeros((1, ts))
    O.spmv(t2l["data"], t2l["indices"], t2l["offsets"], dim, lie_w, log_w, csc=True, out_offset=1, out_size=ts - 1)
    res = {"config": {"paths": n, "width": w, "depth": N, "coefficients": dim, "lie_dimension": ts - 1,
                      "increments_per_path": L, "dtype": "f64", "steps": steps, "warmup": warmup,
                      "pipeline": "signature -> ft_log -> tensor_to_lie -> lie_to_tensor"},
           "ms_per_step": ms, "wall_ms_per_step": wall, "paths_per_s": n / (ms * 1e-3),
           "stage_ms": stage_ms,
           "log_max_rel_err_vs_fp64_oracle": rel_err(lg_all[:k], log_ref),
           "lie_max_rel_err_vs_fp64_oracle": rel_err(got[:k, 1:].cpu().numpy(), lie_ref),
           "round_trip_max_rel_err_all_paths": float(rt[worst]),
           "round_trip_worst_path": {
               "path": worst,
               # the fused signature sums in its own order (1e-16 apart from
               # the reference's chain); the log's conditioning on this path,
               # seen in the oracle's own round trip, amplifies that
               "fused_signature_rel_err_vs_oracle": rel_err(sig[worst:worst + 1].cpu().numpy(), s_w),
               "fused_pipeline_log_rel_err_vs_oracle_pipeline": rel_err(lg_all[worst], log_w[0]),
               "oracle_round_trip_rel_err": rel_err(back_w, log_w),
               "max_abs_log": float(np.abs(log_w).max()),
               "plain_chain_signature_bit_identical": bool(np.array_equal(plain.cpu().numpy(), s_w)),
               "plain_chain_log_bit_identical": bool(np.array_equal(lg_p.cpu().numpy(), log_w)),
               "plain_chain_lie_bit_identical": bool(np.array_equal(lie_p.cpu().numpy()[:, 1:], lie_w[:, 1:]))},
           "ft_log_worst_conditioned": log_worst_conditioned(rpp, torch, O, dev)}
    del sig, lg, lie_out, back
    torch.cuda.empty_cache()
    return res


def log_worst_conditioned(rpp, torch, O, dev):
    """ft_log of the 32 worst-conditioned log-signatures of this workload
    (|log S| up to 30) on the reference's own signatures (rebuilt bit for bit
    by the oracle's chain), against the reference's ft_log from the committed
    fixture: the parity of ft_log itself, on identical inputs."""
    import hashlib
    path = os.path.join(ROOT, "tests", "golden", "log_c3_worst_w3d8.npz")
    with np.load(path) as z:
        inc, expected, sha = z["increments"], z["log"], z["signature_sha256"]
    n, L, w = inc.shape
    sig = O.signature(w, 8, inc.reshape(-1, w), np.arange(n + 1, dtype=np.int64) * L)
    same_input = hashlib.sha256(np.ascontiguousarray(sig).tobytes()).digest() == sha.tobytes()
    out = torch.empty(sig.shape, dtype=torch.float64, device=dev)
    rpp.ft_log(out, torch.tensor(sig, device=dev), rpp.TensorBasis(w, 8))
    got = out.cpu().numpy()
    return {"paths": n, "max_abs_log": float(np.abs(expected).max()), "input_is_reference_signature": bool(same_input),
            "max_rel_err_vs_reference": rel_err(got, expected), "bit_identical": bool(np.array_equal(got, expected))}


# ------------------------------------------------------------- uneven (C4)

def uneven(rpp, torch, O, dev, steps=3, warmup=1):
    w, N, n = 8, 5, 8192
    rng = np.random.default_rng(0xC4)
    lengths = rng.integers(16, 1025, size=n)
    off_np = np.zeros(n + 1, dtype=np.int64)
    off_np[1:] = np.cumsum(lengths)
    inc = np.concatenate([rng.normal(0.0, np.sqrt(1.0 / Lp), size=(Lp, w)) for Lp in lengths])
    basis = rpp.TensorBasis(w, N)
    off = torch.tensor(off_np, device=dev)
    total = int(off_np[-1])
    # the two shortest and the two longest paths against the FP64 oracle
    srt = np.argsort(lengths, kind="stable")
    chk = np.concatenate([srt[:2], srt[-2:]])
    ref = np.concatenate([O.signature(w, N, inc[off_np[p]:off_np[p + 1]], np.array([0, lengths[p]]))
                          for p in chk])
    res = {"config": {"paths": n, "width": w, "depth": N, "coefficients": basis.size(),
                      "lengths": "uniform integers in [16, 1024]", "increments": total,
                      "launch": "hipGraph replay of the fused persistent kernel", "steps": steps,
                      "warmup": warmup}}
    for name, dt, peak in (("fp32", torch.float32, FP32_PEAK_TFLOPS), ("fp64", torch.float64, FP64_PEAK_TFLOPS)):
        g = rpp.SignatureGraph(torch.tensor(inc, dtype=dt, device=dev), off, basis)
        ms, wall, out = timed(torch, g.replay, steps, warmup)
        tflops = total * chen_flops(w, N) / (ms * 1e-3) / 1e12
        got = out[torch.tensor(chk, device=dev)].double().cpu().numpy()
        res[name] = {"ms_per_step": ms, "wall_ms_per_step": wall,
                     "path_increments_per_s": total / (ms * 1e-3), "paths_per_s": n / (ms * 1e-3),
                     "vector_tflops": tflops, "vector_frac": tflops / peak,
                     "max_rel_err_vs_fp64_oracle": rel_err(got, ref),
                     "checked_path_lengths": [int(lengths[p]) for p in chk],
                     "rel_err_longest_path": rel_err(got[-1], ref[-1])}
        del g, out
        torch.cuda.empty_cache()
    return res


def fp32_drift(rpp, torch, O, dev):
    """FP32 gradient error of the batched backward against the FP64 oracle,
    exp(-x) reconstruction against ceil(sqrt(L)) checkpoints, over path
    lengths 16 .. 1024 at three shapes (4 paths each, 2 at 1024 increments;
    increments N(0, 1/L), loss <ell, S> with ell ~ N(0, 1)): the measurement
    behind the library's checkpoint policy (rpp_hip_signature_checkpoint_every)."""
    import math
    rows = []
    for (w, N) in ((6, 6), (5, 6), (3, 8)):
        for L in (16, 32, 64, 128, 256, 512, 1024):
            n = 2 if L == 1024 else 4
            rng = np.random.default_rng(L * 7 + w)
            inc = rng.normal(0, np.sqrt(1 / L), size=(n * L, w))
            off = np.arange(n + 1, dtype=np.int64) * L
            basis = rpp.TensorBasis(w, N)
            dim = basis.size()
            ell = rng.normal(size=dim)
            ref = O.signature_vjp(w, N, inc, off, ell)
            row = {"width": w, "depth": N, "increments": L, "paths": n}
            every = int(math.ceil(math.sqrt(L)))
            for label, k in (("reconstruct", 0), ("checkpoints", every)):
                x = torch.tensor(inc, dtype=torch.float32, device=dev)
                o = torch.tensor(off, device=dev)
                kw = {}
                if k:
                    kw = dict(checkpoint_every=k,
                              checkpoints=torch.empty((n * (L // k), dim), dtype=torch.float32, device=dev),
                              ckpt_offsets=torch.arange(n + 1, dtype=torch.int64, device=dev) * (L // k))
                sig = rpp.signature(x, o, basis, **kw)
                g = rpp.signature_backward(x, o, basis, sig, torch.tensor(ell, dtype=torch.float32, device=dev), **kw)
                row[label] = rel_err(g.double().cpu().numpy(), ref)
            row["checkpoint_every"] = every
            row["ratio"] = row["reconstruct"] / row["checkpoints"]
            print(json.dumps(row), file=sys.stderr, flush=True)
            rows.append(row)
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sections", default=",".join(SECTIONS),
                    help="comma list of extra sections to run (" + ",".join(SECTIONS) + "), or 'none'")
    ap.add_argument("--only", default=None,
                    help="run one headline variant (plain_fp64, fused_fp64, plain_fp32, fused_fp32), "
                         "ft_mul_hbm (2^17 width-4 depth-4 products), capi_call (1024 width-4 depth-4 "
                         "products called from C++), layout, fused_fwd_bwd (the width-6 "
                         "depth-6 fused forward + backward), log_stages (the width-3 depth-8 log-signature "
                         "stages) or fp32_drift (the FP32 gradient error sweep behind the checkpoint policy), "
                         "and print its result only")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported: paths are independent, "
              "so a multi-GPU run would shard the batch by paths with no communication in the hot path",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import oracle as O
    import rpp_hip as rpp

    dev = torch.device("cuda:0")
    if args.copy_mib:
        src = torch.ones(args.copy_mib * (1 << 20) // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    if args.only == "ft_mul_hbm":
        w, N, n = 4, 4, 1 << 17
        basis = rpp.TensorBasis(w, N)
        rng = np.random.default_rng(0xC1)
        L = torch.tensor(rng.uniform(-1, 1, size=(n, basis.size())), device=dev)
        R = torch.tensor(rng.uniform(-1, 1, size=(n, basis.size())), device=dev)
        out = torch.empty_like(L)
        ms, _, _ = timed(torch, lambda: rpp.ft_mul(out, L, R, basis), args.steps, args.warmup)
        print(json.dumps({"variant": args.only, "us_per_call": ms * 1e3,
                          "achieved_gbs": 3 * n * basis.size() * 8 / (ms * 1e-3) / 1e9}))
        return
    if args.only == "capi_call":
        print(json.dumps({"variant": args.only, **(capi_call_cost(1024) or {})}))
        return
    if args.only == "fused_fwd_bwd":
        # the fused forward + backward of the width 6, depth 6 workload (2048
        # paths x 256 increments, FP64, exp(-x) reconstruction) on its own,
        # for profiling its two kernels
        w, N, n, L = 6, 6, 2048, 256
        basis = rpp.TensorBasis(w, N)
        flat = torch.tensor(make_increments(0xC5, n, L, w).reshape(n * L, w), device=dev)
        off = torch.arange(n + 1, dtype=torch.int64, device=dev) * L
        order = rpp.path_order(off)
        ell = torch.tensor(np.random.default_rng(0xC5 + 1).normal(size=basis.size()), device=dev)
        sig = torch.empty((n, basis.size()), dtype=torch.float64, device=dev)
        grad = torch.empty_like(flat)

        def step():
            rpp.signature(flat, off, basis, out=sig, order=order)
            rpp.signature_backward(flat, off, basis, sig, ell, grad=grad, order=order)
        ms, _, _ = timed(torch, step, args.steps, args.warmup)
        print(json.dumps({"variant": args.only, "ms_per_step": ms, "paths_per_s": n / (ms * 1e-3)}))
        return
    if args.only == "log_stages":
        # the four GPU stages of the width 3, depth 8 log-signature workload
        # (16384 paths x 64 increments, FP64) on their own, for profiling
        w, N, n, L = 3, 8, 16384, 64
        basis = rpp.TensorBasis(w, N)
        flat = torch.tensor(make_increments(0xC3, n, L, w).reshape(n * L, w), device=dev)
        off = torch.arange(n + 1, dtype=torch.int64, device=dev) * L
        order = rpp.path_order(off)
        sig = torch.empty((n, basis.size()), dtype=torch.float64, device=dev)
        lg, back = torch.empty_like(sig), torch.empty_like(sig)
        lie = torch.empty((n, rpp.lie_basis(w, N)["true_size"]), dtype=torch.float64, device=dev)

        def step():
            rpp.signature(flat, off, basis, out=sig, order=order)
            rpp.ft_log(lg, sig, basis)
            rpp.tensor_to_lie(lie, lg, basis)
            rpp.lie_to_tensor(back, lie, basis)
        ms, _, _ = timed(torch, step, args.steps, args.warmup)
        print(json.dumps({"variant": args.only, "ms_per_step": ms, "paths_per_s": n / (ms * 1e-3)}))
        return
    if args.only == "fp32_drift":
        print(json.dumps({"variant": args.only, "rows": fp32_drift(rpp, torch, O, dev)}))
        return
    if args.only == "layout":
        print(json.dumps({"variant": args.only, **layout_section(rpp, torch, O, dev, reps=args.steps)}))
        return
    if args.only:
        inc = make_increments(0xC2, PATHS, LENGTH, WIDTH)
        cls = Plain if args.only.startswith("plain") else Fused
        dt = torch.float64 if args.only.endswith("fp64") else torch.float32
        runner = cls(rpp, torch, inc, dt, dev)
        ms, _, _ = timed(torch, runner.step, args.steps, args.warmup)
        print(json.dumps({"variant": args.only, "ms_per_step": ms,
                          "path_increments_per_s": PATHS * LENGTH / (ms * 1e-3)}))
        return

    inc, results, launch_ms = headline(rpp, torch, O, dev, args.steps, args.warmup)
    plain = results["plain_fp64"]
    dim = dim_of(WIDTH, DEPTH)
    alg_bytes = PATHS * (2 * dim + WIDTH) * 8
    achieved = alg_bytes / (launch_ms * 1e-3) / 1e9
    tile = rpp.fmexp_tile(WIDTH, DEPTH, torch.float64)
    kernel = f"{tile[0]}<double,{WIDTH},{DEPTH},{tile[1]},{tile[2]}>" if tile else "k_fmexp_horner"
    fused = results["fused_fp64"]
    fused_tflops = PATHS * LENGTH * chen_flops(WIDTH, DEPTH) / (fused["ms_per_step"] * 1e-3) / 1e12
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
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": load_traffic("k_fmexp_coal_w5d6_4096objects_fp64_traffic.json"),
                     "kernel": kernel, "avg_launch_ms": launch_ms,
                     "algorithmic_bytes_per_launch": alg_bytes},
        "max_rel_err_vs_fp64_oracle": plain["max_rel_err_vs_fp64_oracle"],
        "plain_vs_fused": results,
        "fused_speedup_fp64": fused["path_increments_per_s"] / plain["path_increments_per_s"],
        "fused_fp64_vector_tflops": fused_tflops,
        "fused_fp64_vector_frac": fused_tflops / FP64_PEAK_TFLOPS,
    }
    del inc
    wanted = [] if args.sections == "none" else [s for s in args.sections.split(",") if s]
    runners = {"fwd_bwd": fwd_bwd, "ft_mul": ft_mul_section, "log_signature": log_signature, "uneven": uneven,
               "layout": layout_section}
    for name in wanted:
        print(f"section {name}", file=sys.stderr, flush=True)
        line[name] = runners[name](rpp, torch, O, dev)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(make_increments(0xC2, PATHS, LENGTH, WIDTH))
    print(json.dumps(line))


if __name__ == "__main__":
    main()
