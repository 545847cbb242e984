"""Time the reference's cpu::single_thread chain and the oracle's restatement
of it on the same inputs, on 1 and 8 cores, in a container that holds the
reference (oracle/_ref built by `make -C oracle ref`).

Workload: the headline signature chain, width 5, depth 6, 128 increments per
path drawn iid N(0, 1/128) per channel (seed 0xC2, the benchmark's
generator), one ft_fmexp per increment.  Multi-core runs split the batch into
contiguous slices, one std::thread (reference) or Python thread around the C
call (restatement) per core.  Each figure is the median of 7 timed runs after
one warm-up run; the two outputs are checked equal bit for bit.

The benchmark times the restatement on the GPU host's cores (the reference
never leaves this container) and reports the reference-to-restatement ratio
measured here beside it, from profiles/cpu_baseline_w5d6_chain.json.

    python3 oracle/time_cpu_baseline.py
"""
import json
import os
import platform
import statistics
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, HERE)

import oracle as O  # noqa: E402

WIDTH, DEPTH, LENGTH = 5, 6, 128
RUNS = 7


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def median_time(fn):
    fn()  # warm-up
    times = []
    for _ in range(RUNS):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    return statistics.median(times), times


def main():
    if not O.has_ref():
        raise SystemExit("the reference driver (oracle/_ref) is not built here")
    rng = np.random.default_rng(0xC2)
    inc = rng.normal(0.0, np.sqrt(1.0 / LENGTH), size=(512, LENGTH, WIDTH))
    res = {"workload": "signature chain, width 5, depth 6, 128 increments N(0, 1/128), seed 0xC2, FP64",
           "unit": "path-increments/s", "runs": RUNS, "statistic": "median after one warm-up",
           "host": {"cpu": cpu_model(), "logical_cpus": os.cpu_count()},
           "build": {"reference": "g++ -O3 -DNDEBUG -std=c++17 (oracle/Makefile ref)",
                     "restatement": "gcc -O3 -ffp-contract=off -std=c11 (oracle/Makefile)"},
           "cores": {}}
    for cores, paths in ((1, 64), (8, 512)):
        x = np.ascontiguousarray(inc[:paths])
        flat = x.reshape(-1, WIDTH)
        off = np.arange(paths + 1, dtype=np.int64) * LENGTH
        ref_out = O.ref_signature(WIDTH, DEPTH, x, threads=cores)
        port_out = O.signature_threads(WIDTH, DEPTH, flat, off, cores)
        assert np.array_equal(ref_out, port_out), "restatement and reference differ"
        t_ref, runs_ref = median_time(lambda: O.ref_signature(WIDTH, DEPTH, x, threads=cores))
        t_port, runs_port = median_time(lambda: O.signature_threads(WIDTH, DEPTH, flat, off, cores))
        units = paths * LENGTH
        res["cores"][str(cores)] = {
            "paths": paths, "reference": units / t_ref, "restatement": units / t_port,
            "reference_over_restatement": t_port / t_ref,
            "reference_runs_s": runs_ref, "restatement_runs_s": runs_port, "bit_identical": True}
        print(f"{cores} core(s), {paths} paths: reference {units / t_ref:,.0f}, restatement "
              f"{units / t_port:,.0f} path-increments/s (ratio {t_port / t_ref:.3f})", flush=True)
    out = os.path.join(ROOT, "profiles", "cpu_baseline_w5d6_chain.json")
    with open(out, "w") as f:
        json.dump(res, f, indent=1)
    lines = ["# CPU baseline: the width 5, depth 6 signature chain (128 increments, FP64)", "",
             "Command: `python3 oracle/time_cpu_baseline.py` (in the container that holds the reference).", "",
             f"Host: {res['host']['cpu']}, {res['host']['logical_cpus']} logical CPUs.  "
             f"Median of {RUNS} runs after one warm-up; both outputs equal bit for bit.", "",
             "| cores | paths | reference (path-increments/s) | restatement (path-increments/s) | "
             "reference / restatement |", "|---|---|---|---|---|"]
    for c, v in res["cores"].items():
        lines.append(f"| {c} | {v['paths']} | {v['reference']:,.0f} | {v['restatement']:,.0f} | "
                     f"{v['reference_over_restatement']:.3f} |")
    lines += ["", "The benchmark times the restatement on the GPU host's cores and reports this ratio "
              "beside it (`cpu_baseline.reference_over_port`)."]
    with open(os.path.join(ROOT, "profiles", "cpu_baseline_w5d6_chain.md"), "w") as f:
        f.write("\n".join(lines) + "\n")
    print("wrote", out)


if __name__ == "__main__":
    main()
