#!/usr/bin/env python3
"""Benchmark of the distance-distribution hot path on one MI355X.

Workload (one "step"): the exhaustive AN pair-distance histogram of the code
A=58659, k=24 -- all 2^24 data words, 40-bit codewords c = A*d, every
unordered pair (C(2^24, 2) = 1.407e14 pairs), bins 0..40, uint64 counts --
computed by the HIP pair kernel through the C ABI (cr_an_pairs).  There are no
input tensors: codewords are generated in registers, so the inputs are
trivially resident before the timed region starts.

Prints ONE JSON line (rank 0).  Extra keys beyond the required ones:
  roofline       VALU roofline of the pair kernel (events measured live)
  cpu_baseline   the oracle's restatement of the reference CPU loop on a
                 bounded row-shard sample of the same workload, host cores
  batched        batched-vs-per-launch comparison (when --batched is given)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

A_HEADLINE = 58659
K_HEADLINE = 24
PEAK_CLOCK_GHZ = 2.4
LANES_PER_CU_CLK = 64


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(A: int, k: int, target_s: float) -> dict:
    """Times the oracle's restatement of the reference CPU loop nest
    (an_coding.cpp:50-102, OpenMP dynamic over 64-word row shards) on the first
    row shards of the same (A, k) problem, for about target_s seconds."""
    from oracle import oracle as O
    O.lib()
    threads = O.num_threads()
    # calibrate on one shard (64 rows x 2^k columns)
    t0 = time.perf_counter()
    _, pairs = O.an_cpu_counts(k, A, 0, 1)
    dt = max(time.perf_counter() - t0, 1e-6)
    shards = max(1, int(target_s / dt))
    nshards_total = (1 << k) // 64
    shards = min(shards, nshards_total)
    t0 = time.perf_counter()
    _, pairs = O.an_cpu_counts(k, A, 0, shards)
    dt = time.perf_counter() - t0
    return {
        "value": round(pairs / dt / 1e9, 4),
        "unit": "Gpairs/s",
        "cores": threads,
        "kind": "port",
        "sample": f"A={A} k={k}: row shards 0..{shards - 1} of {nshards_total} "
                  f"(64 words each vs all later words), {pairs} pairs in {dt:.2f} s on {cpu_model()}",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--k", type=int, default=K_HEADLINE, help="data bits (default 24)")
    ap.add_argument("--A", type=int, default=A_HEADLINE)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--rows-per-lane", type=int, default=0)
    ap.add_argument("--chunk-blocks", type=int, default=0)
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is design-only, see DESIGN.md)"}))
        return 2

    import coding_reliability_amd as cra

    A, k = args.A, args.k
    N = 1 << k
    pairs_per_step = N * (N - 1) // 2
    ctx = cra.Context(0)
    info = ctx.device_info()
    log(f"device {info['name']} {info['arch']} CUs={info['compute_units']} clock={info['max_clock_mhz']} MHz")
    tune = dict(rows_per_lane=args.rows_per_lane, chunk_blocks=args.chunk_blocks)

    for w in range(args.warmup):
        _, ms = ctx.an_pairs(A, N, **tune)
        log(f"warmup {w}: kernel {ms:.1f} ms")

    kernel_ms = []
    raw = None
    t0 = time.perf_counter()
    for s in range(args.steps):
        raw, ms = ctx.an_pairs(A, N, **tune)
        kernel_ms.append(ms)
        log(f"step {s}: kernel {ms:.1f} ms")
    elapsed = time.perf_counter() - t0
    assert sum(raw) == pairs_per_step, "histogram does not sum to C(2^k,2)"

    ms_per_step = elapsed * 1e3 / args.steps
    gpairs = pairs_per_step * args.steps / elapsed / 1e9
    avg_kernel_s = sum(kernel_ms) / len(kernel_ms) / 1e3

    # VALU roofline: per pair the kernel needs one xor, one popcount and one
    # histogram increment (3 int32 lane-ops); for 33..64-bit codewords whose
    # column high dword is constant over 2^32/A columns, the high-dword
    # distance is a per-segment constant folded into the popcount.
    width_bits = ((N - 1) * A).bit_length()
    W = 1 if width_bits <= 32 else 2
    ops_per_pair = 3
    cus = info["compute_units"]
    peak = cus * LANES_PER_CU_CLK * PEAK_CLOCK_GHZ * 1e9 / 1e12  # T lane-ops/s
    achieved = pairs_per_step / avg_kernel_s * ops_per_pair / 1e12
    roofline = {
        "bound": "valu",
        "achieved": round(achieved, 3),
        "peak": round(peak, 3),
        "unit": "Tops/s (int32 lane-ops)",
        "frac": round(achieved / peak, 4),
        "traffic": None,
        "kernel": "pair_kernel<R>",
        "kernel_ms": round(avg_kernel_s * 1e3, 3),
        "ops_per_pair": ops_per_pair,
        "survey_ops_per_pair_2W_plus_1": 2 * W + 1,
        "peak_clock_ghz": PEAK_CLOCK_GHZ,
    }

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        cpu = cpu_baseline(A, k, args.cpu_seconds)
        log(f"cpu baseline {cpu['value']} Gpairs/s on {cpu['cores']} threads")

    out = {
        "metric": "pair-comparisons/sec (Gpairs/s), exhaustive AN histogram",
        "value": round(gpairs, 3),
        "unit": "Gpairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32/u64 integer (exact)",
        "data": "synthetic (codewords c=A*d generated on the GPU)",
        "config": {"workload": f"exhaustive AN histogram, A={A}, k={k}, {width_bits}-bit codewords, "
                               f"all C(2^{k},2) pairs, {width_bits + 1} bins",
                   "A": A, "k": k, "pairs_per_step": pairs_per_step, "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
