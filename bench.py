#!/usr/bin/env python3
"""Benchmark of the distance-distribution hot path on one MI355X.

Workload (one "step"): the exhaustive AN pair-distance histogram of the code
A=58659, k=24 -- all 2^24 data words, 40-bit codewords c = A*d, every
unordered pair (C(2^24, 2) = 1.407e14 pairs), bins 0..40, uint64 counts --
computed by the HIP pair kernel through the C ABI (cr_an_pairs).  There are no
input tensors: codewords are generated in registers, so the inputs are
trivially resident before the timed region starts.

Prints ONE JSON line (rank 0).  Extra keys beyond the required ones:
  roofline       LDS-increment roofline of the pair kernel (its binding one)
                 and its VALU roofline, each against a ceiling measured in
                 this run (cr_probe_ceiling), kernel time from HIP events
  ceilings       the measured ceilings (VALU xor+popcount, Philox, LDS increments)
  cpu_baseline   the oracle's restatement of the reference CPU loop on a
                 bounded row-shard sample of the same workload, host cores
  spectral       the same histogram through the Walsh-Hadamard path
  batched        4096 mixed-size codes batched (one launch + captured hipGraph) vs one code per
                 launch, and the super-A k=8 sweep (skip with --no-batched)
  kernels        the other hot kernels on their workloads (skip with --no-kernels)
  served         continuous-batching request mix with Poisson arrivals, p50/p99
                 latency, and a burst of small requests against one per launch
                 (skip with --no-served)
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
HBM_PEAK_TBS = 8.0  # MI355X HBM3E, published


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
    nshards_total = max(1, (1 << k) // 64)
    # calibrate: grow the leading row-shard range until it takes >= 1/8 of the
    # target, then scale it to the target (shards near the top hold the most
    # pairs, so the per-pair rate is what scales, not the shard count)
    shards = min(nshards_total, 4 * threads)
    while True:
        t0 = time.perf_counter()
        _, pairs = O.an_cpu_counts(k, A, 0, shards)
        dt = max(time.perf_counter() - t0, 1e-6)
        if dt >= target_s / 8 or shards >= nshards_total:
            break
        shards = min(nshards_total, shards * 4)
    want_pairs = pairs / dt * target_s
    lo, hi = shards, nshards_total
    N = 1 << k
    while lo < hi:  # smallest shard count whose pair total reaches want_pairs
        mid = (lo + hi) // 2
        rows = min(N, 64 * mid)
        if rows * N - rows * (rows + 1) // 2 >= want_pairs:
            hi = mid
        else:
            lo = mid + 1
    shards = lo
    t0 = time.perf_counter()
    _, pairs = O.an_cpu_counts(k, A, 0, shards)
    dt = time.perf_counter() - t0
    # the CPU Hamming loop (hamming.cpp:43-63) on the (39,32) code: all 2^32
    # data words, the exhaustive run the GPU does in half a millisecond
    hw = 1 << 32
    t0 = time.perf_counter()
    hh = O.hamming_weights(32, 0, hw)
    dt_h = time.perf_counter() - t0
    assert sum(hh) == hw
    return {
        "value": round(pairs / dt / 1e9, 4),
        "unit": "Gpairs/s",
        "cores": threads,
        "kind": "port",
        "sample": f"A={A} k={k}: row shards 0..{shards - 1} of {nshards_total} "
                  f"(64 words each vs all later words), {pairs} pairs in {dt:.2f} s on {cpu_model()}",
        "build": "gcc -O3 -march=x86-64-v3 -fopenmp: the oracle is built once where the tree is built and runs on "
                 "whichever x86-64 host runs the GPU (its model is not known at build time); its hot loop is xor + "
                 "popcount + increment, which x86-64-v3 (POPCNT, AVX2) covers",
        "hamming_39_32": {"value": round(hw / dt_h / 1e9, 4), "unit": "Gwords/s", "cores": threads,
                          "sample": f"(39,32) weights of all 2^32 data words by the restatement of "
                                    f"hamming.cpp:43-63, {dt_h:.2f} s"},
    }


def run_batched(ctx, ceil: dict | None = None, autotune: bool = True) -> dict:
    """4096 AN codes (odd A uniform in [3, 2^16), k uniform in 8..18, seed
    MIXED_CODES_SEED) -- one code per launch (cr_an_pairs per code: zero, kernel,
    copy-back, sync) against one batched launch and its captured hipGraph;
    plus the reference's published per-launch sweep shape (super-A, k=8,
    16384 codes).  Every result is checked bit-exact against the per-code run."""
    from coding_reliability_amd import workloads as WL
    out = {}
    codes = WL.mixed_codes()
    pairs = WL.pairs_of(codes)
    tune = {}
    if autotune:
        best, ms = ctx.autotune_batch(codes[::16], repeats=1)
        tune = best
        log(f"autotune on 256 codes: {best} ({ms:.1f} ms)")
    t0 = time.perf_counter()
    per = [ctx.an_pairs(A, N)[0] for A, N in codes]
    t_per = time.perf_counter() - t0
    log(f"4096 codes, one launch per code: {t_per:.3f} s")
    t0 = time.perf_counter()
    got, kms = ctx.an_batch(codes, **tune)
    t_batch = time.perf_counter() - t0
    plan = ctx.batch_plan(codes, **tune)
    plan.run()
    reps, t_graph = 3, 0.0
    for _ in range(reps):
        t0 = time.perf_counter()
        res, _ = plan.run()
        t_graph += time.perf_counter() - t0
        assert res == per, "graph replay differs from per-code launches"
    t_graph /= reps
    plan.close()
    assert got == per, "batched launch differs from per-code launches"
    log(f"4096 codes batched: {t_batch:.3f} s, graph replay: {t_graph:.3f} s")
    out["codes4096"] = {
        "workload": "4096 AN codes, odd A uniform in [3,2^16), k uniform in {8..18}, seed 20261202",
        "pairs": pairs,
        "per_launch_s": round(t_per, 4), "per_launch_gpairs": round(pairs / t_per / 1e9, 2),
        "batched_s": round(t_batch, 4), "batched_gpairs": round(pairs / t_batch / 1e9, 2),
        "batched_kernel_ms": round(kms, 3),
        "graph_s": round(t_graph, 4), "graph_gpairs": round(pairs / t_graph / 1e9, 2),
        "speedup_graph_vs_per_launch": round(t_per / t_graph, 3),
        "tuning": tune, "bit_exact": True,
    }
    # the bound batching can reach on this mix: its codes with k >= 17 hold
    # most of the pairs and each fills the device alone, so a batched launch
    # can at best run every pair at the LDS-increment ceiling; what batching
    # removes is the per-code launch and synchronisation time
    big = WL.pairs_of([(A, N) for A, N in codes if N >= 1 << 17])
    if ceil:
        t_floor = pairs / ceil["lds"]
        out["codes4096"].update({
            "share_of_pairs_k17_18": round(big / pairs, 4),
            "batched_floor_s": round(t_floor, 4),
            "speedup_bound": round(t_per / t_floor, 3),
            "bound_note": "batched_floor_s = pairs / the run's LDS-increment ceiling (one increment per pair); the "
                          "per-launch run is already within per_launch_s - batched_floor_s of it, so no batching "
                          "of this mix can exceed speedup_bound",
        })
    # the reference's published per-launch sweep: super-A k=8, 16384 codes
    sweep = WL.supera_codes(8, 16)
    spairs = WL.pairs_of(sweep)
    t0 = time.perf_counter()
    per8 = [ctx.an_pairs(A, N)[0] for A, N in sweep]
    t_per8 = time.perf_counter() - t0
    plan = ctx.batch_plan(sweep)
    plan.run()
    t0 = time.perf_counter()
    res8, _ = plan.run()
    t_g8 = time.perf_counter() - t0
    plan.close()
    assert res8 == per8
    out["supera_k8"] = {
        "workload": "super-A sweep k=8: all 16384 odd A in (2^15, 2^16)",
        "per_launch_ms_per_code": round(t_per8 * 1e3 / len(sweep), 4),
        "graph_ms_per_code": round(t_g8 * 1e3 / len(sweep), 5),
        "speedup_graph_vs_per_launch": round(t_per8 / t_g8, 1),
        "reference_per_launch_ms_per_code": 2.08,
        "bit_exact": True,
    }
    log(f"super-A k=8: per-launch {t_per8:.3f} s, graph {t_g8 * 1e3:.2f} ms")
    return out


# kernel executes `valu` VALU instructions -- the SQ_INSTS_VALU counts of its
# profile under profiles/, divided by the units of the launch --, `lds`
# histogram increments (wave-wide ds_add_u32 lanes) and `reads` syndrome-table
# lookups (ds_read_u8 lanes).  frac_valu is the VALU issue utilisation
# (executed instructions/s over the ceiling), frac_lds the share of time the
# LDS is busy (increments over the increment ceiling plus lookups over the
# table-read ceiling); frac is the larger (the two pipelines run
# concurrently), named by "bound".  `ops` is the algorithmic int32 work per
# unit for reference: xor + popcount + increment per enumerated word or pair
# (2W+1 with the high dword folded), per sampled 32-bit word 5 field lookups
# + 4 xors + 2 popcounts + 1 increment (10 + 9 + 3 + 1 for 64-bit), 40 per
# Philox block (10 per 32-bit Monte Carlo word, 20 per 64-bit), 9 per
# ANcandidates step.
KERNEL_WORK = {
    # name: (executed VALU per unit, LDS increments per unit, LDS table reads per unit,
    #        algorithmic ops per unit, profile)
    "hamming_exhaustive_39_32": (3.07, 1, 0, 3, "profiles/ham_kernel_39_32_exhaustive_grid_mc_and_72_64_mc.md"),
    "hamming_grid_39_32": (None, 1, 5, 12, "latency-bound launch"),
    "hamming_mc_39_32": (21.92, 1, 5, 22, "profiles/ham_kernel_39_32_exhaustive_grid_mc_and_72_64_mc.md"),
    "hamming_mc_72_64": (42.47, 1, 10, 43, "profiles/ham_kernel_39_32_exhaustive_grid_mc_and_72_64_mc.md"),
    "an_grid1d_ref": (3.01, 1, 0, 3, "profiles/rect_kernel_grid1d_A61_k32_m1001.md"),
    "an_mc": (16.02, 1, 0, 14, "profiles/an_mc_kernel_A61_k24_m65536.md"),
    "an_candidates": (None, 0, 0, 9, "lower bound: algorithmic steps of the maximising A"),
}


def _kernel_entry(name: str, units: float, ms: float, unit: str, ceil: dict, workload: str) -> dict:
    valu, lds, reads, ops, src = KERNEL_WORK[name]
    rate = units / (ms / 1e3)
    f_valu = rate * (valu if valu is not None else ops) / ceil["valu"]
    f_lds = rate * lds / ceil["lds"] + rate * reads / ceil["lds_read"]
    bound = "valu" if f_valu >= f_lds else "lds"
    return {"workload": workload, "kernel_ms": round(ms, 3), "rate": round(rate / 1e9, 3), "rate_unit": unit,
            "valu_per_unit": valu if valu is not None else ops, "valu_source": src if valu is not None else
            "algorithmic ops (" + src + ")", "lds_increments_per_unit": lds, "lds_table_reads_per_unit": reads,
            "ops_per_unit": ops, "achieved_tops": round(rate * ops / 1e12, 3),
            "frac_valu": round(f_valu, 4), "frac_lds": round(f_lds, 4), "bound": bound,
            "frac": round(max(f_valu, f_lds), 4)}


def _ceilings_json(ceil: dict) -> dict:
    return {"valu": round(ceil["valu"] / 1e12, 4), "lds": round(ceil["lds"] / 1e12, 4),
            "lds_read": round(ceil["lds_read"] / 1e12, 4),
            "write": round(ceil["write"] / 1e12, 4), "read": round(ceil["read"] / 1e12, 4),
            "unit": "T/s: VALU lane-instructions, LDS lane increments, LDS lane-reads from 128-byte tables, TB/s "
                    "streaming stores and loads over 256 MiB (cr_probe_ceiling, this run)",
            "probes": {kk: round(v / 1e12, 4) for kk, v in ceil["probes"].items()}}


def measure_ceilings(ctx) -> dict:
    """The device ceilings of this run (cr_probe_ceiling), each the best of
    three launches after one untimed one: VALU issue (the best of the
    xor+popcount, Philox and mixed probes, in lane-instructions/s), LDS
    increments (lane increments/s, the better of the two lane patterns), LDS
    reads from 128-byte tables (lane-reads/s) and streaming stores and loads
    over a 256 MiB buffer (bytes/s)."""
    iters = {0: 1 << 15, 1: 1 << 12, 2: 1 << 14, 3: 1 << 12, 4: 1 << 14, 5: 1 << 14, 6: 16, 7: 16, 8: 1 << 13}

    def best(probe):
        ctx.probe_ceiling(probe, iters[probe])
        return max(ctx.probe_ceiling(probe, iters[probe])[0] for _ in range(3))

    r = {p: best(p) for p in range(9)}
    return {"valu": max(r[0], r[1], r[3], r[4]), "lds": max(r[2], r[5]), "lds_read": r[8], "write": r[6],
            "read": r[7],
            "probes": {"valu_xor_popcount_8": r[0], "valu_philox": r[1], "lds_scattered_rows": r[2],
                       "valu_philox_pair_mix": r[3], "valu_xor_popcount_16": r[4], "lds_same_row": r[5],
                       "stream_write_bytes": r[6], "stream_read_bytes": r[7], "lds_table_reads": r[8]}}


def _warm(f):
    """Runs f twice (the first call loads the code object and warms the
    caches) and returns the second call's result."""
    f()
    return f()


def run_kernels(ctx, ceil: dict) -> dict:
    """The other hot kernels on their workloads, each checked for a complete
    histogram, timed with HIP events on the library's stream after one
    untimed call."""
    import coding_reliability_amd as cra
    out = {}
    k = 32
    h, ms = _warm(lambda: ctx.hamming_weights(k, cra.HAM_RANGE, 0, 1 << k))
    assert sum(h) == 1 << k
    out["hamming_exhaustive_39_32"] = _kernel_entry("hamming_exhaustive_39_32", 1 << k, ms, "Gwords/s", ceil,
                                                    "(39,32) weights of all 2^32 data words")
    h_grid, ms = _warm(lambda: ctx.hamming_weights(k, cra.HAM_STRIDE, 0, 1 << 20, stride=65537))
    assert sum(h_grid) == 1 << 20
    e = _kernel_entry("hamming_grid_39_32", 1 << 20, ms, "Gwords/s", ceil,
                      "(39,32) integer grid d = i*65537 mod 2^32, i < 2^20")
    # 2^20 words are well under a microsecond of device work: the launch is
    # latency-bound, so its floor (the same kernel over 4096 words) is given
    _, floor_ms = _warm(lambda: ctx.hamming_weights(k, cra.HAM_STRIDE, 0, 1 << 12, stride=65537))
    e["launch_floor_ms"] = round(floor_ms, 4)
    out["hamming_grid_39_32"] = e
    S = 1 << 34
    h, ms = _warm(lambda: ctx.hamming_weights(k, cra.HAM_MC, 0, S, seed=2026, stream=0))
    assert sum(h) == S
    out["hamming_mc_39_32"] = _kernel_entry("hamming_mc_39_32", S, ms, "Gwords/s", ceil, "(39,32) 2^34 Philox words")
    # the grid as a second job of the Monte Carlo launch (cr_hamming_weights_batch):
    # it then costs its words, not a launch.  The Monte Carlo launch alone
    # varies by a few percent from call to call, so five alternating pairs
    # (alone, with the grid) are timed and the median difference reported
    reqs = [dict(k=k, source=cra.HAM_STRIDE, count=1 << 20, stride=65537),
            dict(k=k, source=cra.HAM_MC, count=S, seed=2026)]
    rows, _ = _warm(lambda: ctx.hamming_weights_batch(reqs))
    assert rows == [h_grid, h]
    pairs = []
    for _ in range(5):
        alone = ctx.hamming_weights(k, cra.HAM_MC, 0, S, seed=2026, stream=0)[1]
        pairs.append((alone, ctx.hamming_weights_batch(reqs)[1]))
    diffs = sorted(b - a for a, b in pairs)
    e["with_mc_2pow34_one_launch"] = {
        "ms_median": round(sorted(b for _, b in pairs)[2], 4),
        "mc_alone_ms_median": round(sorted(a for a, _ in pairs)[2], 4),
        "grid_marginal_ms_median": round(diffs[2], 4),
        "grid_marginal_ms_range": [round(diffs[0], 4), round(diffs[-1], 4)]}
    S = 1 << 30
    h, ms = _warm(lambda: ctx.hamming_weights(64, cra.HAM_MC, 0, S, seed=2026, stream=1))
    assert sum(h) == S
    out["hamming_mc_72_64"] = _kernel_entry("hamming_mc_72_64", S, ms, "Gwords/s", ceil, "(72,64) 2^30 Philox words")
    h, ms = _warm(lambda: ctx.an_grid_ref(1, 32, 1001, 0, 61))
    pairs = sum(h)
    out["an_grid1d_ref"] = _kernel_entry("an_grid1d_ref", pairs, ms, "Gpairs/s", ceil,
                                         f"reference grid 1D, A=61, k=32, m=1001 (38-bit codewords, {pairs} pairs)")
    m = 1 << 16
    h, ms = _warm(lambda: ctx.an_mc(61, 24, m, 16384))
    assert sum(h) == (1 << 24) * m
    out["an_mc"] = _kernel_entry("an_mc", (1 << 24) * m, ms, "Gpairs/s", ceil,
                                 "AN Monte Carlo, A=61, k=24, 2^16 samples")
    best, cands, ms = _warm(lambda: ctx.an_candidates(24, 16))
    # pruning skips most (A, d) steps, so the work is counted as the steps the
    # candidates (A attaining the maximum, never pruned) must run in full: a
    # lower bound of the executed steps, hence a lower bound of the fraction
    steps = len(cands) * ((1 << 24) - 1)
    e = _kernel_entry("an_candidates", steps, ms, "Gsteps/s", ceil,
                      "ANcandidates cell k=24, |A|=16 (16384 odd A); rate = steps of the A attaining the maximum "
                      "(each runs all 2^24-1 steps) per second, a lower bound")
    e["best"], e["candidates"] = best, len(cands)
    out["an_candidates"] = e
    return out


# Timings the reference records in its own result files (distance_distribution/
# results/...), one K80-class GPU unless stated: (seconds, source).
REFERENCE_RECORDED = {
    "an_A61_k24": (1306.17, "runtimes/4/ancoding_gpu_A61_n24_h6.csv:34"),
    "an_A32771_k24": (8058.97, "As/ancoding_gpu_A32771_n24_h16.csv:44"),
    "grid1d_A61_k32_m1001": (107.01, "runtimes/4/ancoding_grid1d_gpu_m1001_A61_n32_h6.csv:42"),
    "grid1d_A61_k40_m101": (33952.62, "runtimes/4/ancoding_grid1d_gpu_m101_A61_n40_h6.csv:50"),
    "supera_k16_sweep": (1933.3, "superAs/superA_n16_h16.txt:37 (sum of 16384 per-code kernel times)"),
    "supera_k8_sweep": (34.14, "superAs/superA_n8_h16.txt:34"),
}


def run_spectral(ctx, A: int, k: int, raw_pairs: list, pair_kernel_s: float, ceil: dict | None = None) -> dict:
    """The headline histogram again, from the code's Walsh-Hadamard spectrum
    (cr_an_pairs_spectral, the path cr_an_pairs takes for this code): one
    untimed call, one timed call, checked equal to the pair kernel's."""
    ctx.an_pairs_spectral(A, k)
    t0 = time.perf_counter()
    raw, ms = ctx.an_pairs_spectral(A, k)
    dt = time.perf_counter() - t0
    assert raw == raw_pairs, "spectral histogram differs from the pair kernel's"
    N = 1 << k
    Ao = A // (A & -A)  # odd part: the same distances
    n = max((Ao * (N - 1)).bit_length(), k)
    # memory roofline of the passes: each transformed element is written once
    # by the first pass and read once by the last: int16 for k = 24 (two
    # transforms per first-pass workgroup), int32 for 17 <= k <= 23
    elements = 1 << n
    bytes_per_element = 4 if (k == 24 and n > k) else (8 if 17 <= k <= 24 else None)
    memory = None
    if bytes_per_element:
        achieved = elements * bytes_per_element / (ms / 1e3) / 1e12
        half = elements * bytes_per_element / 2
        floor_s = half / ceil["write"] + half / ceil["read"] if ceil else None
        memory = {"bytes_per_element": bytes_per_element, "achieved": round(achieved, 3), "peak": HBM_PEAK_TBS,
                  "unit": "TB/s", "frac": round(achieved / HBM_PEAK_TBS, 4)}
        if floor_s:
            memory.update({"measured_stream_floor_s": round(floor_s, 4),
                           "frac_of_measured_stream": round(floor_s / (ms / 1e3), 4),
                           "note": "floor = the bytes written at the streaming-store ceiling plus the bytes read at "
                                   "the streaming-load ceiling, both measured in this run over 256 MiB"})
    return {
        "method": "Walsh-Hadamard spectrum of the code: 2^(n-k) transforms of length 2^k, exact integer sums, "
                  "Krawtchouk transform on the host",
        "seconds": round(dt, 4),
        "device_seconds": round(ms / 1e3, 4),
        "equivalent_gpairs_per_s": round(N * (N - 1) / 2 / dt / 1e9, 1),
        "speedup_vs_pair_kernel": round(pair_kernel_s / dt, 2),
        "transforms": 1 << (n - k),
        "transform_length": N,
        "memory_roofline": memory,
        "bit_exact_vs_pair_kernel": True,
    }


def run_vs_reference(ctx) -> dict:
    """The workloads the reference records timings for, on this device, each
    checked (complete histograms; batched = per-code results)."""
    from coding_reliability_amd import workloads as WL
    out = {}

    def entry(key, secs, extra=None, out_key=None):
        ref, src = REFERENCE_RECORDED[key]
        e = {"seconds": round(secs, 4), "reference_seconds": ref, "reference_source": src,
             "speedup_vs_reference": round(ref / secs, 1)}
        e.update(extra or {})
        out[out_key or key] = e
        log(f"{out_key or key}: {secs:.4f} s (reference {ref} s)")

    for A, key in ((61, "an_A61_k24"), (32771, "an_A32771_k24")):
        N = 1 << 24
        raw, ms = ctx.an_pairs(A, N)
        assert sum(raw) == N * (N - 1) // 2
        entry(key, ms / 1e3, {"gpairs_per_s": round(N * (N - 1) / 2 / (ms / 1e3) / 1e9, 1), "method": "pair kernel"})
        # the same histogram from the code's spectrum (what cr_an_pairs picks)
        ctx.an_pairs_spectral(A, 24)
        t0 = time.perf_counter()
        raw_s, ms_s = ctx.an_pairs_spectral(A, 24)
        dt = time.perf_counter() - t0
        assert raw_s == raw
        entry(key, dt, {"device_seconds": round(ms_s / 1e3, 4),
                        "method": "Walsh-Hadamard spectrum (cr_an_pairs_spectral), bit-exact against the pair kernel"},
              out_key=key + "_spectral")
    # the exact k = 32 distribution the grid above approximates (spectral path)
    ctx.an_pairs_spectral(3, 32)
    t0 = time.perf_counter()
    raw32, ms32 = ctx.an_pairs_spectral(61, 32)
    dt = time.perf_counter() - t0
    exact = [1 << 32] + [2 * x for x in raw32[1:]]
    assert sum(exact) == 1 << 64
    grids = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                                         "an_grid1d_runs.json")))["runs"]
    errs = {}
    for m in (1001, 10000):
        g = [r for r in grids if r["n"] == 32 and r["A"] == 61 and r["m"] == m][0]["counts"]
        errs[f"recorded_grid_m{m}_max_rel_error"] = round(max(abs(g[d] / exact[d] - 1) for d in range(1, 38)
                                                              if exact[d] > 0), 6)
    entry("grid1d_A61_k32_m1001", dt, dict(errs, device_seconds=round(ms32 / 1e3, 4),
                                           method="exact distribution (spectral path, 64 transforms of 2^32) "
                                                  "against the reference's m=1001 grid approximation"),
          out_key="exact_A61_k32")
    for k, m, key in ((32, 1001, "grid1d_A61_k32_m1001"), (40, 101, "grid1d_A61_k40_m101")):
        h, ms = ctx.an_grid_ref(1, k, m, 0, 61)
        entry(key, ms / 1e3, {"gpairs_per_s": round(sum(h) / (ms / 1e3) / 1e9, 1)})
    for k, key in ((16, "supera_k16_sweep"), (8, "supera_k8_sweep")):
        sweep = WL.supera_codes(k, 16)
        plan = ctx.batch_plan(sweep)
        plan.run()
        t0 = time.perf_counter()
        res, _ = plan.run()
        dt = time.perf_counter() - t0
        plan.close()
        assert all(sum(r) == N * (N - 1) // 2 for r, (A, N) in zip(res, sweep))
        entry(key, dt, {"codes": len(sweep), "gpairs_per_s": round(WL.pairs_of(sweep) / dt / 1e9, 1)})
    return out


def run_served(ctx, njobs: int, rate: float, seed: int = 7, max_wait_s: float = 300.0) -> dict:
    """Served mix: Poisson arrivals (rate jobs/s) of a uniform mix of exhaustive AN
    (k in 18..24, odd 16-bit A), integer-grid AN (k=32, stride 4097, odd 16-bit
    A) and Monte Carlo (72,64) Extended Hamming (2^30 samples) requests."""
    import numpy as np
    import coding_reliability_amd as cra
    g = np.random.Generator(np.random.Philox(seed))
    jobs, arrivals, t = [], [], 0.0
    for i in range(njobs):
        kind = int(g.integers(0, 3))
        A = int(2 * g.integers(1 << 14, 1 << 15) + 1)
        if kind == 0:
            jobs.append(dict(kind=cra.JOB_AN_EXHAUSTIVE, k=int(g.integers(18, 25)), A=A))
        elif kind == 1:
            jobs.append(dict(kind=cra.JOB_AN_GRID, k=32, A=A, stride=4097))
        else:
            jobs.append(dict(kind=cra.JOB_HAMMING_MC, k=64, samples=1 << 30, seed=seed, stream=i))
        t += float(g.exponential(1.0 / rate))
        arrivals.append(t)
    srv = cra.Server(ctx, iteration_ms=20)
    t0 = srv.now()
    ids = []
    for j, at in zip(jobs, arrivals):
        delay = at - (srv.now() - t0)
        if delay > 0:
            time.sleep(delay)
        ids.append(srv.submit(**j))
    deadline = time.perf_counter() + max_wait_s
    for rid in ids:
        while True:
            try:
                srv.wait(rid, timeout_ms=15000)
                break
            except cra.CRError as e:
                if e.status != 9:  # CR_ERR_BUSY: still running
                    raise
                st = srv.stats()
                log(f"served: waiting for request {rid}: {st['completed']}/{len(ids)} done, "
                    f"{st['active']} active, {st['queued']} queued, {st['iterations']} iterations")
                if time.perf_counter() > deadline:
                    srv.close()
                    raise RuntimeError(f"served mix unfinished after {max_wait_s} s")
        log(f"served: request {rid} done")
    infos = [srv.result(rid)[1] for rid in ids]
    st = srv.stats()
    srv.close()
    lat = sorted(i["finish_s"] - i["submit_s"] for i in infos)
    makespan = max(i["finish_s"] for i in infos) - min(i["submit_s"] for i in infos)
    an_pairs = sum(i["units"] for j, i in zip(jobs, infos) if j["kind"] != cra.JOB_HAMMING_MC)
    samples = sum(i["units"] for j, i in zip(jobs, infos) if j["kind"] == cra.JOB_HAMMING_MC)
    pct = lambda q: lat[min(len(lat) - 1, int(round(q * (len(lat) - 1))))]
    return {
        "workload": f"{njobs} requests, Poisson rate {rate}/s, mix of exhaustive AN k 18..24, grid AN k=32 "
                    f"stride 4097, MC (72,64) 2^30 samples; seed {seed}",
        "latency_p50_s": round(pct(0.5), 3), "latency_p99_s": round(pct(0.99), 3),
        "makespan_s": round(makespan, 3),
        "an_gpairs_per_s": round(an_pairs / makespan / 1e9, 2),
        "mc_gsamples_per_s": round(samples / makespan / 1e9, 3),
        "iterations": st["iterations"], "device_busy_frac": round(st["busy_device_ms"] / 1e3 / makespan, 3),
    }


def run_served_small(ctx, njobs: int = 2048, seed: int = 11) -> dict:
    """A burst of small exhaustive requests (k uniform in 8..14, odd 16-bit A)
    through the request loop, against the same requests one per launch; every
    served result equals its one-per-launch result."""
    import numpy as np
    import coding_reliability_amd as cra
    g = np.random.Generator(np.random.Philox(seed))
    jobs = [(int(2 * g.integers(1 << 14, 1 << 15) + 1), int(g.integers(8, 15))) for _ in range(njobs)]
    t0 = time.perf_counter()
    per = [ctx.an_pairs(A, 1 << k)[0] for A, k in jobs]
    t_per = time.perf_counter() - t0
    srv = cra.Server(ctx, iteration_ms=5, max_active=1024)
    t0 = time.perf_counter()
    ids = [srv.submit(kind=cra.JOB_AN_EXHAUSTIVE, k=k, A=A) for A, k in jobs]
    for rid in ids:
        srv.wait(rid, timeout_ms=60000)
    t_srv = time.perf_counter() - t0
    got = [srv.result(rid)[0] for rid in ids]
    st = srv.stats()
    srv.close()
    assert all(g_[: len(p)] == p[: len(g_)] for g_, p in zip(got, per)), "served result differs"
    pairs = sum((1 << k) * ((1 << k) - 1) // 2 for _, k in jobs)
    log(f"served small burst: {t_srv:.3f} s vs one per launch {t_per:.3f} s")
    return {"workload": f"{njobs} exhaustive AN requests at once, k uniform in 8..14, odd 16-bit A, seed {seed}",
            "served_s": round(t_srv, 4), "per_launch_s": round(t_per, 4),
            "served_gpairs_per_s": round(pairs / t_srv / 1e9, 2), "per_launch_gpairs_per_s": round(pairs / t_per / 1e9, 2),
            "speedup_served_vs_per_launch": round(t_per / t_srv, 1), "iterations": st["iterations"], "bit_exact": True}


def traffic_probe(A: int, k: int) -> int:
    """For rocprofv3 --pmc runs: one device copy of exactly 1 GiB (torch, to
    scale FETCH_SIZE / WRITE_SIZE into bytes) followed by one pair-kernel
    launch of the headline problem (A, k)."""
    import torch
    import coding_reliability_amd as cra
    x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    log("copy of 1 GiB done")
    ctx = cra.Context(0)
    raw, ms = ctx.an_pairs(A, 1 << k)
    log(f"pair kernel A={A} k={k}: {ms:.3f} ms, sum {sum(raw)}")
    ctx.close()
    return 0


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pair_kernel_A58659_k24_hbm_traffic.txt")


def measured_traffic(A: int, k: int):
    """HBM bytes per pair-kernel launch from the committed FETCH_SIZE /
    WRITE_SIZE summary (counter values scaled by the same counters read on a
    1 GiB copy), or None when no summary exists for this (A, k)."""
    try:
        kv = {}
        for line in open(TRAFFIC_FILE):
            if ":" in line and not line.startswith("#"):
                key, val = line.split(":", 1)
                kv[key.strip()] = val.strip()
        if int(kv["A"]) != A or int(kv["k"]) != k:
            return None
        return int(float(kv["traffic_bytes"]))
    except (OSError, KeyError, ValueError):
        return None


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
    ap.add_argument("--no-batched", action="store_true", help="skip the batched-vs-per-launch comparison")
    ap.add_argument("--no-spectral", action="store_true", help="skip the spectral computation of the headline histogram")
    ap.add_argument("--spectral-only", nargs=2, type=int, metavar=("A", "K"),
                    help="one cr_an_pairs_spectral call for (A, K), checked to sum to C(2^K, 2) (for counter runs)")
    ap.add_argument("--no-served", action="store_true", help="skip the served request mix")
    ap.add_argument("--served-only", action="store_true", help="only the served request mix")
    ap.add_argument("--ceilings-only", action="store_true", help="only the device ceiling probes")
    ap.add_argument("--no-kernels", action="store_true", help="skip the other kernels' measurements")
    ap.add_argument("--kernels-only", action="store_true", help="only the other kernels (for counter runs)")
    ap.add_argument("--vs-reference", action="store_true",
                    help="only the workloads the reference records timings for")
    ap.add_argument("--served-jobs", type=int, default=24)
    ap.add_argument("--served-rate", type=float, default=2.0)
    ap.add_argument("--traffic-probe", action="store_true", help="copy + one pair launch, for PMC counter runs")
    args = ap.parse_args()
    if args.traffic_probe:
        return traffic_probe(args.A, args.k)

    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus must be 1 "
                                   "(the pair space would shard by work items across GPUs; not built)"}))
        return 2

    import coding_reliability_amd as cra

    A, k = args.A, args.k
    N = 1 << k
    pairs_per_step = N * (N - 1) // 2
    ctx = cra.Context(0)
    info = ctx.device_info()
    log(f"device {info['name']} {info['arch']} CUs={info['compute_units']} clock={info['max_clock_mhz']} MHz")
    tune = dict(rows_per_lane=args.rows_per_lane, chunk_blocks=args.chunk_blocks)
    ceil = measure_ceilings(ctx)
    log("ceilings: " + ", ".join(f"{kk} {v / 1e12:.3f} T/s" for kk, v in ceil["probes"].items()))
    if args.ceilings_only:
        print(json.dumps({"ceilings": _ceilings_json(ceil)}), flush=True)
        ctx.close()
        return 0
    if args.kernels_only:
        print(json.dumps({"kernels": run_kernels(ctx, ceil)}), flush=True)
        ctx.close()
        return 0
    if args.served_only:
        served = run_served(ctx, args.served_jobs, args.served_rate)
        served["small_burst"] = run_served_small(ctx)
        print(json.dumps({"served": served}), flush=True)
        ctx.close()
        return 0
    if args.spectral_only:
        sA, sk = args.spectral_only
        raw, ms = ctx.an_pairs_spectral(sA, sk)
        assert sum(raw) == (1 << sk) * ((1 << sk) - 1) // 2
        print(json.dumps({"spectral": {"A": sA, "k": sk, "device_ms": round(ms, 3)}}), flush=True)
        ctx.close()
        return 0
    if args.vs_reference:
        print(json.dumps({"vs_reference": run_vs_reference(ctx)}), flush=True)
        ctx.close()
        return 0

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

    # Rooflines of the pair kernel, each against a ceiling measured in this
    # run.  Per pair it executes one xor, one popcount (for 33..64-bit
    # codewords whose column high dword is constant over 2^32/A columns, the
    # high-dword distance is a per-segment constant folded into the
    # popcount) and one histogram increment: 3 int32 VALU ops (the 2W+1 of
    # W = 1; the increment costs its LDS address computation) and one
    # wave-wide ds_add_u32 lane.  The LDS increment rate binds: one
    # increment per pair runs at 16 lanes per clock per CU against 64 VALU
    # lanes, so 3 VALU per pair cannot reach the VALU ceiling first.
    width_bits = ((N - 1) * A).bit_length()
    W = 1 if width_bits <= 32 else 2
    ops_per_pair = 3
    cus = info["compute_units"]
    nominal = cus * LANES_PER_CU_CLK * PEAK_CLOCK_GHZ * 1e9 / 1e12  # T lane-ops/s
    pair_rate = pairs_per_step / avg_kernel_s
    roofline = {
        "bound": "lds_atomic",
        "achieved": round(pair_rate / 1e12, 4),
        "peak": round(ceil["lds"] / 1e12, 4),
        "unit": "T LDS increments/s (one per pair)",
        "frac": round(pair_rate / ceil["lds"], 4),
        "traffic": measured_traffic(A, k),
        "kernel": "pair_kernel<R>",
        "kernel_ms": round(avg_kernel_s * 1e3, 3),
        "peak_source": "cr_probe_ceiling(2) in this run: wave-wide ds_add_u32 to lane-private LDS counters, "
                       "the pair kernel's histogram layout",
        "valu": {
            "achieved": round(pair_rate * ops_per_pair / 1e12, 3),
            "peak": round(ceil["valu"] / 1e12, 3),
            "unit": "T int32 lane-ops/s",
            "frac": round(pair_rate * ops_per_pair / ceil["valu"], 4),
            "ops_per_pair": ops_per_pair,
            "ops_per_pair_2W_plus_1": 2 * W + 1,
            "peak_source": "cr_probe_ceiling in this run: the best VALU issue rate of the xor+popcount, Philox "
                           "and mixed probes",
            "nominal_peak": round(nominal, 3),
            "nominal_frac": round(pair_rate * ops_per_pair / 1e12 / nominal, 4),
            "nominal_note": f"{cus} CU x 64 lanes x {PEAK_CLOCK_GHZ} GHz, one VALU instruction per CU per clock",
        },
    }

    spectral = None
    if not args.no_spectral:
        log("spectral ...")
        spectral = run_spectral(ctx, A, k, raw, avg_kernel_s, ceil)
        log(f"spectral: {spectral['seconds']} s, {spectral['equivalent_gpairs_per_s']} Gpairs/s equivalent")

    batched = None
    if not args.no_batched:
        log("batched vs per-launch ...")
        batched = run_batched(ctx, ceil)

    kernels = None
    if not args.no_kernels:
        log("other kernels ...")
        kernels = run_kernels(ctx, ceil)
        for name, e in kernels.items():
            log(f"{name}: {e['rate']} {e['rate_unit']}, frac {e['frac']}")

    served = None
    if not args.no_served:
        log("served mix ...")
        served = run_served(ctx, args.served_jobs, args.served_rate)
        served["small_burst"] = run_served_small(ctx)

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        cpu = cpu_baseline(A, k, args.cpu_seconds)
        log(f"cpu baseline {cpu['value']} Gpairs/s on {cpu['cores']} threads")

    out = {
        "metric": "pair-comparisons/sec (Gpairs/s) + %VALU peak at k=24; batched vs per-launch speedup",
        "value": round(gpairs, 3),
        "unit": "Gpairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "dtype_note": "exact integer arithmetic: codewords of up to 64 bits in 32-bit VALU lanes (the high dword's "
                      "distance folded per column segment), histogram counts uint64, bit-exact",
        "data": "synthetic (codewords c=A*d generated on the GPU)",
        "config": {"workload": f"exhaustive AN histogram, A={A}, k={k}, {width_bits}-bit codewords, "
                               f"all C(2^{k},2) pairs, {width_bits + 1} bins",
                   "A": A, "k": k, "pairs_per_step": pairs_per_step, "parallelism": "single GPU"},
        "roofline": roofline,
        "ceilings": _ceilings_json(ceil),
        "spectral": spectral,
        "cpu_baseline": cpu,
        "batched": batched,
        "kernels": kernels,
        "served": served,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
