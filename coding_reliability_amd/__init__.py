"""MI355X engine for the distance-distribution path of brics-db/coding_reliability.

Python host API over the C ABI of libcodrel.so (hand-written gfx950 HIP
kernels).  The reference-compatible C++ entry points (run_ancoding, ...) are
exported by the same library; see include/cr_reference_api.h.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import (CRError, check, cr_an_code, cr_device_info, cr_ham_request, cr_job, cr_job_result,
                   cr_pair_tuning, cr_server_config, cr_server_stats, cr_supera_request, lib)

JOB_AN_EXHAUSTIVE, JOB_AN_GRID, JOB_HAMMING_MC = 0, 1, 2

HAM_RANGE, HAM_STRIDE, HAM_LIST, HAM_MC = 0, 1, 2, 3

__all__ = ["Context", "CRError", "an_bins", "ceil_log2", "grid_report", "hamming_parity_bits", "hamming_N",
           "hamming_report", "hamming_sdc_table", "HAM_RANGE", "HAM_STRIDE", "HAM_LIST", "HAM_MC"]


def hamming_parity_bits(k: int) -> int:
    return int(lib().cr_hamming_parity_bits(k))


def hamming_N(k: int) -> int:
    """Extended codeword length (data + parity + overall parity)."""
    return k + hamming_parity_bits(k) + 1


def hamming_report(k: int, weights: list[int], with_1bit: bool) -> list[int]:
    """Reported counts of run_hamming (hamming.cu:182-208) from a non-extended
    weight histogram (uint128 values)."""
    N = hamming_N(k)
    w = np.zeros(max(N + 1, len(weights)), dtype=np.uint64)
    w[: len(weights)] = weights
    out = np.zeros(2 * (N + 1), dtype=np.uint64)
    check(lib().cr_hamming_report(k, _u64p(w), int(with_1bit), _u64p(out)))
    return _u128(out)


def hamming_sdc_table(k: int, method: int = 2, ctx: "Context | None" = None) -> tuple[str, list[int]]:
    """SDC table text of the (N,k) Extended Hamming code in the legacy table
    layout (legacy/hamming_count_undetectable_errors.cpp:226-339) and the
    non-extended weight distribution it was built from.  method 1 enumerates
    all 2^k words on the device (k <= 40, needs ctx), 2 uses the MacWilliams
    identity, 0 picks."""
    N = hamming_N(k)
    w = np.zeros(2 * (N + 1), dtype=np.uint64)
    txt = C.c_void_p()
    check(lib().cr_hamming_sdc_table(ctx.handle if ctx is not None else None, k, method, _u64p(w), N + 1,
                                      C.byref(txt)))
    try:
        s = C.string_at(txt).decode()
    finally:
        lib().cr_free(txt)
    return s, _u128(w)


def _u64p(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


def _u128(buf) -> list[int]:
    b = buf.reshape(-1, 2)
    return [int(lo) | (int(hi) << 64) for lo, hi in b]


def grid_report(gdim: int, n: int, m1: int, m2: int, hist: list[int], count_counts: int) -> list[int]:
    """Scaled counts of run_ancoding_grid (an_coding_grid.cu:316-322) from a raw
    grid histogram (uint128 values)."""
    h = np.zeros(max(count_counts, len(hist)), dtype=np.uint64)
    h[: len(hist)] = hist
    out = np.zeros(2 * count_counts, dtype=np.uint64)
    check(lib().cr_an_grid_report(gdim, n, m1, m2, _u64p(h), count_counts, _u64p(out)))
    return _u128(out)


def ceil_log2(A: int) -> int:
    """h = ceil(log(A)/log(2.0)) in double, as the reference drivers compute it
    (distance_distribution/src/an_coding.cu:185)."""
    import math
    return int(math.ceil(math.log(A) / math.log(2.0))) if A > 1 else 0


def an_bins(A: int, num_words: int) -> int:
    """Histogram bins needed for codewords A*i, i < num_words."""
    return max(((num_words - 1) * A).bit_length(), 0) + 1


class Context:
    """One device, one HIP stream (cr_context)."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        check(lib().cr_context_create(device, C.byref(self._h)))

    def close(self):
        if self._h:
            lib().cr_context_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def device_info(self) -> dict:
        info = cr_device_info()
        check(lib().cr_device_query(self._h, C.byref(info)))
        return dict(name=info.name.decode(), arch=info.arch.decode(), compute_units=info.compute_units,
                    max_clock_mhz=info.max_clock_mhz, hbm_bytes=info.hbm_bytes)

    def device_memory(self) -> tuple[int, int]:
        """(free, total) device memory in bytes."""
        f, t = C.c_uint64(0), C.c_uint64(0)
        check(lib().cr_device_memory(self._h, C.byref(f), C.byref(t)))
        return f.value, t.value

    (PROBE_VALU, PROBE_PHILOX, PROBE_LDS, PROBE_MIXED, PROBE_VALU16, PROBE_LDS_ROW, PROBE_WRITE, PROBE_READ,
     PROBE_LDS_TABLE) = range(9)

    def probe_ceiling(self, probe: int, iters: int = 1 << 14) -> tuple[float, float]:
        """(units per second, device ms) of a ceiling probe (cr_probe_ceiling):
        VALU lane-instructions (probes 0, 1, 3, 4), LDS lane increments (2, 5),
        bytes (6, 7: streaming stores / loads over 256 MiB, iters passes) or LDS
        lane-reads from 128-byte tables (8)."""
        rate, ms = C.c_double(0), C.c_double(0)
        check(lib().cr_probe_ceiling(self._h, probe, iters, C.byref(rate), C.byref(ms)))
        return rate.value, ms.value

    def an_pairs(self, A: int, num_words: int, nbins: int | None = None, rows_per_lane: int = 0,
                 chunk_blocks: int = 0, waves: int = 0) -> tuple[list[int], float]:
        """Raw unordered-pair histogram of codewords A*i, i < num_words, and the
        pair-kernel time in ms."""
        if nbins is None:
            nbins = an_bins(A, num_words)
        out = np.zeros(nbins, dtype=np.uint64)
        ms = C.c_double(0)
        t = cr_pair_tuning(rows_per_lane, chunk_blocks, waves, 0)
        check(lib().cr_an_pairs_tuned(self._h, A, num_words, _u64p(out), nbins, C.byref(t), C.byref(ms)))
        return [int(x) for x in out], ms.value

    def an_pairs_auto(self, A: int, num_words: int, nbins: int | None = None) -> tuple[list[int], float]:
        """The same histogram as an_pairs through cr_an_pairs, which picks the
        spectral path for 2^16 / 2^24 words when it is estimated cheaper."""
        if nbins is None:
            nbins = an_bins(A, num_words)
        out = np.zeros(nbins, dtype=np.uint64)
        ms = C.c_double(0)
        check(lib().cr_an_pairs(self._h, A, num_words, _u64p(out), nbins, C.byref(ms)))
        return [int(x) for x in out], ms.value

    def an_pairs_spectral(self, A: int, k: int, nbins: int | None = None) -> tuple[list[int], float]:
        """Same raw histogram as an_pairs(A, 2^k), from the Walsh-Hadamard
        spectrum of the code (cr_an_pairs_spectral), and the device time in ms."""
        if nbins is None:
            nbins = an_bins(A, 1 << k)
        out = np.zeros(nbins, dtype=np.uint64)
        ms = C.c_double(0)
        check(lib().cr_an_pairs_spectral(self._h, A, k, _u64p(out), nbins, C.byref(ms)))
        return [int(x) for x in out], ms.value

    # --- batched exhaustive AN -------------------------------------------------
    @staticmethod
    def _codes(codes):
        arr = (cr_an_code * len(codes))()
        for i, (A, N) in enumerate(codes):
            arr[i].A, arr[i].num_words = A, N
        return arr

    @staticmethod
    def _stride(codes):
        return max(an_bins(A, N) for A, N in codes)

    def an_batch(self, codes, rows_per_lane: int = 0, items_per_wave: int = 0) -> tuple[list[list[int]], float]:
        """Raw histograms of many (A, num_words) codes in one launch."""
        stride = self._stride(codes)
        out = np.zeros(len(codes) * stride, dtype=np.uint64)
        ms = C.c_double(0)
        t = cr_pair_tuning(rows_per_lane, 0, 0, items_per_wave)
        check(lib().cr_an_batch(self._h, self._codes(codes), len(codes), _u64p(out), stride, C.byref(t),
                                C.byref(ms)))
        o = out.reshape(len(codes), stride)
        return [[int(x) for x in row[: an_bins(A, N)]] for row, (A, N) in zip(o, codes)], ms.value

    def batch_plan(self, codes, rows_per_lane: int = 0, items_per_wave: int = 0) -> "BatchPlan":
        return BatchPlan(self, codes, rows_per_lane, items_per_wave)

    def autotune_batch(self, codes, repeats: int = 1) -> tuple[dict, float]:
        best = cr_pair_tuning()
        ms = C.c_double(0)
        check(lib().cr_autotune_batch(self._h, self._codes(codes), len(codes), repeats, C.byref(best), C.byref(ms)))
        return dict(rows_per_lane=best.rows_per_lane, items_per_wave=best.items_per_wave), ms.value

    def supera_search(self, style: int, k: int, h_bits: int = 0, A_start: int = 0, A_end: int = 0,
                      As=None, with_grid: int = 0, m1: int = 0, m2: int = 0, verbose: int = 0) -> tuple[str, float]:
        """Super-A search output text (style 0: distance_distribution stdout,
        style 1: super_a CSV file body) and the device seconds it took."""
        lst = np.ascontiguousarray(As if As is not None else [], dtype=np.uint64)
        q = cr_supera_request(style, k, h_bits, with_grid, A_start, A_end,
                              _u64p(lst) if len(lst) else None, len(lst), verbose, m1, m2)
        txt = C.c_void_p()
        secs = C.c_double(0)
        check(lib().cr_supera_search(self._h, C.byref(q), C.byref(txt), C.byref(secs)))
        try:
            s = C.string_at(txt).decode()
        finally:
            lib().cr_free(txt)
        return s, secs.value

    def hamming_sdc_table(self, k: int, method: int = 0) -> tuple[str, list[int]]:
        """hamming_sdc_table on this context's device (method 0/1)."""
        return hamming_sdc_table(k, method, self)

    def an_candidates(self, k: int, abits: int) -> tuple[int, list[int], float]:
        """(best minimum NAF weight, sorted A attaining it, kernel ms) for one
        (data width, A width) cell of ANcandidates."""
        cap = 1 << 20
        best = np.zeros(1, dtype=np.uint64)
        n = np.zeros(1, dtype=np.uint64)
        cands = np.zeros(cap, dtype=np.uint64)
        ms = C.c_double(0)
        check(lib().cr_an_candidates(self._h, k, abits, _u64p(best), _u64p(cands), cap, _u64p(n), C.byref(ms)))
        return int(best[0]), [int(x) for x in cands[: min(int(n[0]), cap)]], ms.value

    def ancandidates_tsv(self, dmin: int, dmax: int, amin: int, amax: int) -> str:
        """Output of the ANcandidates CLI for -d dmin -D dmax -a amin -A amax
        (ANcandidates.cpp:246-299), candidates sorted."""
        lines = [f"# bit width data = [{dmin},{dmax}]", f"# bit width A = [{amin},{amax}]",
                 "# n=|C|\tk=|D|\t|A|\t|SDR|\t|candidates|\tcandidates"]
        nd, na = dmax - dmin + 1, amax - amin + 1
        for cell in range(nd * na):
            k = cell % nd + dmin
            a = cell // nd + amin
            best, cands, _ = self.an_candidates(k, a)
            lines.append(f"{k + a}\t{k}\t{a}\t{best}\t{len(cands)}\t" + ",".join(str(x) for x in cands))
        return "\n".join(lines) + "\n"

    def hamming_weights(self, k: int, source: int = HAM_RANGE, first: int = 0, count: int | None = None,
                        stride: int = 0, words=None, seed: int = 0, stream: int = 0) -> tuple[list[int], float]:
        """Non-extended codeword weight histogram over a stream of data words."""
        lst = None
        if source == HAM_LIST:
            lst = np.ascontiguousarray(words, dtype=np.uint64)
            count = len(lst)
        if count is None:
            count = 1 << k
        req = cr_ham_request(k, source, first, count, stride, _u64p(lst) if lst is not None else None, seed,
                             stream, 0)
        nbins = hamming_N(k)
        out = np.zeros(nbins, dtype=np.uint64)
        ms = C.c_double(0)
        check(lib().cr_hamming_weights(self._h, C.byref(req), _u64p(out), nbins, C.byref(ms)))
        return [int(x) for x in out], ms.value

    def hamming_weights_batch(self, requests) -> tuple[list[list[int]], float]:
        """hamming_weights of several requests in one call
        (cr_hamming_weights_batch): each request is a dict of hamming_weights'
        keyword arguments (k required); row i has hamming_N(k_i) bins."""
        reqs = (cr_ham_request * len(requests))()
        keep = []
        for r, q in zip(reqs, requests):
            k, source = q["k"], q.get("source", HAM_RANGE)
            lst, count = None, q.get("count")
            if source == HAM_LIST:
                lst = np.ascontiguousarray(q["words"], dtype=np.uint64)
                keep.append(lst)
                count = len(lst)
            if count is None:
                count = 1 << k
            r.k, r.source, r.first, r.count, r.stride = k, source, q.get("first", 0), count, q.get("stride", 0)
            r.list = _u64p(lst) if lst is not None else None
            r.seed, r.stream = q.get("seed", 0), q.get("stream", 0)
        nbins = max((hamming_N(q["k"]) for q in requests), default=1)
        out = np.zeros((len(requests), nbins), dtype=np.uint64)
        ms = C.c_double(0)
        check(lib().cr_hamming_weights_batch(self._h, reqs, len(requests), _u64p(out), nbins, C.byref(ms)))
        return [[int(x) for x in row[: hamming_N(q["k"])]] for row, q in zip(out, requests)], ms.value

    def an_mc(self, A: int, n: int, m: int, seed: int, nbins: int = 65) -> tuple[list[int], float]:
        """Monte Carlo (word, sample) histogram of run_ancoding_mc with Philox samples."""
        out = np.zeros(nbins, dtype=np.uint64)
        ms = C.c_double(0)
        check(lib().cr_an_mc(self._h, A, n, m, seed, _u64p(out), nbins, C.byref(ms)))
        return [int(x) for x in out], ms.value

    def an_mc_ranges(self, A: int, n: int, bounds, seed: int, nbins: int = 65) -> tuple[list[list[int]], float]:
        """Histograms of the sample ranges [bounds[i], bounds[i+1]) of an_mc
        (same A, n, seed) in one launch (cr_an_mc_ranges)."""
        b = np.ascontiguousarray(bounds, dtype=np.uint64)
        count = len(b) - 1
        out = np.zeros((max(count, 0), nbins), dtype=np.uint64)
        ms = C.c_double(0)
        check(lib().cr_an_mc_ranges(self._h, A, n, _u64p(b), count, seed, _u64p(out), nbins, C.byref(ms)))
        return [[int(x) for x in row] for row in out], ms.value

    def an_mc_prefixes(self, A: int, n: int, ms_list, seed: int, nbins: int = 65) -> tuple[list[list[int]], float]:
        """an_mc histograms for each sample count in ms_list (non-decreasing),
        every sample drawn once: the cumulative sums of an_mc_ranges."""
        rows, ms = self.an_mc_ranges(A, n, [0, *ms_list], seed, nbins)
        acc, out = [0] * nbins, []
        for r in rows:
            acc = [a + x for a, x in zip(acc, r)]
            out.append(list(acc))
        return out, ms

    def an_grid_ref(self, gdim: int, n: int, m1: int, m2: int, A: int, nbins: int = 65) -> tuple[list[int], float]:
        """Kernel-level histogram of the reference grid approximation
        (run_ancoding_grid) and the kernel time in ms."""
        out = np.zeros(nbins, dtype=np.uint64)
        ms = C.c_double(0)
        check(lib().cr_an_grid_ref(self._h, gdim, n, m1, m2, A, _u64p(out), nbins, C.byref(ms)))
        return [int(x) for x in out], ms.value


class BatchPlan:
    """A captured hipGraph replaying one batch (cr_batch_plan)."""

    def __init__(self, ctx: Context, codes, rows_per_lane: int = 0, items_per_wave: int = 0):
        self.codes = list(codes)
        self.stride = Context._stride(self.codes)
        self._h = C.c_void_p()
        t = cr_pair_tuning(rows_per_lane, 0, 0, items_per_wave)
        check(lib().cr_batch_plan_create(ctx.handle, Context._codes(self.codes), len(self.codes), self.stride,
                                         C.byref(t), C.byref(self._h)))
        self._out = np.zeros(len(self.codes) * self.stride, dtype=np.uint64)

    def run(self, fetch: bool = True):
        ms = C.c_double(0)
        check(lib().cr_batch_plan_run(self._h, _u64p(self._out) if fetch else None, C.byref(ms)))
        if not fetch:
            return None, ms.value
        o = self._out.reshape(len(self.codes), self.stride)
        return [[int(x) for x in row[: an_bins(A, N)]] for row, (A, N) in zip(o, self.codes)], ms.value

    def close(self):
        if self._h:
            lib().cr_batch_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request loop (cr_server) on one context."""

    def __init__(self, ctx: Context, hbm_budget_bytes: int = 0, max_active: int = 0, rows_per_lane: int = 0,
                 iteration_ms: float = 0.0):
        self._h = C.c_void_p()
        cfg = cr_server_config(hbm_budget_bytes, max_active, rows_per_lane, iteration_ms)
        check(lib().cr_server_create(ctx.handle, C.byref(cfg), C.byref(self._h)))

    def submit(self, kind: int, k: int, A: int = 0, stride: int = 0, samples: int = 0, seed: int = 0,
               stream: int = 0) -> int:
        j = cr_job(kind, k, A, stride, samples, seed, stream, 0)
        rid = C.c_uint64()
        check(lib().cr_server_submit(self._h, C.byref(j), C.byref(rid)))
        return rid.value

    def wait(self, rid: int, timeout_ms: float = -1.0) -> None:
        check(lib().cr_server_wait(self._h, rid, timeout_ms))

    def result(self, rid: int) -> tuple[list[int], dict]:
        res = cr_job_result()
        out = np.zeros(130, dtype=np.uint64)
        check(lib().cr_server_result(self._h, rid, _u64p(out), 130, C.byref(res)))
        info = {f: getattr(res, f) for f, _ in cr_job_result._fields_}
        return [int(x) for x in out[: res.nbins]], info

    def state(self, rid: int) -> dict:
        res = cr_job_result()
        lib().cr_server_result(self._h, rid, None, 0, C.byref(res))
        return {f: getattr(res, f) for f, _ in cr_job_result._fields_}

    def cancel(self, rid: int) -> None:
        check(lib().cr_server_cancel(self._h, rid))

    def stats(self) -> dict:
        st = cr_server_stats()
        check(lib().cr_server_stats_get(self._h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in cr_server_stats._fields_}

    def now(self) -> float:
        return lib().cr_server_now(self._h)

    def close(self):
        if self._h:
            lib().cr_server_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
