"""CPU oracle for the distance-distribution path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module.  The product library never calls it.

It wraps ``liboracle.so`` (built from ``cr_oracle.c`` by ``oracle/Makefile``),
a plain-C/OpenMP restatement of the reference CPU path, plus a few pure-Python
checks (MacWilliams, signed-digit weights) that pin it.

Pinning of each oracle function (what checks it, see ``tests/test_oracle.py``):

* ``an_cpu_counts`` (an_coding.cpp:50-102): golden vectors of
  distance_distribution/src/solutions.h:26-114, every exact fixture under
  results/As (k<=18), results/runtimes/*, the 5x5 matrices of
  legacy/print_codeword_matrix.out, brute force, and the A=1 closed form.
* ``an_pairs_raw``: equals ``an_cpu_counts`` (independent loop nest) and the
  same fixtures; brute force on tiny inputs.
* ``grid1d_ref_samples`` / ``grid2d_ref_samples`` + ``rect_hist``: the
  deterministic fixtures results/converge/geom/*grid1d*, results/GridMs/*grid1d*
  and results/runtimes/*/ancoding_grid1d_*; m = 2^k reduces to exhaustive.
* ``hamming_*``: solutions.h:18-24 weight distributions, brute-force sphere
  enumeration (hamming_native_short.cpp:52-77), MacWilliams identity of the
  dual code (all k, including the (72,64) code the reference cannot run).
* ``philox4x32_10``: known-answer vectors produced from the open-source
  rocRAND header (tests/golden/make_philox_vectors.py).
* ``mc_word`` / ``hamming_mc`` / ``an_mc`` / ``an_mc_sample``: parity
  unpinned against the reference (it uses closed cuRAND Sobol32 streams);
  pinned statistically (4-sigma binomial bounds vs exhaustive / MacWilliams
  truth, the binomial closed form of A = 1), and the sample words against
  the Philox known-answer vectors through their counter layout.
* ``an_candidates``: independent NAF-weight recurrence in pure Python
  (``naf_weight``) brute-forced for small (k, |A|).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

u64p = C.POINTER(C.c_uint64)


def build() -> str:
    """Compile liboracle.so in place (gcc, no GPU)."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.or_an_cpu_counts.restype = C.c_uint64
        L.or_an_cpu_counts.argtypes = [C.c_uint64, C.c_uint64, u64p, C.c_uint64, C.c_uint64, C.c_uint64]
        L.or_an_pairs_raw.argtypes = [C.c_uint64, C.c_uint64, u64p, C.c_uint]
        L.or_grid1d_ref_samples.restype = C.c_uint64
        L.or_grid1d_ref_samples.argtypes = [C.c_uint64, C.c_uint64, C.c_double, C.c_uint, C.c_int, u64p, C.c_uint64]
        L.or_grid2d_ref_samples.restype = C.c_uint64
        L.or_grid2d_ref_samples.argtypes = [C.c_uint64, C.c_uint64, C.c_double, C.c_int, u64p, C.c_uint64]
        L.or_rect_hist.argtypes = [u64p, C.c_uint64, u64p, C.c_uint64, u64p, C.c_uint]
        L.or_rect_hist_arith.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_int, u64p, C.c_uint64, u64p, C.c_uint]
        L.or_grid_report.argtypes = [C.c_int, C.c_uint64, C.c_uint64, C.c_uint64, u64p, C.c_uint, u64p]
        L.or_hamming_parity_bits.restype = C.c_uint
        L.or_hamming_parity_bits.argtypes = [C.c_uint]
        L.or_hamming_weight_nonext.restype = C.c_uint
        L.or_hamming_weight_nonext.argtypes = [C.c_uint64, C.c_uint]
        L.or_hamming_codeword_ext.restype = C.c_uint64
        L.or_hamming_codeword_ext.argtypes = [C.c_uint64, C.c_uint]
        L.or_hamming_weights.argtypes = [C.c_uint, C.c_uint64, C.c_uint64, u64p, C.c_uint]
        L.or_hamming_weights_list.argtypes = [C.c_uint, u64p, C.c_uint64, u64p, C.c_uint]
        L.or_hamming_report.argtypes = [C.c_uint, u64p, C.c_int, u64p]
        L.or_hamming_bruteforce.argtypes = [C.c_uint, C.c_int, u64p]
        L.or_hamming_grid_ref_samples.restype = C.c_uint64
        L.or_hamming_grid_ref_samples.argtypes = [C.c_uint64, C.c_double, C.c_uint64, u64p, C.c_uint64]
        L.or_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.or_mc_word.restype = C.c_uint64
        L.or_mc_word.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint]
        L.or_hamming_mc.argtypes = [C.c_uint, C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint64, u64p, C.c_uint]
        L.or_an_mc.argtypes = [C.c_uint64, C.c_uint, C.c_uint64, C.c_uint64, u64p, C.c_uint]
        L.or_an_mc_rows.argtypes = [C.c_uint64, C.c_uint, C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint64, u64p,
                                    C.c_uint]
        L.or_an_mc_sample.restype = C.c_uint64
        L.or_an_mc_sample.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint]
        L.or_an_min_naf.restype = C.c_uint64
        L.or_an_min_naf.argtypes = [C.c_uint64, C.c_uint]
        L.or_an_candidates.restype = C.c_uint64
        L.or_an_candidates.argtypes = [C.c_uint, C.c_uint, u64p, u64p, C.c_uint64]
        L.or_num_threads.restype = C.c_int
        _LIB = L
    return _LIB


def _p(a: np.ndarray):
    return a.ctypes.data_as(u64p)


def _u128_list(buf: np.ndarray) -> list[int]:
    b = buf.reshape(-1, 2)
    return [int(lo) | (int(hi) << 64) for lo, hi in b]


def ceil_log2(A: int) -> int:
    """h = ceil(log(A)/log(2.0)) as computed in double (an_coding.cu:185)."""
    import math
    return int(math.ceil(math.log(A) / math.log(2.0))) if A > 1 else 0


# --- exhaustive AN ---------------------------------------------------------

def an_cpu_counts(n: int, A: int, shard_begin: int = 0, shard_end: int = 1 << 62,
                  count_counts: int | None = None) -> tuple[list[int], int]:
    """Reported-form counts of run_ancoding_cpu (an_coding.cpp:104-142) and the
    number of unordered pairs evaluated."""
    if count_counts is None:
        count_counts = n + ceil_log2(A) + 1
    buf = np.zeros(2 * max(count_counts, 1), dtype=np.uint64)
    pairs = lib().or_an_cpu_counts(n, A, _p(buf), count_counts, shard_begin, shard_end)
    return _u128_list(buf), int(pairs)


def an_pairs_raw(Ap: int, N: int, nbins: int = 65) -> list[int]:
    raw = np.zeros(nbins, dtype=np.uint64)
    lib().or_an_pairs_raw(Ap & ((1 << 64) - 1), N, _p(raw), nbins)
    return [int(x) for x in raw]


def an_report(n: int, raw: list[int], count_counts: int) -> list[int]:
    """an_coding.cu:287-292."""
    out = [0] * count_counts
    out[0] = 1 << n
    for i in range(1, count_counts):
        out[i] = raw[i] << 1
    return out


def an_minb(counts: list[int], limit: int) -> tuple[int, int]:
    """an_coding.cu:295-308."""
    for i in range(1, limit):
        if counts[i] != 0:
            return i, counts[i]
    return 0xFFFF, (1 << 128) - 1


# --- reference grid ----------------------------------------------------------

def grid_ref_config(A: int, n: int, m: int, m2: int, gdim: int) -> dict:
    """Host set-up of run_ancoding_grid (an_coding_grid.cu:215-296, one device)."""
    if (A & (A - 1)) == 0:
        A = 1
    msgs = 1 << n
    m = min(m, msgs)
    m2 = msgs if gdim == 1 else min(m2, msgs)
    cc = n + ceil_log2(A) + 1
    return dict(A=A, n=n, m=m, m2=m2, cc=cc, step=1.0 * msgs / m, step2=1.0 * msgs / m2)


def grid1d_ref_samples(A: int, n: int, step: float, unroll: int, t32: bool) -> np.ndarray:
    cap = 1 << 22
    out = np.zeros(cap, dtype=np.uint64)
    cnt = lib().or_grid1d_ref_samples(A, n, step, unroll, int(t32), _p(out), cap)
    assert cnt <= cap
    return out[:cnt].copy()


def grid2d_ref_samples(A: int, n: int, step: float, t32: bool) -> np.ndarray:
    cap = 1 << 22
    out = np.zeros(cap, dtype=np.uint64)
    cnt = lib().or_grid2d_ref_samples(A, n, step, int(t32), _p(out), cap)
    assert cnt <= cap
    return out[:cnt].copy()


def rect_hist(rows: np.ndarray, cols: np.ndarray, nbins: int = 65) -> list[int]:
    rows = np.ascontiguousarray(rows, dtype=np.uint64)
    cols = np.ascontiguousarray(cols, dtype=np.uint64)
    h = np.zeros(nbins, dtype=np.uint64)
    lib().or_rect_hist(_p(rows), len(rows), _p(cols), len(cols), _p(h), nbins)
    return [int(x) for x in h]


def rect_hist_arith(Ap: int, r0: int, r1: int, t32: bool, cols: np.ndarray, nbins: int = 65) -> list[int]:
    cols = np.ascontiguousarray(cols, dtype=np.uint64)
    h = np.zeros(nbins, dtype=np.uint64)
    lib().or_rect_hist_arith(Ap, r0, r1, int(t32), _p(cols), len(cols), _p(h), nbins)
    return [int(x) for x in h]


def grid_report(gdim: int, n: int, m: int, m2: int, hist: list[int], cc: int) -> list[int]:
    h = np.array(hist[:cc] + [0] * max(0, cc - len(hist)), dtype=np.uint64)
    out = np.zeros(2 * cc, dtype=np.uint64)
    lib().or_grid_report(gdim, n, m, m2, _p(h), cc, _p(out))
    return _u128_list(out)


def an_grid_ref(gdim: int, n: int, m: int, m2: int, A: int) -> tuple[list[int], list[int], dict]:
    """Kernel-level histogram and reported counts of run_ancoding_grid."""
    cfg = grid_ref_config(A, n, m, m2, gdim)
    A, m, m2, cc = cfg["A"], cfg["m"], cfg["m2"], cfg["cc"]
    if gdim == 1:
        t32 = cc <= 32
        unroll = 3 if m <= 16 else 16
        cols = grid1d_ref_samples(A, n, cfg["step"], unroll, t32)
        hist = rect_hist_arith(A, 0, 1 << n, t32, cols)
    else:
        t32 = (A << n) < (1 << 32)
        cols = grid2d_ref_samples(A, n, cfg["step"], t32)
        mask = 0xFFFFFFFF if t32 else (1 << 64) - 1
        rows = np.array([(A * int(float(i) * cfg["step2"])) & mask for i in range(m2)], dtype=np.uint64)
        hist = rect_hist(rows, cols)
        hist[0] = 0  # dancoding_grid_2d flushes bins 1.. only (an_coding_grid.cu:138)
    return hist, grid_report(gdim, n, m, m2, hist, cc), cfg


# --- Hamming -----------------------------------------------------------------

def hamming_parity_bits(k: int) -> int:
    return int(lib().or_hamming_parity_bits(k))


def hamming_N(k: int) -> int:
    """Codeword length k + parity bits + overall parity."""
    return k + hamming_parity_bits(k) + 1


def hamming_weight_nonext(d: int, k: int) -> int:
    return int(lib().or_hamming_weight_nonext(d, k))


def hamming_codeword_ext(d: int, k: int) -> int:
    return int(lib().or_hamming_codeword_ext(d, k))


def hamming_weights(k: int, d0: int = 0, d1: int | None = None, nbins: int = 130) -> list[int]:
    if d1 is None:
        d1 = 1 << k
    h = np.zeros(nbins, dtype=np.uint64)
    lib().or_hamming_weights(k, d0, d1, _p(h), nbins)
    return [int(x) for x in h]


def hamming_weights_list(k: int, words: np.ndarray, nbins: int = 130) -> list[int]:
    words = np.ascontiguousarray(words, dtype=np.uint64)
    h = np.zeros(nbins, dtype=np.uint64)
    lib().or_hamming_weights_list(k, _p(words), len(words), _p(h), nbins)
    return [int(x) for x in h]


def hamming_report(k: int, h: list[int], with_1bit: bool) -> list[int]:
    N = hamming_N(k)
    hh = np.array((list(h) + [0] * 130)[:130], dtype=np.uint64)
    out = np.zeros(2 * (N + 1), dtype=np.uint64)
    lib().or_hamming_report(k, _p(hh), int(with_1bit), _p(out))
    return _u128_list(out)


def hamming_bruteforce(k: int, with_1bit: bool) -> list[int]:
    N = hamming_N(k)
    out = np.zeros(2 * (N + 1), dtype=np.uint64)
    lib().or_hamming_bruteforce(k, int(with_1bit), _p(out))
    return _u128_list(out)


def hamming_grid_ref_samples(m: int, step: float, S: int) -> np.ndarray:
    cap = max(1, (m // S) * S)
    out = np.zeros(cap, dtype=np.uint64)
    cnt = lib().or_hamming_grid_ref_samples(m, step, S, _p(out), cap)
    return out[:cnt].copy()


# --- Philox / Monte Carlo ---------------------------------------------------------

def philox4x32_10(ctr, key) -> list[int]:
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().or_philox4x32_10(c, k, o)
    return list(o)


def mc_word(seed: int, stream: int, s: int, k: int) -> int:
    return int(lib().or_mc_word(seed, stream, s, k))


def hamming_mc(k: int, seed: int, stream: int, s0: int, s1: int, nbins: int = 130) -> list[int]:
    h = np.zeros(nbins, dtype=np.uint64)
    lib().or_hamming_mc(k, seed, stream, s0, s1, _p(h), nbins)
    return [int(x) for x in h]


def an_mc(A: int, n: int, m: int, seed: int, nbins: int = 65) -> list[int]:
    h = np.zeros(nbins, dtype=np.uint64)
    lib().or_an_mc(A, n, m, seed, _p(h), nbins)
    return [int(x) for x in h]


def an_mc_rows(A: int, n: int, m: int, seed: int, i0: int, i1: int, nbins: int = 65) -> list[int]:
    """The AN Monte Carlo histogram restricted to the rows i in [i0, i1)."""
    h = np.zeros(nbins, dtype=np.uint64)
    lib().or_an_mc_rows(A, n, m, seed, i0, i1, _p(h), nbins)
    return [int(x) for x in h]


def an_mc_sample(seed: int, i: int, j: int, n: int) -> int:
    return int(lib().or_an_mc_sample(seed, i, j, n))


# --- ANcandidates -----------------------------------------------------------------

def an_min_naf(A: int, k: int) -> int:
    return int(lib().or_an_min_naf(A, k))


def an_candidates(k: int, abits: int) -> tuple[int, list[int]]:
    cap = 1 << 20
    best = np.zeros(1, dtype=np.uint64)
    cands = np.zeros(cap, dtype=np.uint64)
    cnt = lib().or_an_candidates(k, abits, _p(best), _p(cands), cap)
    return int(best[0]), [int(x) for x in cands[: min(cnt, cap)]]


def naf_weight(x: int) -> int:
    """Number of nonzero digits of the non-adjacent form, by the textbook
    right-to-left recurrence (independent of the popcount(x ^ 3x) identity)."""
    w = 0
    while x:
        if x & 1:
            d = 2 - (x & 3)  # +1 or -1
            x -= d
            w += 1
        x >>= 1
    return w


def num_threads() -> int:
    return int(lib().or_num_threads())
