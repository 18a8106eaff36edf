"""GPU parity of the Extended Hamming weight kernel (all sources) and of the AN
Monte Carlo kernel: bit-exact against the oracle on shared inputs / Philox
counters, exact against the golden vectors and the MacWilliams closed form,
and within 4-sigma binomial bounds at full Monte Carlo sizes."""
import math

import numpy as np
import pytest

import coding_reliability_amd as cra
from conftest import macwilliams_weights
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = cra.Context(0)
    yield c
    c.close()


def merged_even(h):
    """Extended-codeword weights from non-extended weights (odd -> next even)."""
    n = len(h) + 1
    return [(h[w] + (h[w - 1] if w else 0)) if w % 2 == 0 and w < len(h) else 0 for w in range(n)]


@pytest.mark.parametrize("k", [4, 8, 12, 16, 20, 24, 25, 26, 27, 28])
def test_exhaustive_vs_oracle(ctx, k):
    h, _ = ctx.hamming_weights(k)
    assert h == O.hamming_weights(k)[: len(h)]


@pytest.mark.parametrize("k,first,count", [(32, 0, 3 << 24), (32, 1 << 30, 5 << 22), (32, (1 << 32) - (1 << 24), 1 << 25),
                                           (30, 12345 << 10, 1 << 26), (28, 1 << 20, (1 << 28) - (1 << 20))])
def test_exhaustive_windows_vs_oracle(ctx, k, first, count):
    # aligned and unaligned windows, wrap-around past 2^k, partial last items:
    # the exhaustive fast path (aligned full items) and the generic path meet
    h, _ = ctx.hamming_weights(k, cra.HAM_RANGE, first=first, count=count)
    mask = (1 << k) - 1
    lo, hi = first & mask, (first + count) & mask
    if first + count <= (first & ~mask) + (1 << k):
        ref = O.hamming_weights(k, lo, lo + count)
    else:  # wraps: [lo, 2^k) and [0, hi)
        a, b = O.hamming_weights(k, lo, 1 << k), O.hamming_weights(k, 0, hi)
        ref = [x + y for x, y in zip(a, b)]
    assert h == ref[: len(h)]


@pytest.mark.parametrize("k,name", [(8, "sol_hamming_n8"), (16, "sol_hamming_n16"), (24, "sol_hamming_n24"),
                                    (32, "sol_hamming_n32")])
def test_exhaustive_golden(ctx, solutions, k, name):
    # run_hamming's reported even bins equal solutions.h << k (hamming.cu:183-208)
    h, ms = ctx.hamming_weights(k)
    assert sum(h) == 1 << k
    rep = cra.hamming_report(k, h, True)
    sol = solutions[name]
    assert all(rep[i] == sol[i] << k for i in range(4, len(sol), 2))


def test_exhaustive_k32_macwilliams(ctx):
    h, _ = ctx.hamming_weights(32)
    E = merged_even(h)
    assert E[: 40] == macwilliams_weights(32)[: 40]


def test_range_subrange_and_k64_window(ctx):
    # a window of the (72,64) code: d in [2^63 - 2^20, 2^63 + 2^20)
    k, first, count = 64, (1 << 63) - (1 << 20), 1 << 21
    h, _ = ctx.hamming_weights(k, cra.HAM_RANGE, first=first, count=count)
    words = np.arange(count, dtype=np.uint64) + np.uint64(first)
    ref = O.hamming_weights_list(k, words)
    assert h == ref[: len(h)]


def test_stride_grid_c3(ctx):
    # (39,32) integer grid d_i = i*65537 mod 2^32, i < 2^20
    k, stride, m = 32, 65537, 1 << 20
    h, _ = ctx.hamming_weights(k, cra.HAM_STRIDE, first=0, count=m, stride=stride)
    words = (np.arange(m, dtype=np.uint64) * np.uint64(stride)) & np.uint64(0xFFFFFFFF)
    assert h == O.hamming_weights_list(k, words)[: len(h)]
    assert sum(h) == m


def test_list_reference_grid_samples(ctx):
    # the sample points of the reference grid kernel (hamming_grid.cu:53-74), k=32, m=2^20
    k, m, S = 32, 1 << 20, 256
    step = 1.0 * (1 << k) / m
    words = O.hamming_grid_ref_samples(m, step, S)
    h, _ = ctx.hamming_weights(k, cra.HAM_LIST, words=words)
    assert h == O.hamming_weights_list(k, words)[: len(h)]


@pytest.mark.parametrize("k,first,count", [(32, 0, 1 << 16), (16, 4096, 100003), (64, 0, 1 << 15), (64, 2, 33333),
                                           (57, 10, 4096)])
def test_mc_bitexact_vs_oracle(ctx, k, first, count):
    h, _ = ctx.hamming_weights(k, cra.HAM_MC, first=first, count=count, seed=0xC0FFEE, stream=7)
    assert h == O.hamming_mc(k, 0xC0FFEE, 7, first, first + count)[: len(h)]


def test_mc_c3_2pow34_within_4sigma(ctx):
    # Monte Carlo over 2^34 Philox words of the (39,32) code vs the exhaustive truth
    k, S = 32, 1 << 34
    h, ms = ctx.hamming_weights(k, cra.HAM_MC, first=0, count=S, seed=20261202, stream=0)
    truth, _ = ctx.hamming_weights(k)
    print(f"MC 2^34 samples (39,32): {ms:.1f} ms", flush=True)
    assert sum(h) == S
    for w, t in enumerate(truth):
        p = t / (1 << k)
        assert abs(h[w] - S * p) <= 4 * math.sqrt(S * p * (1 - p)) + 1, w


def test_mc_72_64_within_4sigma_macwilliams(ctx):
    # (72,64): no exhaustive run is possible; truth from the MacWilliams identity
    k, S = 64, 1 << 30
    h, ms = ctx.hamming_weights(k, cra.HAM_MC, first=0, count=S, seed=99, stream=3)
    A = macwilliams_weights(64)
    E = merged_even(h)
    print(f"MC 2^30 samples (72,64): {ms:.1f} ms", flush=True)
    for w in range(0, 73, 2):
        p = A[w] / 2.0 ** 64
        assert abs(E[w] - S * p) <= 4 * math.sqrt(S * p * (1 - p)) + 1, w


def test_batch_equals_separate_calls(ctx):
    # one call: the (39,32) stride grid of 2^20 words beside its 2^34-word
    # Monte Carlo run, an exhaustive (22,16) code, a list, an empty request
    # and two 64-bit requests (the second launch); every row equals the
    # request's own cr_hamming_weights call bit for bit
    words = (np.arange(5000, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)) >> np.uint64(40)
    reqs = [dict(k=32, source=cra.HAM_STRIDE, count=1 << 20, stride=65537),
            dict(k=32, source=cra.HAM_MC, count=1 << 34, seed=20261202),
            dict(k=16),
            dict(k=64, source=cra.HAM_MC, first=2, count=33333, seed=5, stream=1),
            dict(k=24, source=cra.HAM_LIST, words=words),
            dict(k=8, count=0),
            dict(k=57, source=cra.HAM_RANGE, first=(1 << 56) - 999, count=1 << 20)]
    rows, t_batch = ctx.hamming_weights_batch(reqs)
    t_sep = 0.0
    for q, row in zip(reqs, rows):
        q = dict(q)
        h, t = ctx.hamming_weights(q.pop("k"), **q)
        t_sep += t
        assert row == h, q
    assert rows[5] == [0] * len(rows[5])
    print(f"batch of {len(reqs)}: {t_batch:.2f} ms, separate calls {t_sep:.2f} ms", flush=True)


def test_batch_c3_grid_in_mc_launch(ctx):
    # the grid alone is launch-bound; in the Monte Carlo launch it costs its
    # words only (checked against the oracle, not against another call)
    grid = dict(k=32, source=cra.HAM_STRIDE, count=1 << 20, stride=65537)
    rows, ms = ctx.hamming_weights_batch([grid, dict(k=32, source=cra.HAM_MC, count=1 << 30, seed=3)])
    words = (np.arange(1 << 20, dtype=np.uint64) * np.uint64(65537)) & np.uint64(0xFFFFFFFF)
    assert rows[0] == O.hamming_weights_list(32, words)[: len(rows[0])]
    assert sum(rows[1]) == 1 << 30


def test_batch_many_small_requests(ctx):
    # 300 seeded requests of every source, k from 1 to 64, counts from 0 to
    # 40000 (partial items, empty jobs), so both launches binary-search a few
    # hundred jobs; each row equals the request's own call, and the lists and
    # a sample of ranges also equal the oracle
    rng = np.random.default_rng(20261203)
    reqs = []
    for _ in range(300):
        k = int(rng.integers(1, 65))
        src = int(rng.integers(0, 4))
        count = int(rng.integers(0, 40001)) if rng.random() > 0.05 else 0
        q = dict(k=k, source=src, count=count)
        if src == cra.HAM_STRIDE:
            q.update(first=int(rng.integers(0, 1 << 62)), stride=int(rng.integers(1, 1 << 62)))
        elif src == cra.HAM_RANGE:
            q.update(first=int(rng.integers(0, 1 << min(k, 62))))
        elif src == cra.HAM_MC:
            per = 4 if k <= 32 else 2
            q.update(first=per * int(rng.integers(0, 1 << 30)), seed=int(rng.integers(0, 1 << 63)),
                     stream=int(rng.integers(0, 1 << 32)))
        else:
            q = dict(k=k, source=src, words=rng.integers(0, 1 << 63, size=count, dtype=np.uint64))
        reqs.append(q)
    rows, ms = ctx.hamming_weights_batch(reqs)
    checked = 0
    for q, row in zip(reqs, rows):
        q = dict(q)
        k = q.pop("k")
        h, _ = ctx.hamming_weights(k, **q)
        assert row == h, (k, q.get("source"), q.get("count"))
        if q.get("source") == cra.HAM_LIST and checked < 20:
            assert row == O.hamming_weights_list(k, q["words"])[: len(row)]
            checked += 1
    assert checked == 20
    with pytest.raises(cra.CRError):
        ctx.hamming_weights_batch([dict(k=8)] * 4097)  # beyond 4096 requests
    print(f"batch of 300: {ms:.3f} ms", flush=True)


def test_batch_refusals(ctx):
    with pytest.raises(cra.CRError):
        ctx.hamming_weights_batch([])
    with pytest.raises(cra.CRError):
        ctx.hamming_weights_batch([dict(k=16), dict(k=65)])
    with pytest.raises(cra.CRError):
        ctx.hamming_weights_batch([dict(k=32, source=cra.HAM_MC, first=2, count=8)])
    rows, ms = ctx.hamming_weights_batch([dict(k=16, count=0), dict(k=64, count=0)])  # nothing launched
    assert ms == 0 and not any(rows[0]) and not any(rows[1])


@pytest.mark.parametrize("A,n,m", [(61, 8, 16), (641, 10, 100), (58659, 12, 64), (3, 16, 5),
                                   (61, 12, 1024), (58659, 16, 512), (641, 14, 700), (65533, 9, 1280)])
def test_an_mc_bitexact_vs_oracle(ctx, A, n, m):
    h, _ = ctx.an_mc(A, n, m, seed=1234)
    assert h == O.an_mc(A, n, m, 1234)[: len(h)]


@pytest.mark.parametrize("A", [61, 3])
def test_an_mc_64bit_samples_bitexact_vs_oracle(ctx, A):
    # n = 33: 64-bit samples (two Philox words each, the row's high bits in
    # the counter's fourth word) and 64-bit codewords, every one of the 2^33
    # rows against the oracle's restatement of the same sample definition
    n, m = 33, 2
    h, ms = ctx.an_mc(A, n, m, seed=4321)
    print(f"AN MC A={A} n={n} m={m}: {ms:.1f} ms", flush=True)
    assert sum(h) == (1 << n) * m
    assert h == O.an_mc(A, n, m, 4321)[: len(h)]


@pytest.mark.parametrize("A,n,m", [(1, 40, 4), (1024, 40, 2), (1, 36, 64)])
def test_an_mc_64bit_binomial_closed_form(ctx, A, n, m):
    # for A a power of two the distance of A*i and A*u is popcount(i ^ u), and
    # u uniform on n bits makes it Binomial(n, 1/2) for every row: each bin is
    # within 4 sigma of 2^n m C(n, d) / 2^n
    h, ms = ctx.an_mc(A, n, m, seed=2026)
    print(f"AN MC A={A} n={n} m={m}: {ms:.1f} ms ({(1 << n) * m / ms / 1e9:.2f} Tpairs/s)", flush=True)
    tot = (1 << n) * m
    assert sum(h) == tot
    for d in range(n + 1):
        p = math.comb(n, d) / 2.0 ** n
        assert abs(h[d] - tot * p) <= 4 * math.sqrt(tot * p * (1 - p)) + 1, d
    assert not any(h[n + 1:])


@pytest.mark.parametrize("A,n,bounds", [(61, 10, [0, 7, 100, 333, 1024]), (58659, 12, [5, 5, 64, 257, 700]),
                                        (3, 9, [0, 1, 2, 3, 4, 255, 256, 257])])
def test_an_mc_ranges_bitexact_vs_oracle(ctx, A, n, bounds):
    # each range [b_i, b_{i+1}) is the difference of two oracle prefixes;
    # bounds off the 4-word Philox blocks and the 256-sample chunks, and an
    # empty range, exercise the partial chunks
    rows, _ = ctx.an_mc_ranges(A, n, bounds, seed=99)
    pre = [O.an_mc(A, n, b, 99) for b in bounds]
    for i, r in enumerate(rows):
        want = [x - y for x, y in zip(pre[i + 1], pre[i])]
        assert r == want[: len(r)], (i, bounds[i], bounds[i + 1])


@pytest.mark.parametrize("A,n,ms", [(61, 16, [1000, 3000, 9000, 27000]), (58659, 20, [7, 21, 63, 189, 567]),
                                    (61, 33, [1, 3, 9])])
def test_an_mc_prefixes_equal_separate_calls(ctx, A, n, ms):
    # the search's nested prefixes from one launch (every sample drawn once)
    # equal separate an_mc calls with each sample count, bit for bit; n = 33
    # draws 64-bit samples, two per Philox block
    pre, t_one = ctx.an_mc_prefixes(A, n, ms, seed=16384)
    t_sep = 0.0
    for m, p in zip(ms, pre):
        h, t = ctx.an_mc(A, n, m, seed=16384)
        t_sep += t
        assert p == h, m
        assert sum(p) == (1 << n) * m
    print(f"A={A} n={n} prefixes {ms}: one launch {t_one:.2f} ms, separate calls {t_sep:.2f} ms", flush=True)


def test_an_mc_ranges_refusals(ctx):
    with pytest.raises(cra.CRError):
        ctx.an_mc_ranges(61, 10, [10, 5], seed=1)  # decreasing
    with pytest.raises(cra.CRError):
        ctx.an_mc_ranges(61, 10, [0], seed=1)  # no range
    with pytest.raises(cra.CRError):
        ctx.an_mc_ranges(61, 10, list(range(66)), seed=1)  # 65 ranges
    rows, ms = ctx.an_mc_ranges(61, 10, [4, 4, 4], seed=1)  # all empty: nothing launched
    assert rows == [[0] * 65, [0] * 65] and ms == 0


def test_an_mc_4sigma_vs_exhaustive(ctx):
    A, n, m = 58659, 16, 4096
    h, _ = ctx.an_mc(A, n, m, seed=77)
    raw, _ = ctx.an_pairs(A, 1 << n)
    tot = (1 << n) * m
    assert sum(h) == tot
    for b in range(len(raw)):
        p = (2 * raw[b] + ((1 << n) if b == 0 else 0)) / 4 ** n
        assert abs(h[b] - tot * p) <= 4 * math.sqrt(tot * p * (1 - p)) + 1, b


def test_invalid(ctx):
    with pytest.raises(cra.CRError):
        ctx.hamming_weights(65)
    with pytest.raises(cra.CRError):
        ctx.hamming_weights(32, cra.HAM_MC, first=1, count=10)
    with pytest.raises(cra.CRError):
        ctx.an_mc(61, 60, 10, 1)  # codewords beyond 64 bits
    with pytest.raises(cra.CRError):
        ctx.an_mc(1, 64, 10, 1)
