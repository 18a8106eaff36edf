"""GPU parity of the exhaustive AN pair kernel (cr_an_pairs) against the oracle,
the reference's golden vectors and its recorded k=20/24 result files."""
import math
import sys

import pytest

from conftest import golden, padded_eq
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import coding_reliability_amd as cra
    c = cra.Context(0)
    yield c
    c.close()


def report(k, raw, cc):
    return O.an_report(k, raw, cc)


_ORACLE = {}


def oracle_raw(A, N, nbins):
    key = (A, N)
    if key not in _ORACLE:
        _ORACLE[key] = O.an_pairs_raw(A, N, 65)
    return _ORACLE[key][:nbins]


def test_loaded_in_tree_library(ctx):
    from coding_reliability_amd import _lib
    assert _lib.LIB_PATH.endswith("coding_reliability_amd/libcodrel.so")
    info = ctx.device_info()
    print("device", info, flush=True)
    assert info["arch"].startswith("gfx950")


@pytest.mark.parametrize("A,k", [(61, 8), (641, 8), (61, 16), (641, 16)])
def test_golden_vectors(ctx, solutions, A, k):
    raw, ms = ctx.an_pairs(A, 1 << k)
    cc = k + O.ceil_log2(A) + 1
    assert padded_eq(report(k, raw, cc), solutions[f"solution_an{k}_A{A}"])


@pytest.mark.parametrize("N", [1, 2, 3, 63, 64, 65, 127, 128, 129, 255, 256, 257, 1000, 4097])
@pytest.mark.parametrize("A", [1, 3, 61, 58659, (1 << 31) + 11, (1 << 40) + 3, 0xFFFFFFFFFF1])
def test_small_sizes_all_modes(ctx, N, A):
    if (N - 1) * A >= 1 << 64:
        pytest.skip("codewords exceed 64 bits")
    raw, _ = ctx.an_pairs(A, N)
    ref = oracle_raw(A, N, len(raw))
    assert raw == ref, (A, N)


@pytest.mark.parametrize("R", [2, 4, 8, 16])
@pytest.mark.parametrize("Q", [1, 3, 7])
@pytest.mark.parametrize("A,N", [(58659, 5000), (61, 3000), ((1 << 33) + 5, 2500), (3, 70000)])
def test_tuning_grid(ctx, R, Q, A, N):
    # many items, partial chunks and row blocks, every kernel variant
    raw, _ = ctx.an_pairs(A, N, rows_per_lane=R, chunk_blocks=Q)
    assert raw == oracle_raw(A, N, len(raw))


@pytest.mark.parametrize("A,N", [(58659, 80000), (65533, 70000), (32771, 140000)])
def test_w2seg_crosses_2_32(ctx, A, N):
    # codewords cross 2^32 (segmented high-dword path) and many column low
    # dwords have bit 31 set
    assert ((N - 1) * A).bit_length() > 32
    raw, _ = ctx.an_pairs(A, N)
    assert raw == oracle_raw(A, N, len(raw))


def test_two_waves(ctx):
    # 2 persistent waves over a k=16 code: each wave walks hundreds of items of
    # different row blocks into one LDS histogram before its single flush
    A, k = 58659, 16
    raw, _ = ctx.an_pairs(A, 1 << k, rows_per_lane=16, chunk_blocks=2, waves=2)
    ref = oracle_raw(A, 1 << k, len(raw))
    assert raw == ref


def test_k16_58659_partial_pin(ctx):
    # results/superAs/superA_n16_h16.txt:33: A=58659, k=16 -> first nonzero bin 6 = 733232
    raw, _ = ctx.an_pairs(58659, 1 << 16)
    rep = report(16, raw, 16 + 16 + 1)
    assert all(rep[i] == 0 for i in range(1, 6)) and rep[6] == 733232
    assert raw == oracle_raw(58659, 1 << 16, len(raw))


def test_exact_k20_fixtures(ctx):
    runs = [r for r in golden("an_exhaustive_runs.json")["runs"] if r["n"] == 20]
    assert len(runs) == 4
    for r in runs:
        raw, ms = ctx.an_pairs(r["A"], 1 << 20)
        assert report(20, raw, len(r["counts"])) == r["counts"], r["src"]


def test_k24_A61_golden(ctx, solutions):
    raw, ms = ctx.an_pairs(61, 1 << 24)
    print(f"k=24 A=61: {ms:.1f} ms", flush=True)
    assert padded_eq(report(24, raw, 31), solutions["solution_an24_A61"])


def test_k24_40bit_fixture(ctx):
    # 40-bit codewords (the 58659 k=24 shape): results/As/ancoding_gpu_A65533_n24_h16.csv
    run = [r for r in golden("an_exhaustive_runs.json")["runs"] if r["n"] == 24 and r["A"] == 65533][0]
    raw, ms = ctx.an_pairs(65533, 1 << 24)
    print(f"k=24 A=65533: {ms:.1f} ms", flush=True)
    assert report(24, raw, len(run["counts"])) == run["counts"]


def test_k22_invariants(ctx):
    # size-independent properties at a size the oracle cannot run here:
    # sum = C(2^k,2), bin 0 empty, bins above the codeword width empty,
    # and A*2^s has the same distances as A
    A, k = 58659, 22
    raw, _ = ctx.an_pairs(A, 1 << k)
    N = 1 << k
    assert sum(raw) == N * (N - 1) // 2
    assert raw[0] == 0
    assert all(x == 0 for x in raw[((N - 1) * A).bit_length() + 1:])
    raw2, _ = ctx.an_pairs(A << 3, 1 << k)
    assert raw2[: len(raw)] == raw


def test_invalid_arguments(ctx):
    import coding_reliability_amd as cra
    with pytest.raises(cra.CRError):
        ctx.an_pairs(0, 16)
    with pytest.raises(cra.CRError):
        ctx.an_pairs(61, 1 << 16, nbins=5)
    with pytest.raises(cra.CRError):
        ctx.an_pairs((1 << 63) + 1, 4)


def test_all_exact_fixtures_k_le_20_in_one_batch(ctx):
    """Every exact exhaustive result file of the reference with k <= 20 (A odd
    1..127 at k=8..16, A=35..63 at k=18, four at k=20, A=61 runtimes) in one
    batched launch, bit-exact in the reported form."""
    runs = [r for r in golden("an_exhaustive_runs.json")["runs"] if r["n"] <= 20]
    assert len(runs) >= 600
    codes = [((r["A"] if r["A"] & (r["A"] - 1) else 1), 1 << r["n"]) for r in runs]
    got, ms = ctx.an_batch(codes)
    print(f"{len(codes)} recorded codes in one launch: {ms:.1f} ms", flush=True)
    for r, raw in zip(runs, got):
        assert report(r["n"], raw + [0] * 64, len(r["counts"])) == r["counts"], r["src"]


def test_more_k24_fixtures_batched(ctx):
    # results/As/ancoding_gpu_A17_n24_h5.csv (30-bit words) and A55831 (40-bit)
    runs = [r for r in golden("an_exhaustive_runs.json")["runs"] if r["n"] == 24 and r["A"] in (17, 55831)]
    assert len(runs) == 2
    got, ms = ctx.an_batch([(r["A"], 1 << 24) for r in runs])
    print(f"two k=24 codes: {ms:.1f} ms", flush=True)
    for r, raw in zip(runs, got):
        assert report(24, raw + [0] * 64, len(r["counts"])) == r["counts"], r["src"]


@pytest.mark.parametrize("A,N", [((1 << 64) - 1, 2), ((1 << 63) + 1, 2), ((1 << 62) + 7, 4), ((1 << 32) - 1, 1 << 16),
                                 (0x5555555555555555, 3)])
def test_extreme_widths(ctx, A, N):
    # codewords up to 64 bits (65 bins), a single pair, the widest words
    raw, _ = ctx.an_pairs(A, N, nbins=65)
    assert raw == oracle_raw(A, N, 65), (A, N)
    assert sum(raw) == N * (N - 1) // 2


def test_empty_and_single_word(ctx):
    import coding_reliability_amd as cra
    raw, _ = ctx.an_pairs(61, 1)
    assert sum(raw) == 0
    with pytest.raises(cra.CRError):
        ctx.an_pairs(61, 0)


def test_two_contexts_from_two_threads(ctx):
    # contexts share no state: two threads, one context each, run at once
    import threading

    import coding_reliability_amd as cra
    jobs = [(58659, 1 << 16), (61, 1 << 17)]
    want = [ctx.an_pairs(A, N)[0] for A, N in jobs]
    got = [None, None]

    def run(i):
        c = cra.Context(0)
        try:
            got[i] = [c.an_pairs(A, N)[0] for A, N in jobs[i:] + jobs[:i]][0]
        finally:
            c.close()

    ts = [threading.Thread(target=run, args=(i,)) for i in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=120)
    assert got == want
