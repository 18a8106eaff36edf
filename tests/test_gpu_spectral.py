"""GPU parity of the spectral exhaustive AN histogram (cr_an_pairs_spectral)
against the reference's recorded k=16 and k=24 result files, its golden
vectors, and the pair kernel bit for bit."""
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


def test_k16_recorded_runs(ctx):
    runs = [r for r in golden("an_exhaustive_runs.json")["runs"] if r["n"] == 16]
    assert len(runs) == 79
    done = {}
    for r in runs:
        if r["A"] not in done:
            done[r["A"]], _ = ctx.an_pairs_spectral(r["A"], 16)
        assert report(16, done[r["A"]], len(r["counts"])) == r["counts"], r["src"]


@pytest.mark.parametrize("k,nruns", [(14, 64), (15, 64), (18, 15), (20, 4)])
def test_recorded_runs_other_k(ctx, k, nruns):
    runs = [r for r in golden("an_exhaustive_runs.json")["runs"] if r["n"] == k]
    assert len(runs) == nruns
    done = {}
    for r in runs:
        if r["A"] not in done:
            done[r["A"]], _ = ctx.an_pairs_spectral(r["A"], k)
        assert report(k, done[r["A"]], len(r["counts"])) == r["counts"], r["src"]


@pytest.mark.parametrize("A,k", [(61, 17), (641, 19), (61, 21), (3, 23), (61, 23), (3, 17), (58659, 18),
                                 (1021, 20), (3, 21)])
def test_int16_pairs_k17_to_23_match_pair_kernel(ctx, A, k):
    # k = 17 .. 21: the one-wave first pass over bits 0 .. k-11 (its top 22 - k
    # index bits of each 4096-element block left as rows), the 10-bit last
    # pass; A = 3 at k = 17 has the fewest transforms that path takes (four)
    raw_s, _ = ctx.an_pairs_spectral(A, k)
    raw_p, _ = ctx.an_pairs(A, 1 << k)
    assert raw_s == raw_p


@pytest.mark.parametrize("A,k", [(58659, 22), (3, 22), (58659, 23), (1021, 23)])
def test_int16_pairs_k22_k23_match_pair_kernel(ctx, A, k):
    # k = 22 / 23 run the int16-pair passes with a first pass of 12 / 13 bits
    # (one / two waves per workgroup) and the 10-bit last pass
    raw_s, ms_s = ctx.an_pairs_spectral(A, k)
    raw_p, ms_p = ctx.an_pairs(A, 1 << k)
    print(f"A={A} k={k}: spectral {ms_s:.1f} ms, pair kernel {ms_p:.1f} ms", flush=True)
    assert raw_s == raw_p


def test_k25_multi_transform_matches_pair_kernel(ctx):
    # A = 61 at k = 25: 31-bit codewords, 64 transforms of 2^25 through the
    # middle pass and the narrow last pass (per-transform weight base and high
    # bits), bit for bit against the pair kernel's 2^49 pairs (about a minute)
    raw_s, ms_s = ctx.an_pairs_spectral(61, 25)
    raw_p, ms_p = ctx.an_pairs(61, 1 << 25)
    print(f"k=25 A=61: spectral {ms_s:.1f} ms, pair kernel {ms_p:.1f} ms", flush=True)
    assert raw_s == raw_p


@pytest.mark.parametrize("k", [17, 18, 19, 20, 21, 25, 26, 27, 28, 29, 30, 31])
def test_closed_form_k17_to_31(ctx, k):
    # A = 1: D(d) = 2^k C(k, d); one transform, so k = 17 .. 21 take the
    # int32 passes (the 16-bit first pass and the narrow last pass), k >= 25
    # the middle pass too
    import math
    raw, _ = ctx.an_pairs_spectral(1, k)
    assert raw[1:k + 1] == [(1 << (k - 1)) * math.comb(k, d) for d in range(1, k + 1)]


def test_k28_totals(ctx):
    raw, ms = ctx.an_pairs_spectral(61, 28)
    print(f"k=28 A=61: {ms:.1f} ms", flush=True)
    N = 1 << 28
    assert sum(raw) == N * (N - 1) // 2 and raw[1] == 0


def test_k24_recorded_runs(ctx):
    # results/runtimes/*/ancoding_gpu_A*_n24_*.csv and results/As/ancoding_gpu_A*_n24_*.csv:
    # 30- to 40-bit codewords
    runs = [r for r in golden("an_exhaustive_runs.json")["runs"] if r["n"] == 24]
    assert len(runs) == 23
    done = {}
    for r in runs:
        if r["A"] not in done:
            done[r["A"]], ms = ctx.an_pairs_spectral(r["A"], 24)
            print(f"k=24 A={r['A']}: {ms:.1f} ms", flush=True)
        assert report(24, done[r["A"]], len(r["counts"])) == r["counts"], r["src"]


def test_k24_golden_vectors(ctx, solutions):
    for A in (61, 641):
        raw, _ = ctx.an_pairs_spectral(A, 24)
        assert padded_eq(report(24, raw, 31 if A == 61 else 35), solutions[f"solution_an24_A{A}"])


@pytest.mark.parametrize("A", [1, 3, 61, 64 * 61, 641, 32771, 58659, 65521, 65535, 3 << 20])
def test_k16_matches_pair_kernel(ctx, A):
    # even A (3 << 20: odd part 3), A = 1, and 32-bit codewords (n = 2k) included
    raw_s, _ = ctx.an_pairs_spectral(A, 16)
    raw_p, _ = ctx.an_pairs(A, 1 << 16)
    assert raw_s == raw_p


def test_k24_headline_code_matches_pair_kernel(ctx):
    # A = 58659: 40-bit codewords, 2^16 transforms of length 2^24
    raw_s, ms_s = ctx.an_pairs_spectral(58659, 24)
    raw_p, ms_p = ctx.an_pairs(58659, 1 << 24)
    print(f"k=24 A=58659: spectral {ms_s:.1f} ms, pair kernel {ms_p:.1f} ms", flush=True)
    assert raw_s == raw_p
    N = 1 << 24
    assert sum(raw_s) == N * (N - 1) // 2


@pytest.mark.parametrize("A,k", [(61, 16), (58659, 16), (641, 18)])
def test_auto_dispatch_matches_pair_kernel(ctx, A, k):
    # cr_an_pairs takes the spectral path for (61, 16), the pair kernel for
    # the others; either way the histogram is the pair kernel's
    raw_a, _ = ctx.an_pairs_auto(A, 1 << k)
    raw_p, _ = ctx.an_pairs(A, 1 << k)
    assert raw_a == raw_p


def test_auto_dispatch_k24(ctx, solutions):
    raw, ms = ctx.an_pairs_auto(61, 1 << 24)
    print(f"cr_an_pairs k=24 A=61: {ms:.1f} ms", flush=True)
    assert ms < 1000  # the spectral path, not the 15 s pair kernel
    assert padded_eq(report(24, raw, 31), solutions["solution_an24_A61"])


@pytest.mark.parametrize("A", [1, 1 << 7])
def test_k32_closed_form(ctx, A):
    # all 2^32 words (odd part 1): D(d) = 2^32 C(32, d), so raw[d] = 2^31 C(32, d);
    # the transform reaches F = 2^32 at S = 0, the 64-bit end of the last pass
    import math
    free_before, _ = ctx.device_memory()
    raw, ms = ctx.an_pairs_spectral(A, 32)
    print(f"k=32 A={A}: {ms:.1f} ms", flush=True)
    assert raw[0] == 0
    assert raw[1:33] == [(1 << 31) * math.comb(32, d) for d in range(1, 33)]
    assert all(x == 0 for x in raw[33:])
    # the 16 GiB transform buffer is not kept by the context
    free_after, _ = ctx.device_memory()
    assert free_after >= free_before - (64 << 20), (free_before, free_after)


@pytest.mark.parametrize("A", [61, 3])
def test_k32_exact_totals(ctx, A):
    # exact k = 32 distributions (beyond the pair kernel: 2^63 pairs); the
    # library checks that every Krawtchouk sum divides by 2^n and D(0) = 2^32
    raw, ms = ctx.an_pairs_spectral(A, 32)
    print(f"k=32 A={A}: {ms:.1f} ms", flush=True)
    N = 1 << 32
    assert sum(raw) == N * (N - 1) // 2
    n = (A * (N - 1)).bit_length()
    assert raw[0] == 0 and all(x == 0 for x in raw[n + 1:])
    # no multiple of an odd A > 1 is a power of two: no distance 1; distance 2
    # occurs (A = 3: 3 = 2 + 1; A = 61: 2^30 + 1 = 61 * 17602325, below 2^32)
    assert raw[1] == 0 and raw[2] > 0
    raw_auto, _ = ctx.an_pairs_auto(A, N)
    assert raw_auto == raw


def test_k32_exact_against_recorded_grid_approximations(ctx):
    # results/runtimes/*/ancoding_grid1d_*_m{1001,10000}_A61_n32_h6.csv: the
    # reference's own k=32 approximations converge on the exact distribution
    raw, _ = ctx.an_pairs_spectral(61, 32)
    exact = [1 << 32] + [2 * x for x in raw[1:]]
    assert sum(exact) == 1 << 64
    runs = golden("an_grid1d_runs.json")["runs"]
    err, l1 = {}, {}
    for m in (1001, 10000):
        g = [r for r in runs if r["n"] == 32 and r["A"] == 61 and r["m"] == m][0]["counts"]
        rel = {d: abs(g[d] / exact[d] - 1) for d in range(1, 38) if exact[d] > 0}
        err[m] = max(rel.values())
        l1[m] = sum(abs(g[d] - exact[d]) for d in range(39)) / (1 << 64)
        # the bulk (bins 4..33, counts above 3e13) within 1e-3; the tails carry
        # the grid's sampling noise
        assert all(rel[d] < 1e-3 for d in range(4, 34)), m
    assert err[1001] < 0.02 and err[10000] < 0.005 and l1[10000] < l1[1001]


def test_unsupported_and_invalid(ctx):
    import coding_reliability_amd as cra
    for k in (13, 33):
        with pytest.raises(cra.CRError) as e:
            ctx.an_pairs_spectral(61, k)
        assert e.value.status == 2  # CR_ERR_UNSUPPORTED
    with pytest.raises(cra.CRError) as e:
        ctx.an_pairs_spectral(1 << 16 | 1, 32)  # 49-bit codewords
    assert e.value.status == 2  # CR_ERR_UNSUPPORTED
    with pytest.raises(cra.CRError) as e:
        ctx.an_pairs_spectral(65539, 16)  # 65539 * 65535 has 33 bits: more transforms than pairs
    assert e.value.status == 2  # CR_ERR_UNSUPPORTED
    with pytest.raises(cra.CRError) as e:
        ctx.an_pairs_spectral((1 << 33) + 1, 24)
    assert e.value.status == 2
    with pytest.raises(cra.CRError):
        ctx.an_pairs_spectral(0, 16)
    with pytest.raises(cra.CRError):
        ctx.an_pairs_spectral(61, 16, nbins=5)
