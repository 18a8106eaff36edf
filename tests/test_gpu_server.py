"""GPU tests of the continuous-batching request loop: every result is the same
alone and in a mix (bit-exact; Monte Carlo through fixed Philox counters),
admission by HBM footprint, refusal, cancellation."""
import time

import pytest

import coding_reliability_amd as cra
from conftest import padded_eq
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = cra.Context(0)
    yield c
    c.close()


MIX = [
    dict(kind=cra.JOB_AN_EXHAUSTIVE, k=18, A=58659),
    dict(kind=cra.JOB_AN_GRID, k=32, A=61, stride=65537 * 16),
    dict(kind=cra.JOB_HAMMING_MC, k=64, samples=1 << 24, seed=5, stream=1),
    dict(kind=cra.JOB_AN_EXHAUSTIVE, k=16, A=641),
    dict(kind=cra.JOB_HAMMING_MC, k=32, samples=(1 << 22) + 12, seed=9, stream=2),
    dict(kind=cra.JOB_AN_EXHAUSTIVE, k=20, A=32771),
    dict(kind=cra.JOB_AN_GRID, k=24, A=12345, stride=4097),
]


def direct(ctx, j):
    """The same request through the one-shot entry points."""
    if j["kind"] == cra.JOB_AN_EXHAUSTIVE:
        return ctx.an_pairs(j["A"], 1 << j["k"])[0]
    if j["kind"] == cra.JOB_AN_GRID:
        s = j["stride"]
        return ctx.an_pairs(j["A"] * s, ((1 << j["k"]) + s - 1) // s)[0]
    return ctx.hamming_weights(j["k"], cra.HAM_MC, first=0, count=j["samples"], seed=j["seed"],
                               stream=j["stream"])[0]


def test_spectral_requests_in_a_mix(ctx, solutions):
    # exhaustive requests the spectral path is cheaper for (A = 61 at k = 16
    # and 24, A = 58659 at k = 24: 2.4 s instead of 15 s) are served as
    # batches of transforms beside pair-kernel and Monte Carlo requests
    jobs = [dict(kind=cra.JOB_AN_EXHAUSTIVE, k=24, A=58659),
            dict(kind=cra.JOB_AN_EXHAUSTIVE, k=24, A=61),
            dict(kind=cra.JOB_AN_EXHAUSTIVE, k=18, A=58659),
            dict(kind=cra.JOB_AN_EXHAUSTIVE, k=16, A=61),
            dict(kind=cra.JOB_HAMMING_MC, k=32, samples=1 << 22, seed=3, stream=4)]
    t0 = time.perf_counter()
    with cra.Server(ctx, iteration_ms=5) as srv:
        ids = [srv.submit(**j) for j in jobs]
        for rid in ids:
            srv.wait(rid)
        got = [srv.result(rid)[0] for rid in ids]
    dt = time.perf_counter() - t0
    print(f"spectral mix served in {dt:.2f} s", flush=True)
    assert dt < 10  # the k = 24, A = 58659 request alone takes 15 s on the pair kernel
    assert got[0] == ctx.an_pairs_spectral(58659, 24)[0][: len(got[0])]
    assert padded_eq(O.an_report(24, got[1], 31), solutions["solution_an24_A61"])
    for j, g in zip(jobs[2:], got[2:]):
        assert g == direct(ctx, j)[: len(g)], j
    N = 1 << 24
    assert sum(got[0]) == N * (N - 1) // 2


def test_alone_equals_mix(ctx):
    ref = [direct(ctx, j) for j in MIX]
    # alone: one request at a time
    with cra.Server(ctx, iteration_ms=5) as srv:
        alone = []
        for j in MIX:
            rid = srv.submit(**j)
            srv.wait(rid)
            alone.append(srv.result(rid)[0])
    # mix: all at once, short iterations so the requests interleave
    with cra.Server(ctx, iteration_ms=2) as srv:
        ids = [srv.submit(**j) for j in MIX]
        for rid in ids:
            srv.wait(rid)
        mixed = [srv.result(rid) for rid in ids]
        st = srv.stats()
    for j, r, a, (m, info) in zip(MIX, ref, alone, mixed):
        assert a == r[: len(a)], j
        assert m == r[: len(m)], j
    assert max(info["iterations"] for _, info in mixed) > 1
    assert st["completed"] == len(MIX)


def test_exhaustive_request_vs_oracle(ctx):
    with cra.Server(ctx) as srv:
        rid = srv.submit(cra.JOB_AN_EXHAUSTIVE, 12, A=58659)
        srv.wait(rid)
        h, info = srv.result(rid)
    assert h == O.an_pairs_raw(58659, 1 << 12, len(h))
    assert info["units"] == (1 << 12) * ((1 << 12) - 1) // 2


def test_admission_by_footprint(ctx):
    # budget for exactly one histogram slot: requests run one after another
    with cra.Server(ctx, hbm_budget_bytes=1280, iteration_ms=2) as srv:
        ids = [srv.submit(cra.JOB_AN_EXHAUSTIVE, 16, A=A) for A in (61, 641, 58659)]
        for rid in ids:
            srv.wait(rid)
        infos = [srv.result(rid)[1] for rid in ids]
    # each admitted only after the previous one finished
    for a, b in zip(infos, infos[1:]):
        assert b["admit_s"] >= a["finish_s"] - 1e-6


def test_rejected_when_over_budget(ctx):
    with cra.Server(ctx, hbm_budget_bytes=64) as srv:
        with pytest.raises(cra.CRError) as e:
            srv.submit(cra.JOB_AN_EXHAUSTIVE, 10, A=61)
        assert e.value.status == 6
        assert srv.stats()["rejected"] == 1


def test_cancel(ctx):
    with cra.Server(ctx, iteration_ms=5) as srv:
        big = srv.submit(cra.JOB_AN_EXHAUSTIVE, 26, A=58659)  # minutes of work
        small = srv.submit(cra.JOB_AN_EXHAUSTIVE, 14, A=61)
        srv.wait(small, 60000)
        srv.cancel(big)
        srv.wait(big, 60000)
        with pytest.raises(cra.CRError) as e:
            srv.result(big)
        assert e.value.status == 7
        h, _ = srv.result(small)
        assert h == O.an_pairs_raw(61, 1 << 14, len(h))


def test_invalid_requests(ctx):
    with cra.Server(ctx) as srv:
        with pytest.raises(cra.CRError):
            srv.submit(7, 10)
        with pytest.raises(cra.CRError):
            srv.submit(cra.JOB_AN_EXHAUSTIVE, 0, A=3)
        with pytest.raises(cra.CRError):
            srv.result(12345)


def test_large_request_runs_at_device_rate(ctx):
    """A request of 8.8e12 pairs (k=22) served alone finishes in a small
    multiple of the one-shot kernel time: each iteration fills the device (at
    least two work items per wave), so the slice size cannot shrink towards a
    handful of items per iteration."""
    j = dict(kind=cra.JOB_AN_EXHAUSTIVE, k=22, A=58659)
    ref, kms = ctx.an_pairs(j["A"], 1 << j["k"])
    with cra.Server(ctx, iteration_ms=20) as srv:
        t0 = time.perf_counter()
        rid = srv.submit(**j)
        srv.wait(rid, timeout_ms=60000)
        dt = time.perf_counter() - t0
        got, info = srv.result(rid)
    assert got == ref[: len(got)]
    assert dt < 3 * kms / 1e3 + 1.0, (dt, kms, info)


def test_burst_of_small_requests(ctx):
    # many short requests at once: each result equals its one-shot launch,
    # finished requests retire together (one copy-back synchronisation per iteration)
    import random
    rnd = random.Random(3)
    jobs = [(2 * rnd.randrange(1 << 14, 1 << 15) + 1, rnd.randrange(6, 13)) for _ in range(300)]
    with cra.Server(ctx, iteration_ms=2, max_active=128) as srv:
        ids = [srv.submit(kind=cra.JOB_AN_EXHAUSTIVE, k=k, A=A) for A, k in jobs]
        for rid in ids:
            srv.wait(rid, timeout_ms=60000)
        got = [srv.result(rid)[0] for rid in ids]
        st = srv.stats()
    for (A, k), g in zip(jobs, got):
        ref = ctx.an_pairs(A, 1 << k)[0]
        assert g == ref[: len(g)], (A, k)
    assert st["completed"] == len(jobs)
