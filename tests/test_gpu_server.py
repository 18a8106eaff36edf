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


def test_c5_served_mix_at_stated_sizes(ctx):
    """The served configuration at its stated sizes, with seeded Poisson
    arrivals: exhaustive AN requests (k uniform in 18..24, odd 16-bit A),
    integer-grid AN requests (k = 32, stride 2^12 + 1 = 4097: the words
    A*(4097 i) for 4097 i < 2^32, i.e. 1048321 words, every unordered pair)
    and Monte Carlo (72,64) Extended Hamming requests of 2^30 samples each.
    Every result equals the one-shot entry point's (cr_an_pairs, which takes
    the spectral path where it is cheaper; cr_hamming_weights)."""
    import numpy as np
    g = np.random.Generator(np.random.Philox(20261203))
    kinds = [0, 1, 2] * 4
    g.shuffle(kinds)
    jobs, arrivals, t = [], [], 0.0
    for i, kind in enumerate(kinds):
        A = int(2 * g.integers(1 << 14, 1 << 15) + 1)
        if kind == 0:
            jobs.append(dict(kind=cra.JOB_AN_EXHAUSTIVE, k=int(g.integers(18, 25)), A=A))
        elif kind == 1:
            jobs.append(dict(kind=cra.JOB_AN_GRID, k=32, A=A, stride=4097))
        else:
            jobs.append(dict(kind=cra.JOB_HAMMING_MC, k=64, samples=1 << 30, seed=20261203, stream=i))
        t += float(g.exponential(1 / 4.0))  # 4 requests per second
        arrivals.append(t)
    print("C5 jobs:", jobs, flush=True)
    with cra.Server(ctx, iteration_ms=20) as srv:
        t0 = srv.now()
        ids = []
        for j, at in zip(jobs, arrivals):
            delay = at - (srv.now() - t0)
            if delay > 0:
                time.sleep(delay)
            ids.append(srv.submit(**j))
        for rid in ids:
            srv.wait(rid, timeout_ms=120000)
        got = [srv.result(rid) for rid in ids]
        st = srv.stats()
    assert st["completed"] == len(jobs)
    lat = sorted(info["finish_s"] - info["submit_s"] for _, info in got)
    print(f"C5 served: p50 {lat[len(lat) // 2]:.3f} s, max {lat[-1]:.3f} s, {st['iterations']} iterations",
          flush=True)
    for j, (h, info) in zip(jobs, got):
        if j["kind"] == cra.JOB_AN_EXHAUSTIVE:
            want = ctx.an_pairs_auto(j["A"], 1 << j["k"])[0]
            N = 1 << j["k"]
            assert info["units"] == N * (N - 1) // 2
        elif j["kind"] == cra.JOB_AN_GRID:
            m = ((1 << 32) + 4096) // 4097
            assert m == 1048321
            want = ctx.an_pairs(j["A"] * 4097, m)[0]
            assert sum(h) == m * (m - 1) // 2 and info["units"] == m * (m - 1) // 2
        else:
            want = direct(ctx, j)
            assert sum(h) == 1 << 30
        assert h == want[: len(h)] and not any(want[len(h):]), j


@pytest.mark.parametrize("A", [16385, 58659])
def test_k32_stride_grid_request_vs_oracle(ctx, A):
    # the C5 grid request's code path (k = 32, 16-bit A, multiplier A*stride
    # above 2^22: the generic two-dword pair mode) at a stride the oracle
    # enumerates quickly: 4097*64 = 262208 gives 16381 words
    s = 4097 * 64
    m = ((1 << 32) + s - 1) // s
    with cra.Server(ctx, iteration_ms=5) as srv:
        rid = srv.submit(cra.JOB_AN_GRID, 32, A=A, stride=s)
        srv.wait(rid, timeout_ms=60000)
        h, info = srv.result(rid)
    want = O.an_pairs_raw(A * s, m, len(h))
    assert h == want
    assert info["units"] == m * (m - 1) // 2
