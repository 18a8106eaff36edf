"""GPU parity of the batched pair kernel and its captured hipGraph plan:
per-code histograms are bit-exact against the oracle and against one-code
launches, on mixed batches and at the full batched configuration."""
import numpy as np
import pytest

import coding_reliability_amd as cra
from coding_reliability_amd import workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = cra.Context(0)
    yield c
    c.close()


def mixed_codes(n=200, seed=5):
    g = np.random.Generator(np.random.Philox(seed))
    codes = []
    for _ in range(n):
        k = int(g.integers(1, 15))
        A = int(2 * g.integers(1, 1 << 15) + 1)
        codes.append((A, 1 << k))
    # odd sizes, large multipliers (64-bit codewords), duplicates
    codes += [(58659, 5000), (3, 777), ((1 << 40) + 3, 1000), ((1 << 20) + 1, 3000), (61, 256), (61, 256)]
    return codes


@pytest.mark.parametrize("R,ipw", [(0, 0), (4, 4), (8, 64), (16, 16), (2, 1)])
def test_batch_vs_oracle(ctx, R, ipw):
    codes = mixed_codes()
    got, ms = ctx.an_batch(codes, rows_per_lane=R, items_per_wave=ipw)
    for (A, N), h in zip(codes, got):
        assert h == O.an_pairs_raw(A, N, len(h)), (A, N)


def test_plan_graph_replays(ctx):
    codes = mixed_codes(120, seed=9)
    eager, _ = ctx.an_batch(codes)
    plan = ctx.batch_plan(codes, rows_per_lane=4, items_per_wave=8)
    for rep in range(4):
        got, ms = plan.run()
        assert got == eager, rep
        # eager work between replays
        other, _ = ctx.an_pairs(61, 1 << 12)
        assert sum(other) == (1 << 12) * ((1 << 12) - 1) // 2
    plan.close()


def test_c4_full_batch_matches_single_launches(ctx):
    """The batched configuration (4096 codes, k in 8..18): one launch equals
    4096 one-code launches bit for bit; k <= 12 codes also checked on the CPU."""
    codes = workloads.mixed_codes()
    got, ms = ctx.an_batch(codes)
    print(f"4096-code batch: {ms:.1f} ms, {workloads.pairs_of(codes) / ms / 1e9:.2f} Tpairs/s", flush=True)
    for i, ((A, N), h) in enumerate(zip(codes, got)):
        single, _ = ctx.an_pairs(A, N)
        assert h == single, (i, A, N)
        assert sum(h) == N * (N - 1) // 2
    small = [(i, c) for i, c in enumerate(codes) if c[1] <= 1 << 12][:200]
    for i, (A, N) in small:
        assert got[i] == O.an_pairs_raw(A, N, len(got[i]))


def test_autotune_small(ctx):
    codes = workloads.mixed_codes(64, seed=3, kmax=14)
    best, ms = ctx.autotune_batch(codes, repeats=1)
    assert best["rows_per_lane"] in (4, 8, 16) and best["items_per_wave"] in (4, 16, 64) and ms > 0
    got, _ = ctx.an_batch(codes, **best)
    for (A, N), h in zip(codes, got):
        assert h == O.an_pairs_raw(A, N, len(h))


def test_batch_invalid(ctx):
    with pytest.raises(cra.CRError):
        ctx.an_batch([(0, 16)])
    with pytest.raises(cra.CRError):
        ctx.an_batch([((1 << 63) + 1, 4)])


def test_repeated_create_destroy_leaks_nothing():
    # contexts, batch plans and servers created and destroyed many times:
    # device memory in use returns to where it started
    import coding_reliability_amd as cra

    probe = cra.Context(0)

    def used():
        free, total = probe.device_memory()
        return total - free

    codes = [(61, 1 << 10), (58659, 1 << 12)]
    c0 = cra.Context(0)
    ref = c0.an_batch(codes)[0]
    c0.close()
    base = used()
    for _ in range(20):
        c = cra.Context(0)
        assert c.an_batch(codes)[0] == ref
        p = c.batch_plan(codes)
        assert p.run()[0] == ref
        p.close()
        with cra.Server(c, max_active=32) as srv:
            rid = srv.submit(kind=cra.JOB_AN_EXHAUSTIVE, k=10, A=61)
            srv.wait(rid, timeout_ms=30000)
            assert srv.result(rid)[0] == ref[0][: len(srv.result(rid)[0])]
        c.close()
    assert used() - base < (64 << 20)
    probe.close()
