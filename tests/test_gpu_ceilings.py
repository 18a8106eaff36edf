"""The device ceilings the bench states its rooflines against
(cr_probe_ceiling): every probe runs, and the kernels they bound do not
exceed them (a ceiling below a kernel's measured rate would be no ceiling)."""
import pytest

import coding_reliability_amd as cra

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = cra.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def ceilings(ctx):
    iters = {0: 1 << 14, 1: 1 << 11, 2: 1 << 13, 3: 1 << 11, 4: 1 << 13, 5: 1 << 13, 8: 1 << 12}
    out = {}
    for p, it in iters.items():
        ctx.probe_ceiling(p, it)
        out[p] = max(ctx.probe_ceiling(p, it)[0] for _ in range(2))
    print({p: round(v / 1e12, 3) for p, v in out.items()}, flush=True)
    return out


def test_probes_run(ceilings):
    for p, v in ceilings.items():
        assert v > 1e12, p
    # the VALU probes agree within a few percent: one issue ceiling
    valu = [ceilings[p] for p in (0, 3, 4)]
    assert min(valu) > 0.85 * max(valu)


def test_pair_kernel_within_ceilings(ctx, ceilings):
    # one LDS increment and three VALU instructions per pair (xor, popcount,
    # the increment's address)
    A, N = 58659, 1 << 21
    ctx.an_pairs(A, N)
    raw, ms = ctx.an_pairs(A, N)
    rate = N * (N - 1) // 2 / (ms / 1e3)
    lds = max(ceilings[2], ceilings[5])
    valu = max(ceilings[p] for p in (0, 1, 3, 4))
    print(f"pair kernel {rate / 1e12:.3f} Tpairs/s: LDS {rate / lds:.3f}, VALU {3 * rate / valu:.3f}", flush=True)
    assert rate <= 1.03 * lds
    assert 3 * rate <= 1.03 * valu


def test_hamming_mc_within_lds_ceilings(ctx, ceilings):
    # (72,64) Monte Carlo: 10 table reads and one increment per word share the
    # LDS, so the time fractions of the two LDS ceilings add up to at most 1
    S = 1 << 28
    ctx.hamming_weights(64, cra.HAM_MC, 0, S, seed=3, stream=1)
    h, ms = ctx.hamming_weights(64, cra.HAM_MC, 0, S, seed=3, stream=1)
    assert sum(h) == S
    rate = S / (ms / 1e3)
    busy = rate * 10 / ceilings[8] + rate / max(ceilings[2], ceilings[5])
    print(f"(72,64) MC {rate / 1e9:.1f} Gwords/s: LDS busy {busy:.3f}", flush=True)
    assert busy <= 1.03


def test_streaming_probes(ctx):
    # 16 passes of 16-byte stores / loads over 256 MiB, bytes per second
    for p in (6, 7):
        ctx.probe_ceiling(p, 4)
        rate, ms = ctx.probe_ceiling(p, 16)
        print(f"probe {p}: {rate / 1e12:.3f} TB/s ({ms:.3f} ms)", flush=True)
        assert 5e11 < rate < 3e13
    with pytest.raises(cra.CRError):
        ctx.probe_ceiling(6, 65)


def test_invalid_probe(ctx):
    with pytest.raises(cra.CRError):
        ctx.probe_ceiling(9)
    with pytest.raises(cra.CRError):
        ctx.probe_ceiling(0, 0)
