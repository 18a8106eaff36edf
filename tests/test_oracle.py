"""Pins the CPU oracle against the reference's own golden vectors, its recorded
result files, brute force and closed forms.  CPU only (no GPU needed)."""
import math

import numpy as np
import pytest

from conftest import golden, macwilliams_weights, padded_eq
from oracle import oracle as O


# ---------------------------------------------------------------- exhaustive AN

@pytest.mark.parametrize("A,k", [(61, 8), (641, 8), (61, 16), (641, 16)])
def test_an_cpu_counts_golden(solutions, A, k):
    got, pairs = O.an_cpu_counts(k, A)
    assert padded_eq(got, solutions[f"solution_an{k}_A{A}"])
    assert pairs == (1 << k) * ((1 << k) - 1) // 2


@pytest.mark.parametrize("A,k", [(61, 8), (641, 8), (61, 16), (641, 16)])
def test_an_pairs_raw_golden(solutions, A, k):
    h = O.ceil_log2(A)
    raw = O.an_pairs_raw(A, 1 << k)
    assert padded_eq(O.an_report(k, raw, k + h + 1), solutions[f"solution_an{k}_A{A}"])


def test_golden_vectors_sum():
    # every reference vector of the exhaustive AN path sums to 4^k
    sol = golden("solutions.json")["vectors"]
    for k in (8, 16, 24):
        for A in (61, 641):
            assert sum(int(x) for x in sol[f"solution_an{k}_A{A}"]) == 4 ** k


def _runs(max_n):
    return [r for r in golden("an_exhaustive_runs.json")["runs"] if r["n"] <= max_n]


def test_as_fixtures_up_to_k14():
    """All exact results/As and results/runtimes files with k <= 14 (A odd 1..127,
    A=61): the reference CPU loop nest reproduces them bit for bit."""
    runs = _runs(14)
    assert len(runs) >= 400
    for r in runs:
        A = r["A"] if (r["A"] & (r["A"] - 1)) else 1  # GPU driver maps powers of two to 1
        got, _ = O.an_cpu_counts(r["n"], A, count_counts=len(r["counts"]))
        assert got == r["counts"], r["src"]


@pytest.mark.parametrize("A", [35, 61, 101, 127])
def test_as_fixtures_k16(A):
    runs = [r for r in _runs(16) if r["n"] == 16 and r["A"] == A]
    assert runs
    raw = O.an_pairs_raw(A, 1 << 16)
    for r in runs:
        assert O.an_report(16, raw, len(r["counts"])) == r["counts"], r["src"]


def test_bruteforce_tiny():
    for A in (1, 3, 5, 7, 19, 53, 61, 255, 58659):
        for k in (1, 2, 3, 5, 7):
            N = 1 << k
            raw = [0] * 65
            for i in range(N):
                for j in range(i + 1, N):
                    raw[bin((A * i) ^ (A * j)).count("1")] += 1
            assert O.an_pairs_raw(A, N) == raw, (A, k)


def test_closed_form_A1():
    # A=1: codewords are all k-bit words, raw[d] = 2^k * C(k,d) / 2 for d >= 1
    for k in (4, 9, 12):
        raw = O.an_pairs_raw(1, 1 << k)
        for d in range(1, k + 1):
            assert raw[d] == (1 << k) * math.comb(k, d) // 2
        assert raw[0] == 0 and all(x == 0 for x in raw[k + 1:])


def test_invariants_shift_and_sum():
    # A*2^s has the same distances as A; sum of raw = C(2^k, 2)
    k = 10
    base = O.an_pairs_raw(61, 1 << k)
    assert O.an_pairs_raw(61 << 5, 1 << k) == base
    assert sum(base) == (1 << k) * ((1 << k) - 1) // 2


def test_codeword_matrices():
    for m in golden("codeword_matrices.json")["matrices"]:
        raw = O.an_pairs_raw(m["A"], 1 << m["k"])
        assert O.an_report(m["k"], raw, len(m["hist"])) == m["hist"], m["A"]


def test_shard_sample_is_subset():
    # a bounded row-shard sample counts exactly the pairs it reports
    c, pairs = O.an_cpu_counts(12, 58659, 3, 7)
    assert pairs == 4 * (64 * 63 // 2) + sum(64 * 64 * ((1 << 12) // 64 - 1 - s) for s in range(3, 7))
    assert (sum(c) - 4 * 64) // 2 == pairs


# ---------------------------------------------------------------- reference grid

def _grid_runs(max_n):
    return [r for r in golden("an_grid1d_runs.json")["runs"] if r["n"] <= max_n and not r["four_gpu"]]


def test_grid1d_fixtures_k_le_16():
    runs = _grid_runs(16)
    assert len(runs) >= 30
    for r in runs:
        hist, rep, cfg = O.an_grid_ref(1, r["n"], r["m"], 0, r["A"])
        older = "/runtimes/" in r["src"] and r["m"] == 1000
        if older:
            # these files come from an earlier driver that did not overwrite
            # counts[0] = 2^k with the scaled self-hit count (cf. an_coding_grid.cu:317-322)
            assert rep[1:len(r["counts"])] == r["counts"][1:], r["src"]
        else:
            assert rep[:len(r["counts"])] == r["counts"], r["src"]


def test_grid1d_fixture_k24_small_m():
    runs = [r for r in _grid_runs(24) if r["n"] == 24 and r["m"] <= 120][:3]
    assert runs
    for r in runs:
        hist, rep, cfg = O.an_grid_ref(1, r["n"], r["m"], 0, r["A"])
        assert rep[:len(r["counts"])] == r["counts"], r["src"]


def test_grid1d_m_equals_2k_is_exhaustive():
    # m = 2^k: the sample list is every codeword; ordered pairs incl. diagonal
    k, A = 10, 61
    hist, rep, cfg = O.an_grid_ref(1, k, 1 << k, 0, A)
    raw = O.an_pairs_raw(A, 1 << k)
    assert hist[0] == 1 << k
    assert all(hist[b] == 2 * raw[b] for b in range(1, 65))


def test_grid2d_m_equals_2k_is_exhaustive():
    k, A = 9, 53
    hist, rep, cfg = O.an_grid_ref(2, k, 1 << k, 1 << k, A)
    raw = O.an_pairs_raw(A, 1 << k)
    assert hist[0] == 0 and all(hist[b] == 2 * raw[b] for b in range(1, 65))


# ---------------------------------------------------------------- Hamming

@pytest.mark.parametrize("k,name", [(8, "sol_hamming_n8"), (16, "sol_hamming_n16"), (24, "sol_hamming_n24")])
def test_hamming_golden(solutions, k, name):
    rep = O.hamming_report(k, O.hamming_weights(k), True)
    sol = solutions[name]
    for i in range(4, len(sol), 2):
        assert rep[i] == sol[i] << k
    assert O.hamming_N(k) == len(sol) - 1


@pytest.mark.parametrize("k", [8, 12, 16, 24])
def test_hamming_macwilliams(k):
    mw = macwilliams_weights(k)
    h = O.hamming_weights(k)
    # the overall parity bit lifts odd non-extended weights to the next even one
    E = [(h[w] + (h[w - 1] if w else 0)) if w % 2 == 0 else 0 for w in range(len(mw))]
    assert E == mw


def test_hamming_golden_n32_macwilliams(solutions):
    # the (39,32) reference vector equals the closed form (no 2^32 loop needed)
    mw = macwilliams_weights(32)
    sol = solutions["sol_hamming_n32"]
    assert all(mw[i] == sol[i] for i in range(4, len(sol), 2))


@pytest.mark.parametrize("k", [4, 8])
def test_hamming_spheres_bruteforce(k):
    # closed-form sphere bins (hamming.cu:194-205) == explicit sphere enumeration
    # (hamming_native_short.cpp:52-77), for distances >= 3
    closed = O.hamming_report(k, O.hamming_weights(k), True)
    brute = O.hamming_bruteforce(k, True)
    assert closed[3:] == brute[3:]


def test_hamming_72_64_parity_bits():
    assert O.hamming_N(64) == 72 and O.hamming_N(32) == 39 and O.hamming_N(16) == 22 and O.hamming_N(8) == 13


def test_hamming_codeword_layout_k8():
    # computeHamming08 (hamming.h:22-30): data << 5, parity t at bit t+1, overall at bit 0
    for d in range(256):
        cw = O.hamming_codeword_ext(d, 8)
        p = 0
        for t, mk in enumerate((0x5B, 0x6D, 0x8E, 0xF0)):
            p |= (bin(d & mk).count("1") & 1) << (t + 1)
        p |= (bin(d).count("1") + bin(p).count("1")) & 1
        assert cw == (d << 5) | p


# ---------------------------------------------------------------- Philox / MC

def test_philox_known_answers():
    # Random123 known-answer vectors for philox4x32_10
    assert O.philox4x32_10([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert O.philox4x32_10([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert O.philox4x32_10([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_philox_rocrand_fixture():
    fx = golden("philox_vectors.json")
    for v in fx["vectors"]:
        assert O.philox4x32_10(v["ctr"], v["key"]) == v["out"]


def test_mc_word_uniformity():
    # chi-square sanity check of the 8-bit words over 2^16 samples
    k = 8
    cnt = np.zeros(256)
    for s in range(1 << 14):
        cnt[O.mc_word(12345, 0, s, k)] += 1
    exp = (1 << 14) / 256
    chi2 = ((cnt - exp) ** 2 / exp).sum()
    assert chi2 < 255 + 6 * math.sqrt(2 * 255)


def test_hamming_mc_4sigma():
    k, S = 16, 1 << 18
    h = O.hamming_mc(k, 7, 0, 0, S)
    truth = O.hamming_weights(k)
    for w, t in enumerate(truth):
        p = t / (1 << k)
        assert abs(h[w] - S * p) <= 4 * math.sqrt(S * p * (1 - p)) + 1


def test_an_mc_4sigma():
    A, n, m = 61, 10, 64
    h = O.an_mc(A, n, m, 99)
    raw = O.an_pairs_raw(A, 1 << n)
    tot = (1 << n) * m
    for b in range(65):
        # ordered pairs incl. the diagonal
        p = (2 * raw[b] + ((1 << n) if b == 0 else 0)) / (1 << (2 * n))
        assert abs(h[b] - tot * p) <= 4 * math.sqrt(tot * p * (1 - p)) + 1, b


# ---------------------------------------------------------------- ANcandidates

def test_naf_identity():
    for x in range(1, 5000):
        assert O.naf_weight(x) == bin(x ^ (3 * x)).count("1")


@pytest.mark.parametrize("k,abits", [(4, 3), (6, 4), (8, 5), (10, 6)])
def test_an_candidates_bruteforce(k, abits):
    best, cands = O.an_candidates(k, abits)
    vals = {A: min(O.naf_weight(A * d) for d in range(1, 1 << k))
            for A in range((1 << (abits - 1)) + 1, 1 << abits, 2)}
    assert best == max(vals.values())
    assert cands == sorted(A for A, v in vals.items() if v == best)


def test_reported_form_from_macwilliams_equals_sphere_enumeration():
    # the test helper that turns an extended weight distribution into
    # run_hamming's reported form, pinned by the explicit sphere enumeration
    # of hamming_native_short.cpp:52-77 (oracle brute force)
    from conftest import hamming_reported_from_ext
    for k in (4, 8):
        for b in (False, True):
            assert hamming_reported_from_ext(k, macwilliams_weights(k), b) == O.hamming_bruteforce(k, b)


def test_an_mc_sample_counter_layout():
    # the AN Monte Carlo sample words through the Philox generator (pinned by
    # its known-answer vectors): counter (b, b >> 32, i mod 2^32,
    # 0x43520001 ^ (i >> 32)), key (seed, seed >> 32); 4 words per block for
    # n <= 32, 2 (low, high) for n > 32; masked to n bits
    seed = 0x123456789ABCDEF0
    key = [seed & 0xFFFFFFFF, seed >> 32]
    for i, j, n in [(5, 7, 16), (77, 0, 32), (3, 9, 33), ((1 << 33) + 5, 6, 40), ((1 << 48) - 1, 3, 48)]:
        per = 4 if n <= 32 else 2
        b = j // per
        o = O.philox4x32_10([b & 0xFFFFFFFF, b >> 32, i & 0xFFFFFFFF, 0x43520001 ^ (i >> 32)], key)
        w = o[j % 4] if per == 4 else o[2 * (j % 2)] | (o[2 * (j % 2) + 1] << 32)
        assert O.an_mc_sample(seed, i, j, n) == w & ((1 << n) - 1), (i, j, n)


@pytest.mark.parametrize("n,i0", [(20, 0), (36, 1 << 35), (47, (1 << 46) + 12345)])
def test_an_mc_samples_uniform(n, i0):
    # A = 1: the distance of row i to its samples is popcount(i ^ u); uniform
    # samples make it Binomial(n, 1/2) (4 sigma per bin over 2^14 rows x 256)
    rows, m = 1 << 14, 256
    h = O.an_mc_rows(1, n, m, 99, i0, i0 + rows)
    tot = rows * m
    assert sum(h) == tot
    for d in range(n + 1):
        p = math.comb(n, d) / 2.0 ** n
        assert abs(h[d] - tot * p) <= 4 * math.sqrt(tot * p * (1 - p)) + 1, d
