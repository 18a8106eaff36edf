"""The reference's own driver (distance_distribution/src/main.cpp, unchanged)
linked against libcodrel.so.  The CPU test builds and links it where the
reference is mounted; the GPU tests run the linked binary and compare its CSV
output with the reference's recorded result files."""
import os
import subprocess

import pytest

from conftest import ROOT, golden

CLI = os.path.join(ROOT, "_build", "codrel_ref")
REF_MAIN = "/root/reference/distance_distribution/src/main.cpp"


@pytest.mark.skipif(not os.path.exists(REF_MAIN), reason="reference sources not mounted")
def test_reference_main_links_against_library():
    from coding_reliability_amd import build as B
    path = B.build_reference_cli()
    out = subprocess.run([path, "-h"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "USAGE" in out.stdout


def read_counts(path):
    counts = []
    for line in open(path):
        if not line.strip():
            break
        counts.append(int(line.split(",")[1]))
    return counts


def run_cli(tmp_path, *args):
    r = subprocess.run([CLI, *args], cwd=tmp_path, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr
    return r.stdout


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(CLI), reason="relinked reference CLI not built")
def test_cli_exhaustive_matches_recorded_run(tmp_path):
    run_cli(tmp_path, "-a", "61", "-n", "16", "-f")
    got = read_counts(tmp_path / "ancoding_gpu_A61_n16_h6.csv")
    rec = [r for r in golden("an_exhaustive_runs.json")["runs"]
           if r["src"].endswith("runtimes/4/ancoding_gpu_A61_n16_h6.csv")][0]
    assert got == rec["counts"]


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(CLI), reason="relinked reference CLI not built")
def test_cli_exhaustive_k32_is_exact(tmp_path):
    # the reference driver's exhaustive mode at k = 32 (2^63 pairs: out of
    # reach for pair enumeration) now returns the exact distribution, through
    # cr_an_pairs's spectral path
    run_cli(tmp_path, "-a", "61", "-n", "32", "-f")
    got = read_counts(tmp_path / "ancoding_gpu_A61_n32_h6.csv")
    import coding_reliability_amd as cra
    ctx = cra.Context(0)
    raw, _ = ctx.an_pairs_spectral(61, 32)
    ctx.close()
    want = [1 << 32] + [2 * x for x in raw[1:]]
    n = max(len(got), len(want))
    assert got + [0] * (n - len(got)) == want + [0] * (n - len(want))
    assert sum(got) == 1 << 64


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(CLI), reason="relinked reference CLI not built")
def test_cli_grid_matches_recorded_run(tmp_path):
    run_cli(tmp_path, "-a", "61", "-n", "16", "-m", "1001", "-g", "1", "-f")
    got = read_counts(tmp_path / "ancoding_grid1d_gpu_m1001_A61_n16_h6.csv")
    rec = [r for r in golden("an_grid1d_runs.json")["runs"]
           if r["src"].endswith("converge/geom/ancoding_grid1d_gpu_m1001_A61_n16_h6.csv")][0]
    assert got == rec["counts"]


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(CLI), reason="relinked reference CLI not built")
def test_cli_hamming_matches_golden(tmp_path, solutions):
    run_cli(tmp_path, "-e", "1", "-n", "24", "-b", "-f")
    got = read_counts(tmp_path / "hamming_gpu_n24_h6.csv")
    sol = solutions["sol_hamming_n24"]
    assert all(got[i] == sol[i] << 24 for i in range(4, len(sol), 2))


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(CLI), reason="relinked reference CLI not built")
def test_cli_hamming_grid_matches_oracle(tmp_path):
    # run_hamming_grid (hamming_grid.cu:91-219): k=32, m=2^20 sample points,
    # even bins = 2^32 * (w[i] + w[i-1]) over the reference's sample set
    from oracle import oracle as O
    k, m = 32, 1 << 20
    run_cli(tmp_path, "-e", "1", "-n", str(k), "-m", str(m), "-g", "1", "-f")
    got = read_counts(tmp_path / f"hamming_mc_m{m}_n{k}_h7.csv")
    words = O.hamming_grid_ref_samples(m, float(1 << k) / m, 256)
    w = O.hamming_weights_list(k, words)
    for i in range(2, 40, 2):
        assert got[i] == (w[i] + w[i - 1]) << k, i


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(CLI), reason="relinked reference CLI not built")
def test_cli_an_mc_close_to_golden(tmp_path, solutions):
    # run_ancoding_mc with Philox samples: scaled counts sum to ~4^k and stay
    # close to the golden A=61, k=16 vector
    out = run_cli(tmp_path, "-a", "61", "-n", "16", "-m", "4096", "-f")
    got = read_counts(tmp_path / "ancoding_mc_gpu_m4096_A61_n16_h6.csv")
    sol = solutions["solution_an16_A61"]
    assert abs(sum(got) - 4 ** 16) < 4 ** 16 * 1e-6 + 64
    for i in range(5, 20):
        assert abs(got[i] - sol[i]) <= 0.02 * sol[i], i


def _text(path):
    lines = open(path).read().split("\n")
    rows = []
    for ln in lines:
        if not ln.strip():
            break
        rows.append(ln)
    summary = [ln for ln in lines if ln.startswith(('"Total"', '"MaxAbsErr"', '"MaxRelErr"'))]
    return rows, summary


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(CLI), reason="relinked reference CLI not built")
@pytest.mark.parametrize("name,args,fname", [
    ("exhaustive_A61_n16", ["-a", "61", "-n", "16", "-f"], "ancoding_gpu_A61_n16_h6.csv"),
    ("grid1d_m1001_A61_n16", ["-a", "61", "-n", "16", "-m", "1001", "-g", "1", "-f"],
     "ancoding_grid1d_gpu_m1001_A61_n16_h6.csv"),
])
def test_cli_csv_text_matches_recorded_file(tmp_path, name, args, fname):
    # every column of every row, and the total / error summary lines, as text
    rec = golden("recorded_csv_text.json")[name]
    run_cli(tmp_path, *args)
    rows, summary = _text(tmp_path / fname)
    assert rows == rec["rows"], rec["src"]
    assert summary == rec["summary"], rec["src"]


def _as_cases():
    import re
    cases = []
    for name, rec in sorted(golden("recorded_csv_text.json").items()):
        m = re.match(r"As_A(\d+)_n(\d+)$", name)
        if m:
            cases.append((name, m.group(1), m.group(2), os.path.basename(rec["src"])))
    return cases


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(CLI), reason="relinked reference CLI not built")
@pytest.mark.parametrize("name,A,n,fname", _as_cases())
def test_cli_csv_text_matches_recorded_as_files(tmp_path, name, A, n, fname):
    rec = golden("recorded_csv_text.json")[name]
    run_cli(tmp_path, "-a", A, "-n", n, "-f")
    rows, summary = _text(tmp_path / fname)
    assert rows == rec["rows"], rec["src"]
    # Some of these files were written before the driver printed a "Total"
    # line, and the rest by a driver whose total left out the last bin
    # (distance k+h; non-zero for A = 1 and A = 3, for instance).  The current
    # globals.cpp:185-193 sums every bin, as the library does.
    counts = [int(r.split(",")[1]) for r in rows]
    assert summary == [f'"Total",{float(sum(counts)):.6g}']
    if rec["summary"]:
        assert rec["summary"][0] in (f'"Total",{float(sum(counts)):.6g}', f'"Total",{float(sum(counts[:-1])):.6g}')


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(CLI), reason="relinked reference CLI not built")
def test_cli_grid_k40_csv_text_matches_recorded_file(tmp_path):
    # 2^40 words x 101 reference samples (1.2e14 pairs): counts above 2^64
    # printed with 12 digits; the recorded error columns are uninitialised
    # memory, so the first four columns of each row and the total are compared
    rec = golden("recorded_csv_text.json")["grid1d_m101_A61_n40"]
    run_cli(tmp_path, "-a", "61", "-n", "40", "-m", "101", "-g", "1", "-f")
    rows, summary = _text(tmp_path / "ancoding_grid1d_gpu_m101_A61_n40_h6.csv")
    assert [",".join(r.split(",")[:4]) for r in rows] == rec["rows"], rec["src"]
    assert [x for x in summary if x.startswith('"Total"')] == rec["summary"], rec["src"]


# --- the reference CPU implementations' entry points (served by the GPU) ------

@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(CLI), reason="relinked reference CLI not built")
@pytest.mark.parametrize("n", [8, 16, 24])
def test_cli_cpu_path_matches_recorded_file(tmp_path, n):
    # -c: run_ancoding_cpu (an_coding.cpp:104-142) with its own conventions:
    # the ancoding_cpu prefix, a one-row timing table, no device line; every
    # row (counts, probabilities, binomial bases) and the total as recorded
    rec = golden("recorded_csv_text.json")[f"cpu_A61_n{n}"]
    run_cli(tmp_path, "-c", "-a", "61", "-n", str(n), "-f")
    path = tmp_path / f"ancoding_cpu_A61_n{n}_h6.csv"
    rows, summary = _text(path)
    assert rows == rec["rows"], rec["src"]
    assert summary == rec["summary"], rec["src"]
    txt = open(path).read()
    assert '0,"Total Runtime",' in txt and "Kernel Runtime" not in txt and '"Device"' not in txt


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(CLI), reason="relinked reference CLI not built")
def test_cli_cpu_path_keeps_power_of_two_multiplier(tmp_path):
    # the CPU path has no power-of-two -> 1 mapping: A = 8 keeps h = 3 and
    # n + 4 bins, with the distances of A = 1
    from oracle import oracle as O
    n = 10
    run_cli(tmp_path, "-c", "-a", "8", "-n", str(n), "-f")
    got = read_counts(tmp_path / f"ancoding_cpu_A8_n{n}_h3.csv")
    want, _ = O.an_cpu_counts(n, 8)
    assert len(got) == n + 4 and got == want


def _hamming_cli(tmp_path, e, n, with_1bit, prefix, extra=()):
    args = ["-e", str(e), "-n", str(n), "-f", *extra] + (["-b"] if with_1bit else [])
    out = run_cli(tmp_path, *args)
    h = 5 if n == 8 else (6 if n < 32 else 7)
    return read_counts(tmp_path / f"{prefix}_n{n}_h{h}.csv"), out


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(CLI), reason="relinked reference CLI not built")
@pytest.mark.parametrize("with_1bit", [0, 1])
def test_cli_hamming_bruteforce_entry_points(tmp_path, with_1bit):
    # -e 2 / -e 4: run_hamming_cpu_native_short / run_hamming_cpu_native
    # (hamming_native_short.cpp:29-84, hamming_native.cpp:27-127) enumerate
    # every ordered pair of codewords and their 1-bit spheres; the oracle's
    # explicit enumeration of the same spheres; -e 3 counts pairs a < b twice
    # (hamming_native_short_ordering.cpp:58-80), so its bins 0 and 1 are 0
    from oracle import oracle as O
    for n in (8, 16):
        brute = O.hamming_bruteforce(n, bool(with_1bit)) if n == 8 else None
        ref = O.hamming_report(n, O.hamming_weights(n), bool(with_1bit))
        if brute is not None:
            assert ref == brute
        got, _ = _hamming_cli(tmp_path, 2, n, with_1bit, "hamming_cpu_native_short")
        assert got == ref, n
        got, out = _hamming_cli(tmp_path, 4, n, with_1bit, "hamming_cpu_native")
        assert got == ref, n
        assert ("#Distances:" in out) == bool(with_1bit)
        got, _ = _hamming_cli(tmp_path, 3, n, with_1bit, "hamming_cpu_native_short_ordering")
        assert got == [0, 0] + ref[2:], n


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(CLI), reason="relinked reference CLI not built")
@pytest.mark.parametrize("n", [8, 16, 24, 32])
def test_cli_hamming_cpu_entry_point(tmp_path, n, solutions):
    # -e 1 -c: run_hamming_cpu (hamming.cpp:94-126): even bins are the
    # solutions.h weight distributions << n, odd bins the sphere closed form
    from conftest import hamming_reported_from_ext, macwilliams_weights
    out = run_cli(tmp_path, "-e", "1", "-c", "-n", str(n), "-b", "-f")
    assert "Start Hamming Coding Algorithm (CPU)" in out
    want = hamming_reported_from_ext(n, macwilliams_weights(n), True)
    sol = solutions[f"sol_hamming_n{n}"]
    assert all(want[i] == sol[i] << n for i in range(4, len(sol), 2))
    h = 5 if n == 8 else (6 if n < 32 else 7)
    _check_wide_counts(tmp_path / f"hamming_cpu_n{n}_h{h}.csv", want)


def _check_wide_counts(path, want):
    # counts above 2^64 are printed with 12 significant digits when n >= 32
    # (globals.cpp:204-207), every other count exactly
    rows = []
    for line in open(path):
        if not line.strip():
            break
        rows.append(line.split(",")[1].strip())
    assert len(rows) == len(want)
    for i, (r, w) in enumerate(zip(rows, want)):
        if w > 1 << 64:
            assert abs(float(r) / w - 1) < 1e-11, i
        else:
            assert int(r) == w, i


def hamming_cpu_reported(n, wnon, with_1bit):
    """hamming_cpu<N>'s post-processing (hamming.cpp:59-92) of a non-extended
    weight distribution: bins 0 and 1 set, bin 2 the weight-2 count as is,
    odd weights merged into the next even bin from bin 4, the 1-bit spheres,
    bins >= 3 shifted by n."""
    h = 5 if n == 8 else (6 if n < 32 else 7)
    cc = n + h + 1
    c = [0] * cc
    for d in range(2, cc):
        c[d] = wnon[d] if d < len(wnon) else 0
    c[0] = 1 << n
    c[1] = (n + h) * c[0] if with_1bit else 0
    for d in range(4, cc, 2):
        c[d] += c[d - 1]
        c[d - 1] = 0
    if with_1bit:
        for d in range(3, cc, 2):
            c[d] = (n + h - d + 1) * c[d - 1] + ((d + 1) * c[d + 1] if d < n + h else 0)
    return c[:3] + [x << n for x in c[3:]]


def test_cpu_post_processing_equals_extended_distribution():
    # for codes without words of weight 1 or 2 the CPU driver's post-processing
    # of the non-extended weights is the reported form of the extended code's
    # distribution (both MacWilliams, independent duals)
    from conftest import HAMMING_PATTERNS, hamming_reported_from_ext, macwilliams_nonext_weights, macwilliams_weights
    for n in (8, 16, 24, 32):
        wnon = macwilliams_nonext_weights(n, HAMMING_PATTERNS)
        for b in (0, 1):
            assert hamming_cpu_reported(n, wnon, b) == hamming_reported_from_ext(n, macwilliams_weights(n), b)


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(CLI), reason="relinked reference CLI not built")
def test_cli_hamming_cpu_entry_point_n40(tmp_path):
    # run_hamming_cpu for n = 40: apply_w<40> (hamming.cpp:25-41) takes the
    # six 32-bit parity masks, so the 8 high data bits are unprotected and the
    # code has words of weight 1 and 2, where the CPU driver's post-processing
    # differs from the GPU driver's; the non-extended weights come from the
    # MacWilliams identity of the code's dual; counts above 2^64 are printed
    # with 12 significant digits (globals.cpp:204-207)
    from conftest import macwilliams_nonext_weights
    n = 40
    masks32 = [0x56AAAD5B, 0x9B33366D, 0xE3C3C78E, 0x03FC07F0, 0x03FFF800, 0xFC000000]
    wnon = macwilliams_nonext_weights(n, masks32)
    assert wnon[1] == 8 and wnon[2] == 28
    for b in (0, 1):
        want = hamming_cpu_reported(n, wnon, b)
        run_cli(tmp_path, "-e", "1", "-c", "-n", str(n), "-f", *(["-b"] if b else []))
        _check_wide_counts(tmp_path / f"hamming_cpu_n{n}_h7.csv", want)


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(CLI), reason="relinked reference CLI not built")
def test_cli_an_mc_beyond_32_bits(tmp_path):
    # run_ancoding_mc at n = 40 (an_coding_mc.cu:163-188 dispatches up to
    # ccmax 64): 2^40 words x 2 Philox samples of 40 bits; the reported counts
    # are (2^40 / 2) x the raw histogram, summing to exactly 4^40
    run_cli(tmp_path, "-a", "61", "-n", "40", "-m", "2", "-f")
    rows, summary = _text(tmp_path / "ancoding_mc_gpu_m2_A61_n40_h6.csv")
    assert len(rows) == 47
    # the error columns leave the stream at 7 significant digits
    assert summary[0] == '"Total",' + f"{float(4 ** 40):.7g}"
    counts = [float(r.split(",")[1]) for r in rows]
    assert counts[1] == 0  # no two multiples of 61 differ in one bit
    assert abs(sum(counts) / 4 ** 40 - 1) < 1e-9


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(CLI), reason="relinked reference CLI not built")
def test_cli_mc_search_steps_equal_separate_runs(tmp_path, solutions):
    # ancoding_mc_search (main.cpp:179-202): run_ancoding_mc at m, 3m, 9m, ...
    # until the max relative error (offset 0, an_coding_mc.cu's return value)
    # is within the bound.  The drop-in draws each sample once across the
    # steps; every step's error must equal the one a fresh process computes
    # from its own run_ancoding_mc counts at that iteration count.
    A, n, m0, bound = 61, 16, 1000, 0.05
    run_cli(tmp_path, "-a", str(A), "-n", str(n), "-m", str(m0), "-s", str(bound))
    rows = [ln.split(",") for ln in open(tmp_path / "ancoding_mc_search.csv").read().split("\n") if ln.strip()]
    assert len(rows) >= 2
    sol = solutions["solution_an16_A61"]
    for step, r in enumerate(rows):
        assert int(r[0]) == step and int(r[1]) == n and int(r[3]) == m0 * 3 ** step
        iters = int(r[3])
        sub = tmp_path / f"step{step}"
        sub.mkdir()
        run_cli(sub, "-a", str(A), "-n", str(n), "-m", str(iters), "-f")
        # run_ancoding_mc caps the iterations at 2^n (an_coding_mc.cu) and
        # names the file by the capped count
        counts = read_counts(sub / f"ancoding_mc_gpu_m{min(iters, 1 << n)}_A{A}_n{n}_h6.csv")
        err = 0.0
        for k in range(min(n + 6, len(sol))):
            e = abs(counts[k] / sol[k] - 1.0) if sol[k] > 0 else (1.0 if counts[k] > 0 else 0.0)
            err = max(err, e)
        assert abs(float(r[4]) - err) <= 1e-6, (step, r[4], err)
        assert (err <= bound) == (step == len(rows) - 1)
