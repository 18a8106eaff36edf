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
