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
