"""Super-A searches with one batched launch: the distance_distribution driver's
output lines equal the reference's recorded sweeps (results/superAs), and the
super_a CSV rows carry the exact per-code histograms."""
import pytest

import coding_reliability_amd as cra
from conftest import golden
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = cra.Context(0)
    yield c
    c.close()


def _fixture(n, h):
    for f in golden("supera_runs.json")["files"]:
        if f["src"].endswith(f"superA_n{n}_h{h}.txt"):
            return f
    return None


CASES = [(8, h) for h in range(3, 17)] + [(16, h) for h in range(3, 15)]


@pytest.mark.parametrize("n,h", CASES)
def test_dd_search_matches_recorded_sweep(ctx, n, h):
    fx = _fixture(n, h)
    if fx is None or fx["summary"] is None:
        pytest.skip("no recorded sweep")
    txt, secs = ctx.supera_search(0, n, h_bits=h)
    lines = txt.strip().split("\n")
    rows = [ln for ln in lines if ": c[" in ln]
    want = [f"{A}: c[{b}] = {c}" for A, b, c in fx["rows"]]
    assert rows == want
    # summary without the time field
    strip = lambda s: [x for i, x in enumerate(s.split(",")) if i not in (13,)]
    got_sum = lines[-1].split(",")
    want_sum = fx["summary"].split(",")
    assert got_sum[:12] == want_sum[:12] and got_sum[14:] == want_sum[14:]


def test_supera_csv_rows(ctx):
    k, lo, hi = 8, 129, 255
    txt, _ = ctx.supera_search(1, k, A_start=lo, A_end=hi)
    lines = txt.strip().split("\n")
    assert lines[0].startswith('"Total Time [s]"')
    header = lines[2].split(", ")
    rows = [ln.split(",") for ln in lines[3:]]
    assert len(rows) == (hi - lo) // 2 + 1
    for r in rows:
        A, h, cnts = int(r[2]), int(r[3]), int(r[7])
        assert h == A.bit_length() and cnts == k + h + 1
        hist = [int(x) for x in r[8 + cnts: 8 + 2 * cnts]]
        raw = O.an_pairs_raw(A, 1 << k, 65)
        assert hist == [2 * raw[i] for i in range(1, cnts + 1)]
    assert header[-1].strip() == "superA2"


def test_supera_candidate_list(ctx):
    As = [58659, 0, 63877, 59665]
    txt, _ = ctx.supera_search(1, 12, As=As)
    rows = [ln.split(",") for ln in txt.strip().split("\n")[3:]]
    assert [int(r[2]) for r in rows] == [58659, 63877, 59665]


@pytest.mark.parametrize("k,abits", [(4, 3), (6, 4), (8, 5), (10, 6), (12, 7), (14, 9), (16, 8), (18, 12), (20, 10)])
def test_an_candidates_vs_oracle(ctx, k, abits):
    best, cands, _ = ctx.an_candidates(k, abits)
    ob, oc = O.an_candidates(k, abits)
    assert (best, cands) == (ob, oc)


def test_an_candidates_tsv_layout(ctx):
    txt = ctx.ancandidates_tsv(6, 8, 4, 5)
    lines = txt.strip().split("\n")
    assert lines[0] == "# bit width data = [6,8]" and len(lines) == 3 + 3 * 2
    n, k, a, best, cnt, lst = lines[3].split("\t")
    assert int(n) == int(k) + int(a) and int(cnt) == len(lst.split(","))


import os
import subprocess

from conftest import ROOT

BUILD = os.path.join(ROOT, "_build")


@pytest.mark.skipif(not os.path.exists(os.path.join(BUILD, "ANcandidates")), reason="CLI not built")
def test_ancandidates_cli(ctx):
    out = subprocess.run([os.path.join(BUILD, "ANcandidates"), "-d", "6", "-D", "9", "-a", "4", "-A", "6"],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.strip().split("\n")
    assert lines[2] == "# n=|C|\tk=|D|\t|A|\t|SDR|\t|candidates|\tcandidates"
    for ln in lines[3:]:
        n, k, a, best, cnt, lst = ln.split("\t")
        ob, oc = O.an_candidates(int(k), int(a))
        assert int(best) == ob and [int(x) for x in lst.split(",")] == oc and int(cnt) == len(oc)


@pytest.mark.skipif(not os.path.exists(os.path.join(BUILD, "codrel_supera")), reason="CLI not built")
def test_supera_cli(ctx, tmp_path):
    r = subprocess.run([os.path.join(BUILD, "codrel_supera"), "-k", "10", "-s", "33", "-S", "63", "-f", "t"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    body = (tmp_path / "t_k10_a33-63.csv").read_text().strip().split("\n")
    ref, _ = ctx.supera_search(1, 10, A_start=33, A_end=63)
    # identical apart from the time fields
    strip_t = lambda rows: [",".join(x.split(",")[:-11]) for x in rows[3:]]
    assert strip_t(body) == strip_t(ref.strip().split("\n"))
    assert len(body) == 3 + 16


@pytest.mark.skipif(not (os.path.exists(os.path.join(BUILD, "codrel_supera"))
                         and os.path.exists(os.path.join(BUILD, "ANcandidates"))), reason="CLI not built")
def test_supera_cli_reads_ancandidates_tsv(ctx, tmp_path):
    tsv = subprocess.run([os.path.join(BUILD, "ANcandidates"), "-d", "8", "-D", "8", "-a", "6", "-A", "7"],
                         capture_output=True, text=True, timeout=120)
    assert tsv.returncode == 0, tsv.stderr
    (tmp_path / "cands.tsv").write_text(tsv.stdout)
    want = []
    for ln in tsv.stdout.strip().split("\n"):
        if not ln.startswith("#"):
            want += [int(x) for x in ln.split("\t")[-1].split(",")]
    r = subprocess.run([os.path.join(BUILD, "codrel_supera"), "-k", "8", "-c", "cands.tsv", "-f", "t"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    body = (tmp_path / "t_k8_c_cands.tsv.csv").read_text().strip().split("\n")
    assert [int(x.split(",")[2]) for x in body[3:]] == want
