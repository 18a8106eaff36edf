"""SDC tables of the Extended Hamming codes in the legacy table layout
(legacy/hamming_count_undetectable_errors.cpp:226-339).

Pins: the weight tables the legacy program itself seeds its (40,33) and
(52,45) runs with (tests/golden/legacy_hamming_weights.json), the oracle's
exhaustive enumeration for k <= 20, the extended-code MacWilliams enumerator
of the test helpers (dual with the all-ones row) for k up to 64, and the
correcting column against run_hamming's reported counts (hamming.cu:183-208),
a second reference source for the same sphere formula."""
import math

import pytest

import coding_reliability_amd as cra
from conftest import golden, macwilliams_weights
from oracle import oracle as O


def _rows(txt):
    body = [ln for ln in txt.split("\n") if ln and not ln.startswith("#")]
    return [ln.split(",") for ln in body]


@pytest.mark.parametrize("t", golden("legacy_hamming_weights.json")["tables"], ids=lambda t: f"k{t['k']}")
def test_macwilliams_equals_legacy_seed_tables(t):
    txt, w = cra.hamming_sdc_table(t["k"], 2)
    assert w == t["counts"]
    assert [int(r[1]) for r in _rows(txt)] == t["counts"]


@pytest.mark.parametrize("k", list(range(1, 21)))
def test_macwilliams_equals_enumeration(k):
    _, w = cra.hamming_sdc_table(k, 2)
    assert w == O.hamming_weights(k)[: len(w)]


@pytest.mark.parametrize("k", [8, 16, 24, 31, 32, 33, 40, 47, 57, 58, 63, 64])
def test_extended_column_equals_extended_enumerator(k):
    txt, _ = cra.hamming_sdc_table(k, 2)
    assert [int(r[2]) for r in _rows(txt)] == macwilliams_weights(k)


@pytest.mark.parametrize("k", [4, 8, 16, 24, 32, 33, 64])
def test_table_columns(k):
    txt, w = cra.hamming_sdc_table(k, 2)
    N = cra.hamming_N(k)
    rows = _rows(txt)
    assert len(rows) == N + 1 and all(len(r) == 9 for r in rows)
    assert [int(r[0]) for r in rows] == list(range(N + 1))
    assert sum(int(r[1]) for r in rows) == 1 << k
    ext = [int(r[2]) for r in rows]
    cor = [int(r[3]) for r in rows]
    for i, r in enumerate(rows):
        assert int(r[4]) == ext[i] << k and int(r[5]) == cor[i] << k
        assert int(r[6]) == math.comb(N, i) << k
        assert r[7] == f"{(ext[i] << k) / (math.comb(N, i) << k):.10f}"
        assert r[8] == f"{(cor[i] << k) / (math.comb(N, i) << k):.10f}"
    # the correcting column (legacy) is run_hamming's reported form (main driver)
    if max(cor) << k < 1 << 128:
        rep = cra.hamming_report(k, w, True)
        assert [c << k for c in cor[3:]] == rep[3:] and (cor[0] << k, cor[1] << k) == (rep[0], rep[1])
    else:  # the reported form is uint128 like the reference's; the table keeps 256 bits
        with pytest.raises(cra.CRError):
            cra.hamming_report(k, w, True)
    head = txt.split("\n")
    assert head[0] == f"# SDC probabilities for a (n,k) = ({N},{k}) " + \
        ("shortened " if N & (N - 1) == 0 else "") + "correcting Extended Hamming code"
    assert head[11] == "#Distances:"
    tail = [ln for ln in head if ln.startswith(f"# ({N}/{k})")][0]
    assert tail.endswith(f"for {1 << k} code words and {sum(int(r[5]) for r in rows)} transitions.")
    assert txt.endswith(".\n\n\n")


def test_column_widths_right_aligned():
    txt, _ = cra.hamming_sdc_table(33, 2)
    rows = [ln.split(",") for ln in txt.split("\n") if ln and not ln.startswith("#")]
    for col in range(7):
        widths = {len(r[col]) for r in rows}
        assert len(widths) == 1, col  # every field of a column padded to the column's width
    assert rows[0][0] == "   0"


def test_invalid_arguments():
    with pytest.raises(cra.CRError):
        cra.hamming_sdc_table(0, 2)
    with pytest.raises(cra.CRError):
        cra.hamming_sdc_table(65, 2)
    with pytest.raises(cra.CRError):
        cra.hamming_sdc_table(41, 1)  # device enumeration needs k <= 40 (and a context)


@pytest.mark.gpu
@pytest.mark.parametrize("k", [8, 16, 26, 27, 32, 33, 36, 40])
def test_device_enumeration_equals_macwilliams(k):
    ctx = cra.Context(0)
    try:
        t1, w1 = ctx.hamming_sdc_table(k, 1)
    finally:
        ctx.close()
    t2, w2 = cra.hamming_sdc_table(k, 2)
    assert w1 == w2
    strip = lambda t: [ln for ln in t.split("\n") if "Computation took" not in ln]
    assert strip(t1) == strip(t2)


@pytest.mark.gpu
def test_device_enumeration_equals_legacy_40_33():
    t = [x for x in golden("legacy_hamming_weights.json")["tables"] if x["k"] == 33][0]
    ctx = cra.Context(0)
    try:
        _, w = ctx.hamming_sdc_table(33, 0)
    finally:
        ctx.close()
    assert w == t["counts"]


import os
import subprocess

from conftest import ROOT

CLI = os.path.join(ROOT, "_build", "hamming_tables")


def _no_time(t):
    return [ln for ln in t.split("\n") if "Computation took" not in ln]


@pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")
def test_cli_macwilliams_range():
    out = subprocess.run([CLI, "-k", "63", "-K", "64", "-m", "2"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    want = cra.hamming_sdc_table(63, 2)[0] + cra.hamming_sdc_table(64, 2)[0]
    assert _no_time(out.stdout) == _no_time(want)


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")
def test_cli_device_range():
    out = subprocess.run([CLI, "-k", "31", "-K", "33"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    want = "".join(cra.hamming_sdc_table(k, 2)[0] for k in (31, 32, 33))
    assert _no_time(out.stdout) == _no_time(want)
