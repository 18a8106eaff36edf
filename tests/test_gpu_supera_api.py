"""The super_a C++ entry points (distance_distribution_super_a/src/algorithms.h:20-21,
globals.h:58-62) called from a C++ program compiled against
include/cr_reference_api.h (tests/cpp/supera_caller.cpp), checked against the
oracle, the reference's recorded sweeps and the quirks of
super_a/src/an_coding.cu:135-244:
  - counts[0] = 2^k and counts[b] = raw[b] << 1 in 64 bits before widening
    (an_coding.cu:226-231; a bin holds at most C(2^32, 2) < 2^63 pairs for
    k <= 32, so the 64-bit doubling cannot wrap where the library runs);
  - times = {kernel, total} in seconds (an_coding.cu:233-237), the reverse of
    distance_distribution's run_ancoding;
  - Flags::dev selects the device; the default -1 runs on the first device."""
import os
import subprocess

import pytest

from conftest import ROOT, golden
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CALLER = os.path.join(ROOT, "_build", "tests", "supera_caller")


def call(*args):
    assert os.path.exists(CALLER), "tests/cpp/supera_caller.cpp not built (coding_reliability_amd.build)"
    r = subprocess.run([CALLER, *[str(a) for a in args]], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    rows, errors, csv = {}, {}, []
    lines = r.stdout.split("\n")
    i = 0
    while i < len(lines) and lines[i] != "CSV":
        f = lines[i].split(" ", 2)
        if f[0] == "R":
            v = lines[i].split()
            rows[int(v[1])] = dict(h=int(v[2]), t_kernel=float(v[3]), t_total=float(v[4]),
                                   counts=[int(x) for x in v[5:]])
        elif f[0] == "E":
            errors[int(f[1])] = f[2]
        i += 1
    csv = [ln for ln in lines[i + 1:] if ln.strip()]
    return rows, errors, csv


def reported(A, k, h):
    raw = O.an_pairs_raw(A if A & (A - 1) else 1, 1 << k, 65)
    return [1 << k] + [2 * raw[b] for b in range(1, k + h + 1)]


def lowest(counts, cc):
    # get_lowest_prob (super_a/src/main.cpp:157-170)
    for i in range(1, (cc + 1) // 2):
        if counts[i]:
            return i, counts[i]
    return 0xFFFF, (1 << 128) - 1


def test_exhaustive_k8_sweep_against_oracle_and_recorded_search():
    k, hb = 8, 6
    As = list(range((1 << (hb - 1)) + 1, 1 << hb, 2))
    rows, errors, csv = call("exhaustive", k, -1, *As)
    assert not errors
    best = (0, (1 << 128) - 1)
    lines = []
    for A in As:
        r = rows[A]
        assert r["h"] == A.bit_length()
        assert r["counts"] == reported(A, k, r["h"]), A
        # times = {kernel, total}
        assert 0 < r["t_kernel"] <= r["t_total"], r
        b, c = lowest(r["counts"], k + r["h"] + 1)
        if best[0] < b or (best[0] == b and best[1] > c):
            best = (b, c)
            lines.append((A, b, c))
    # the new-best lines equal distance_distribution's recorded sweep of the
    # same A range (results/superAs/superA_n8_h6.txt)
    fx = [f for f in golden("supera_runs.json")["files"] if f["src"].endswith("superA_n8_h6.txt")][0]
    assert [tuple(x) for x in fx["rows"]] == lines
    # process_result rows: header, then one row per A with the hist columns
    # hist_1 .. hist_{k+h+1} (super_a/src/globals.cpp:75-89 reads one bin past
    # the histogram, the zero of the caller's COUNTS_MAX_WIDTH array)
    assert csv[0].startswith("id, k, A, h, A_2, m_1, m_2, hlen")
    for line, A in zip(csv[1:], As):
        v = line.split(",")
        cnts = int(v[7])
        assert int(v[2]) == A and cnts == k + A.bit_length() + 1
        assert [int(x) for x in v[8 + cnts: 8 + 2 * cnts]] == rows[A]["counts"][1:] + [0]


def test_exhaustive_k16_recorded_minimum():
    # superA_n16_h16.txt:33: A = 58659 at k = 16 has no distance below 6 and
    # c[6] = 733232
    rows, errors, _ = call("exhaustive", 16, 0, 58659, 61)
    assert not errors
    c = rows[58659]["counts"]
    assert c[1:6] == [0] * 5 and c[6] == 733232
    assert c == reported(58659, 16, 16)
    assert rows[61]["counts"] == reported(61, 16, 6)


def test_power_of_two_and_unit_multiplier():
    # A power of two runs as A = 1 (an_coding.cu:137-138); h stays as passed
    rows, errors, _ = call("exhaustive", 10, -1, 1, 64)
    assert not errors
    want = reported(1, 10, 1)
    assert rows[1]["counts"] == want
    assert rows[64]["counts"][: len(want)] == want and not any(rows[64]["counts"][len(want):])


def test_device_selection():
    rows, errors, _ = call("exhaustive", 8, 0, 61)
    assert 61 in rows and not errors
    # a device index the box does not have is an error (the reference's
    # cudaSetDevice check exits), reported as an exception
    rows, errors, _ = call("exhaustive", 8, 63, 61)
    assert 61 in errors and not rows


def test_grid1d_against_oracle_and_recorded_file():
    k, m = 16, 1001
    rows, errors, csv = call("grid", 1, k, m, 0, 61)
    assert not errors
    _, rep, _ = O.an_grid_ref(1, k, m, 0, 61)
    got = rows[61]["counts"]
    assert got == rep[: len(got)]
    rec = [r for r in golden("an_grid1d_runs.json")["runs"]
           if r["src"].endswith("converge/geom/ancoding_grid1d_gpu_m1001_A61_n16_h6.csv")][0]
    assert got == rec["counts"][: len(got)]
    v = csv[1].split(",")
    assert int(v[5]) == m and int(v[6]) == 0


def test_grid2d_against_oracle():
    k, m1, m2 = 16, 100, 50
    rows, errors, _ = call("grid", 2, k, m1, m2, 61, 58659)
    assert not errors
    for A in (61, 58659):
        _, rep, _ = O.an_grid_ref(2, k, m1, m2, A)
        got = rows[A]["counts"]
        assert got == rep[: len(got)], A
    # the driver requires A < 2^k (super_a/src/an_coding_grid.cu:182 asserts it)
    rows, errors, _ = call("grid", 2, 12, m1, m2, 58659)
    assert 58659 in errors and not rows
