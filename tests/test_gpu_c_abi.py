"""The C ABI called from plain C (tests/cpp/c_abi_caller.c, compiled as C99
against include/coding_reliability.h and linked against libcodrel.so): its
histograms equal the oracle's and the reference's golden vector, and a
refused call returns its status and message."""
import os
import subprocess

import pytest

from conftest import ROOT
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CALLER = os.path.join(ROOT, "_build", "tests", "c_abi_caller")


@pytest.mark.skipif(not os.path.exists(CALLER), reason="C caller not built (python -m coding_reliability_amd.build)")
def test_c_caller_matches_oracle(solutions):
    r = subprocess.run([CALLER], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = {}
    for line in r.stdout.splitlines():
        name, *rest = line.split(" ", 1)
        rows[name] = rest[0] if rest else ""
    raw = [int(x) for x in rows["an_pairs"].split()]
    assert raw == O.an_pairs_raw(61, 256, 15)
    # reported form (counts[0] = 2^k, counts[i] = 2 raw[i]) = solutions.h:26
    assert [256] + [2 * x for x in raw[1:]] == solutions["solution_an8_A61"][:15]
    assert [int(x) for x in rows["hamming16"].split()] == O.hamming_weights(16)[:23]
    assert rows["batch_hamming16"] == rows["hamming16"]
    assert [int(x) for x in rows["batch_hamming8_100_200"].split()] == O.hamming_weights(8, 100, 200)[:23]
    m100, m300 = O.an_mc(61, 10, 100, 7), O.an_mc(61, 10, 300, 7)
    assert [int(x) for x in rows["an_mc_0_100"].split()] == m100[:65]
    assert [int(x) for x in rows["an_mc_100_300"].split()] == [b - a for a, b in zip(m100, m300)][:65]
    status, msg = rows["refused"].split(" ", 1)
    assert int(status) == 1 and "must be >= 1" in msg
