import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long CPU test")


def golden(name):
    import json
    with open(os.path.join(ROOT, "tests", "golden", name)) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def solutions():
    return {k: [int(x) for x in v] for k, v in golden("solutions.json")["vectors"].items()}


def padded_eq(a, b):
    n = max(len(a), len(b))
    return list(a) + [0] * (n - len(a)) == list(b) + [0] * (n - len(b))


import math


def macwilliams_weights(k):
    """Extended-codeword weight distribution A_w of the implemented (N,k) code
    from the MacWilliams identity of its dual: the dual is spanned by one row
    per parity mask (mask on the data columns, 1 on its parity column) plus the
    all-ones row (overall parity)."""
    from oracle import oracle as O
    r = O.hamming_parity_bits(k)
    N = k + r + 1
    masks = [0xAB55555556AAAD5B, 0xCD9999999B33366D, 0x78F1E1E1E3C3C78E, 0x01FE01FE03FC07F0,
             0x01FFFE0003FFF800, 0x01FFFFFFFC000000, 0xFE00000000000000]
    dm = (1 << k) - 1
    rows = []
    t2 = 0
    for mk in masks:
        if mk & dm:
            rows.append((mk & dm) | (1 << (k + t2)))
            t2 += 1
    rows.append((1 << N) - 1)
    B = [0] * (N + 1)
    for sel in range(1 << len(rows)):
        v = 0
        for i, row in enumerate(rows):
            if sel >> i & 1:
                v ^= row
        B[bin(v).count("1")] += 1

    def kraw(w, j):
        return sum((-1) ** s * math.comb(j, s) * math.comb(N - j, w - s) for s in range(0, w + 1))

    out = []
    for w in range(N + 1):
        tot = sum(B[j] * kraw(w, j) for j in range(N + 1))
        assert tot % (1 << len(rows)) == 0
        out.append(tot >> len(rows))
    assert sum(out) == 1 << k
    return out
