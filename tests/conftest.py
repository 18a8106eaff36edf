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


HAMMING_PATTERNS = [0xAB55555556AAAD5B, 0xCD9999999B33366D, 0x78F1E1E1E3C3C78E, 0x01FE01FE03FC07F0,
                    0x01FFFE0003FFF800, 0x01FFFFFFFC000000, 0xFE00000000000000]


def macwilliams_nonext_weights(k, masks):
    """Weight distribution of the non-extended code (data bits plus the parity
    bits parity(d & mask_t)) from the MacWilliams identity: its dual is
    spanned by the rows [mask_t | e_t]."""
    dm = (1 << k) - 1
    masks = [mk & dm for mk in masks if mk & dm]
    N = k + len(masks)
    rows = [mk | (1 << (k + t)) for t, mk in enumerate(masks)]
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


def macwilliams_weights(k, masks=None):
    """Extended-codeword weight distribution A_w of the (N,k) code whose
    parity bits are parity(d & mask_t), from the MacWilliams identity of its
    dual: the dual is spanned by one row per parity mask (mask on the data
    columns, 1 on its parity column) plus the all-ones row (overall parity).
    masks defaults to the implemented patterns truncated to k bits."""
    dm = (1 << k) - 1
    if masks is None:
        masks = [mk & dm for mk in HAMMING_PATTERNS]
    masks = [mk & dm for mk in masks if mk & dm]
    N = k + len(masks) + 1
    rows = [mk | (1 << (k + t)) for t, mk in enumerate(masks)]
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


def hamming_reported_from_ext(k, A, with_1bit):
    """run_hamming's reported form (hamming.cu:183-208) from the extended
    weight distribution A_w of a code of length N = len(A) - 1: bin 0 = 2^k,
    bin 1 = N 2^k with spheres, even bins A_w 2^k, odd bins >= 3 the 1-bit
    sphere counts ((i+1) A_{i+1} + (N-i+1) A_{i-1}) 2^k."""
    N = len(A) - 1
    out = [0] * (N + 1)
    out[0] = 1 << k
    out[1] = N << k if with_1bit else 0
    for i in range(2, N + 1):
        if i % 2 == 0:
            out[i] = A[i] << k
        elif with_1bit and i >= 3:
            out[i] = ((i + 1) * (A[i + 1] if i + 1 <= N else 0) + (N - i + 1) * A[i - 1]) << k
    return out
