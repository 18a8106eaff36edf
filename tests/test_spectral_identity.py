"""The identity the spectral path computes, checked on the CPU against the
oracle's pair histograms (an_coding.cpp:50-102 restated): for the code
C = {A*i : i < 2^k} of n-bit words, D(d) = 2^-n sum_w K_d(w) W(w) with
W(w) = sum_{|S| = w} F(S)^2, F the Walsh-Hadamard transform of C's indicator,
and for odd A one length-2^k transform per high-part mask s (the low k bits of
A*i run over every residue once).  Small codes only: numpy transforms."""
from math import comb

import numpy as np
import pytest

from oracle import oracle as O


def wht(v):
    v = v.copy()
    h = 1
    while h < len(v):
        v = v.reshape(-1, 2, h)
        a, b = v[:, 0, :].copy(), v[:, 1, :].copy()
        v[:, 0, :] = a + b
        v[:, 1, :] = a - b
        v = v.reshape(-1)
        h *= 2
    return v


def krawtchouk(n, d, w):
    return sum((-1) ** j * comb(w, j) * comb(n - w, d - j) for j in range(0, min(d, w) + 1) if d - j <= n - w)


def spectral_raw(A, k):
    """raw[d] = #{i < j : |A i ^ A j| = d} by the chunked spectrum (odd part of A)."""
    Ao = A // (A & -A)
    M = 1 << k
    n = max((Ao * (M - 1)).bit_length(), k)
    Ainv = pow(Ao, -1, M)
    xlo = np.arange(M, dtype=np.int64)
    h = (Ao * ((xlo * Ainv) % M)) >> k  # high part of the codeword with low bits xlo
    pk = np.array([bin(x).count("1") for x in range(M)])
    W = [0] * (n + 1)
    for s in range(1 << (n - k)):
        sign = 1 - 2 * np.array([bin(int(x) & s).count("1") & 1 for x in h], dtype=np.int64)
        F = wht(sign)
        ws = bin(s).count("1")
        for w in range(k + 1):
            W[ws + w] += int((F[pk == w] ** 2).sum())
    raw = [0] * (n + 1)
    for d in range(n + 1):
        t = sum(krawtchouk(n, d, w) * W[w] for w in range(n + 1))
        assert t % (1 << n) == 0  # the exactness check the library applies
        D = t >> n
        if d == 0:
            assert D == M
        else:
            assert D % 2 == 0
            raw[d] = D // 2
    return raw


@pytest.mark.parametrize("A,k", [(61, 6), (61, 8), (3, 7), (641, 5), (58659, 4), (1, 6), (12, 7), (32771, 3)])
def test_identity_equals_pair_count(A, k):
    raw = spectral_raw(A, k)
    ref = O.an_pairs_raw(A, 1 << k, len(raw))
    assert raw == ref


def test_krawtchouk_orthogonality():
    # sum_w C(n, w) K_d(w) K_e(w) = 2^n C(n, d) [d = e]: a transposed or
    # sign-flipped K fails this
    n = 9
    for d in range(n + 1):
        for e in range(n + 1):
            s = sum(comb(n, w) * krawtchouk(n, d, w) * krawtchouk(n, e, w) for w in range(n + 1))
            assert s == ((1 << n) * comb(n, d) if d == e else 0)


@pytest.mark.parametrize("A,k", [(1, 14), (61, 16), (58659, 24), (641, 20), (61, 32), (3, 32), (16777215, 24),
                                 (16777217, 24), (1021, 17)])
def test_high_part_recurrence(A, k):
    # the first pass steps the high part h(x) = (A * (x * A^-1 mod 2^k)) >> k
    # along each 64-aligned block of x as h' = min(t, t - A) in 32-bit
    # arithmetic, t = h + tc, A * A^-1 = 1 + tc * 2^k; checked here against the
    # definition on blocks at the start, middle and end of the range
    M = 1 << k
    Ainv = pow(A, -1, M)
    tc = (A * Ainv - 1) >> k
    assert A * Ainv == 1 + tc * M and tc < A
    n = max((A * (M - 1)).bit_length(), k)
    if n > min(2 * k, 48):
        pytest.skip("shape outside the spectral path")
    for x0 in (0, 64 * 12345 % M, M // 2, M - 64):
        h = (A * ((x0 * Ainv) % M)) >> k
        for j in range(64):
            assert h == (A * (((x0 + j) * Ainv) % M)) >> k, (x0, j)
            t = (h + tc) & 0xFFFFFFFF
            h = min(t, (t - A) & 0xFFFFFFFF)
