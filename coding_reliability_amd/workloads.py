"""Synthetic workloads of the batched and served configurations.

Every random draw comes from numpy's counter-based Philox generator with a
fixed, stated seed, so a workload is reproducible from its parameters.
"""
from __future__ import annotations

import numpy as np

MIXED_CODES_SEED = 20261202


def mixed_codes(ncodes: int = 4096, seed: int = MIXED_CODES_SEED, kmin: int = 8, kmax: int = 18) -> list[tuple[int, int]]:
    """Batched configuration: odd A uniform in [3, 2^16), k uniform in {kmin..kmax};
    returns (A, 2^k) pairs."""
    g = np.random.Generator(np.random.Philox(seed))
    A = 2 * g.integers(1, 1 << 15, size=ncodes) + 1          # odd, 3 .. 65535
    k = g.integers(kmin, kmax + 1, size=ncodes)
    return [(int(a), 1 << int(kk)) for a, kk in zip(A, k)]


def supera_codes(k: int, abits: int) -> list[tuple[int, int]]:
    """All odd A with exactly `abits` bits (the super-A sweep range
    2^(abits-1) < A < 2^abits of distance_distribution_super_a/src/main.cpp:247-297)."""
    lo = (1 << (abits - 1)) + 1
    return [(A, 1 << k) for A in range(lo, 1 << abits, 2)]


def pairs_of(codes) -> int:
    return sum(N * (N - 1) // 2 for _, N in codes)
