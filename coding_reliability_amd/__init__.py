"""MI355X engine for the distance-distribution path of brics-db/coding_reliability.

Python host API over the C ABI of libcodrel.so (hand-written gfx950 HIP
kernels).  The reference-compatible C++ entry points (run_ancoding, ...) are
exported by the same library; see include/cr_reference_api.h.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import CRError, check, cr_device_info, cr_pair_tuning, lib

__all__ = ["Context", "CRError", "an_bins", "ceil_log2"]


def ceil_log2(A: int) -> int:
    """h = ceil(log(A)/log(2.0)) in double, as the reference drivers compute it
    (distance_distribution/src/an_coding.cu:185)."""
    import math
    return int(math.ceil(math.log(A) / math.log(2.0))) if A > 1 else 0


def an_bins(A: int, num_words: int) -> int:
    """Histogram bins needed for codewords A*i, i < num_words."""
    return max(((num_words - 1) * A).bit_length(), 0) + 1


class Context:
    """One device, one HIP stream (cr_context)."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        check(lib().cr_context_create(device, C.byref(self._h)))

    def close(self):
        if self._h:
            lib().cr_context_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def device_info(self) -> dict:
        info = cr_device_info()
        check(lib().cr_device_query(self._h, C.byref(info)))
        return dict(name=info.name.decode(), arch=info.arch.decode(), compute_units=info.compute_units,
                    max_clock_mhz=info.max_clock_mhz, hbm_bytes=info.hbm_bytes)

    def an_pairs(self, A: int, num_words: int, nbins: int | None = None, rows_per_lane: int = 0,
                 chunk_blocks: int = 0, waves: int = 0) -> tuple[list[int], float]:
        """Raw unordered-pair histogram of codewords A*i, i < num_words, and the
        pair-kernel time in ms."""
        if nbins is None:
            nbins = an_bins(A, num_words)
        out = np.zeros(nbins, dtype=np.uint64)
        ms = C.c_double(0)
        t = cr_pair_tuning(rows_per_lane, chunk_blocks, waves)
        check(lib().cr_an_pairs_tuned(self._h, A, num_words, out.ctypes.data_as(C.POINTER(C.c_uint64)), nbins,
                                      C.byref(t), C.byref(ms)))
        return [int(x) for x in out], ms.value
