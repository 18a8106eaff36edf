"""The allocation-free pair histogram (cr_an_pairs_noalloc, through the C ABI)
against cr_an_pairs_tuned: two calls in a row, the second with a smaller
input, each equal bit for bit; device memory is unchanged across calls."""
import ctypes as C

import pytest

import coding_reliability_amd as cra
from coding_reliability_amd._lib import check, lib

pytestmark = pytest.mark.gpu


def noalloc(A, N, out, stream):
    """One call on `stream` into the device tensor `out`, then a wait."""
    check(lib().cr_an_pairs_noalloc(0, A, N, C.c_void_p(out.data_ptr()), out.numel(), C.c_void_p(stream.cuda_stream)))
    stream.synchronize()
    return [int(x) & ((1 << 64) - 1) for x in out.cpu().tolist()]


def test_noalloc_equals_tuned_two_calls_in_a_row():
    import torch
    ctx = cra.Context(0)
    try:
        stream = torch.cuda.current_stream()
        out = torch.full((80,), -1, dtype=torch.int64, device="cuda")
        # 36-bit codewords (segmented two-dword mode), then a smaller 32-bit code
        for A, N in [(58659, 1 << 20), (61, 1 << 12)]:
            got = noalloc(A, N, out, stream)
            want, _ = ctx.an_pairs(A, N, nbins=80)
            assert got == want, (A, N)
            assert sum(got) == N * (N - 1) // 2
    finally:
        ctx.close()


def test_noalloc_allocates_nothing():
    import torch
    ctx = cra.Context(0)
    try:
        stream = torch.cuda.Stream()
        out = torch.zeros(48, dtype=torch.int64, device="cuda")
        torch.cuda.synchronize()
        noalloc(32771, 1 << 14, out, stream)  # warm-up: the runtime's first launches
        free0, _ = ctx.device_memory()
        a = noalloc(32771, 1 << 16, out, stream)
        b = noalloc(3, 1 << 10, out, stream)
        free1, _ = ctx.device_memory()
        assert free1 == free0
        assert a == ctx.an_pairs(32771, 1 << 16, nbins=48)[0]
        assert b == ctx.an_pairs(3, 1 << 10, nbins=48)[0]
    finally:
        ctx.close()


def test_noalloc_refusals():
    import torch
    out = torch.zeros(16, dtype=torch.int64, device="cuda")
    p = C.c_void_p(out.data_ptr())
    L = lib()
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert L.cr_an_pairs_noalloc(0, 61, 1 << 10, p, 16, s) == 1  # distances reach 16: 17 bins needed
    assert L.cr_an_pairs_noalloc(0, 0, 16, p, 16, s) == 1
    assert L.cr_an_pairs_noalloc(0, 61, 16, None, 16, s) == 1
    assert L.cr_an_pairs_noalloc(99, 61, 16, p, 16, s) == 3  # no such device
