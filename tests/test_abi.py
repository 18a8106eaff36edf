"""CPU-side checks of the drop-in boundary: libcodrel.so loads (no GPU needed)
and exports every function declared in include/*.h."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "coding_reliability_amd", "libcodrel.so")


def c_abi_functions():
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        if 'extern "C"' not in txt:
            continue
        for m in re.finditer(r"^\s*(?:[A-Za-z_][\w ]*?[\s\*]+)(cr_\w+)\s*\(", txt, re.M):
            names.append(m.group(1))
    return sorted(set(names))


def test_library_exists():
    assert os.path.exists(LIB), "run python -m coding_reliability_amd.build"


def test_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    names = c_abi_functions()
    assert "cr_an_pairs" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_symbol():
    from coding_reliability_amd import _lib
    missing = [n for n in c_abi_functions() if n not in _lib.SIGNATURES]
    assert not missing, missing


def test_status_strings_without_gpu():
    lib = ctypes.CDLL(LIB)
    lib.cr_status_string.restype = ctypes.c_char_p
    assert lib.cr_status_string(0) == b"ok"
    assert lib.cr_status_string(6) == b"admission refused"


def test_null_arguments_are_refused_without_gpu():
    # argument checks come before any device call: status 1, message set
    from coding_reliability_amd import _lib
    L = _lib.lib()
    null = None
    assert L.cr_an_pairs(null, 61, 16, null, 0, null) == 1
    assert L.cr_an_batch(null, null, 0, null, 0, null, null) == 1
    assert L.cr_hamming_weights(null, null, null, 0, null) == 1
    assert L.cr_hamming_weights_batch(null, null, 1, null, 0, null) == 1
    assert L.cr_an_candidates(null, 8, 5, null, null, 0, null, null) == 1
    assert L.cr_server_create(null, null, null) == 1
    assert L.cr_batch_plan_create(null, null, 0, 0, null, null) == 1
    assert L.cr_last_error() == b"NULL argument"
    assert L.cr_hamming_report(8, null, 1, null) == 1
    assert L.cr_probe_ceiling(null, 0, 16, null, null) == 1
    assert L.cr_an_mc(null, 61, 16, 10, 1, null, 65, null) == 1
    assert L.cr_an_mc_ranges(null, 61, 16, null, 1, 1, null, 65, null) == 1


def test_context_without_device_is_an_error():
    from coding_reliability_amd import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    s = L.cr_context_create(0, ctypes.byref(h))
    if s == 0:  # a GPU is present here
        L.cr_context_destroy(h)
        pytest.skip("a device is visible")
    assert s in (3, 5) and L.cr_last_error()


ONE_RUNTIME = r"""
import coding_reliability_amd as cra
from coding_reliability_amd._lib import lib
lib()
import torch
paths = set()
for line in open('/proc/self/maps'):
    p = line.split()[-1]
    if 'libamdhip64' in p or 'libhsa-runtime64' in p:
        paths.add(p)
print('\n'.join(sorted(paths)))
"""


def test_library_then_torch_share_one_hip_runtime():
    # loading libcodrel.so before torch must not bring a second HIP/HSA
    # runtime into the process (the second one to start finds no GPU)
    import subprocess
    import sys
    pytest.importorskip("torch")
    r = subprocess.run([sys.executable, "-c", ONE_RUNTIME], capture_output=True, text=True, cwd=ROOT, timeout=300)
    assert r.returncode == 0, r.stderr
    paths = r.stdout.split()
    assert sum("libamdhip64" in p for p in paths) == 1, paths
    assert sum("libhsa-runtime64" in p for p in paths) == 1, paths


ROCM_FIRST = r"""
import ctypes
ctypes.CDLL('/opt/rocm/lib/libamdhip64.so.7', mode=ctypes.RTLD_GLOBAL)
from coding_reliability_amd._lib import lib
lib()
paths = set()
for line in open('/proc/self/maps'):
    p = line.split()[-1]
    if 'libamdhip64' in p:
        paths.add(p)
print('\n'.join(sorted(paths)))
"""


def test_library_keeps_a_runtime_already_mapped():
    # a process that already holds a HIP runtime (here ROCm's) keeps it; the
    # loader does not add torch's beside it
    import subprocess
    import sys
    if not os.path.exists("/opt/rocm/lib/libamdhip64.so.7"):
        pytest.skip("no ROCm HIP runtime at /opt/rocm/lib")
    r = subprocess.run([sys.executable, "-c", ROCM_FIRST], capture_output=True, text=True, cwd=ROOT, timeout=300)
    assert r.returncode == 0, r.stderr
    paths = r.stdout.split()
    assert len(paths) == 1 and "/opt/rocm" in paths[0], paths


def test_c_header_is_c99():
    # the C ABI header compiles as C99 (no C++ in its declarations)
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("no gcc")
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-fsyntax-only", "-x", "c",
                        os.path.join(ROOT, "include", "coding_reliability.h")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
