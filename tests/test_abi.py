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
