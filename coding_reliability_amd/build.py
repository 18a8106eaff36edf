"""Builds libcodrel.so (HIP kernels + C ABI + C++ reference entry points) for gfx950.

Usage: python -m coding_reliability_amd.build   (or build() from Python)

hipcc cross-compiles the device code for gfx950 without a GPU.  Objects go to
_build/ at the repository root; the shared library is written next to this
file so it travels with the package.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(PKG, "libcodrel.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def sources() -> list[str]:
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON, "-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


CLIS = {"ANcandidates": "ancandidates_main.cpp", "codrel_supera": "supera_main.cpp",
        "hamming_tables": "hamming_tables_main.cpp"}


def build_clis() -> list[str]:
    """The ANcandidates, super-A and Hamming-table command lines over the C ABI,
    into _build/."""
    os.makedirs(BUILD, exist_ok=True)
    out = []
    for name, src in CLIS.items():
        exe = os.path.join(BUILD, name)
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"),
               os.path.join(PKG, "cli", src), "-L", PKG, "-lcodrel",
               "-Wl,-rpath,$ORIGIN/../coding_reliability_amd", "-o", exe]
        subprocess.run(cmd, check=True)
        out.append(exe)
    return out


REFERENCE = os.environ.get("CR_REFERENCE_ROOT", "/root/reference")
REF_CLI = os.path.join(BUILD, "codrel_ref")


def build_reference_cli() -> str | None:
    """Compiles the reference's unchanged distance_distribution/src/main.cpp
    (with its own globals.h / algorithms.h) and links it against libcodrel.so,
    which provides every entry point it calls.  Only where the reference
    sources are mounted; the binary lands in _build/."""
    src_dir = os.path.join(REFERENCE, "distance_distribution", "src")
    main = os.path.join(src_dir, "main.cpp")
    if not os.path.exists(main):
        return None
    os.makedirs(BUILD, exist_ok=True)
    cmd = ["g++", "-std=c++17", "-O2", "-w", "-I", src_dir, main, "-L", PKG, "-lcodrel",
           "-Wl,-rpath,$ORIGIN/../coding_reliability_amd", "-o", REF_CLI]
    subprocess.run(cmd, check=True)
    return REF_CLI


TESTS_CPP = os.path.join(ROOT, "tests", "cpp")


def build_test_programs() -> list[str]:
    """Callers of the library used by the GPU tests, into _build/tests/: the
    C++ ones of the reference entry points (tests/cpp/*.cpp, against
    include/cr_reference_api.h) and the plain C ones of the C ABI
    (tests/cpp/*.c, C99 against include/coding_reliability.h), each linked
    against libcodrel.so."""
    out_dir = os.path.join(BUILD, "tests")
    os.makedirs(out_dir, exist_ok=True)
    out = []
    for src in sorted(glob.glob(os.path.join(TESTS_CPP, "*.cpp")) + glob.glob(os.path.join(TESTS_CPP, "*.c"))):
        exe = os.path.join(out_dir, os.path.splitext(os.path.basename(src))[0])
        lang = ["gcc", "-std=c99"] if src.endswith(".c") else ["g++", "-std=c++17"]
        cmd = [*lang, "-O2", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"), src, "-L", PKG,
               "-lcodrel", "-Wl,-rpath,$ORIGIN/../../coding_reliability_amd", "-o", exe]
        subprocess.run(cmd, check=True)
        out.append(exe)
    return out


if __name__ == "__main__":
    build(verbose=True)
    build_clis()
    build_reference_cli()
    build_test_programs()
