"""ctypes binding of libcodrel.so (the C ABI declared in include/coding_reliability.h).

The HIP library is the only compute path: if libcodrel.so is missing or fails
to load, every call raises.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libcodrel.so")

u64p = C.POINTER(C.c_uint64)
u32p = C.POINTER(C.c_uint32)

CR_OK = 0
STATUS_NAMES = {
    0: "CR_OK", 1: "CR_ERR_INVALID_ARGUMENT", 2: "CR_ERR_UNSUPPORTED", 3: "CR_ERR_NO_DEVICE",
    4: "CR_ERR_OUT_OF_MEMORY", 5: "CR_ERR_DEVICE", 6: "CR_ERR_ADMISSION", 7: "CR_ERR_CANCELLED",
    8: "CR_ERR_NOT_FOUND", 9: "CR_ERR_BUSY",
}


class CRError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class cr_device_info(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("arch", C.c_char * 32), ("compute_units", C.c_int),
                ("max_clock_mhz", C.c_int), ("hbm_bytes", C.c_uint64)]


class cr_ham_request(C.Structure):
    _fields_ = [("k", C.c_uint32), ("source", C.c_uint32), ("first", C.c_uint64), ("count", C.c_uint64),
                ("stride", C.c_uint64), ("list", C.POINTER(C.c_uint64)), ("seed", C.c_uint64),
                ("stream", C.c_uint32), ("reserved", C.c_uint32)]


class cr_pair_tuning(C.Structure):
    _fields_ = [("rows_per_lane", C.c_uint32), ("chunk_blocks", C.c_uint32), ("waves", C.c_uint32),
                ("items_per_wave", C.c_uint32)]


class cr_job(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("k", C.c_uint32), ("A", C.c_uint64), ("stride", C.c_uint64),
                ("samples", C.c_uint64), ("seed", C.c_uint64), ("stream", C.c_uint32), ("reserved", C.c_uint32)]


class cr_server_config(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_active", C.c_uint32), ("rows_per_lane", C.c_uint32),
                ("iteration_ms", C.c_double)]


class cr_job_result(C.Structure):
    _fields_ = [("done", C.c_uint32), ("status", C.c_uint32), ("nbins", C.c_uint32), ("iterations", C.c_uint32),
                ("submit_s", C.c_double), ("admit_s", C.c_double), ("finish_s", C.c_double),
                ("units", C.c_uint64)]


class cr_server_stats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("cancelled", C.c_uint64),
                ("rejected", C.c_uint64), ("iterations", C.c_uint64), ("active", C.c_uint64),
                ("queued", C.c_uint64), ("footprint_bytes", C.c_uint64), ("busy_device_ms", C.c_double)]


class cr_supera_request(C.Structure):
    _fields_ = [("style", C.c_uint32), ("k", C.c_uint32), ("h_bits", C.c_uint32), ("with_grid", C.c_int32),
                ("A_start", C.c_uint64), ("A_end", C.c_uint64), ("As", C.POINTER(C.c_uint64)), ("nAs", C.c_uint32),
                ("verbose", C.c_int32), ("m1", C.c_uint64), ("m2", C.c_uint64)]


class cr_an_code(C.Structure):
    _fields_ = [("A", C.c_uint64), ("num_words", C.c_uint64)]


_lib = None

# name -> (restype, argtypes); every symbol declared in include/*.h is listed here
SIGNATURES = {
    "cr_status_string": (C.c_char_p, [C.c_int]),
    "cr_last_error": (C.c_char_p, []),
    "cr_version": (C.c_char_p, []),
    "cr_context_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "cr_context_destroy": (C.c_int, [C.c_void_p]),
    "cr_device_query": (C.c_int, [C.c_void_p, C.POINTER(cr_device_info)]),
    "cr_device_memory": (C.c_int, [C.c_void_p, u64p, u64p]),
    "cr_probe_ceiling": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_double),
                                   C.POINTER(C.c_double)]),
    "cr_an_pairs": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, u64p, C.c_uint32, C.POINTER(C.c_double)]),
    "cr_an_pairs_tuned": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, u64p, C.c_uint32,
                                    C.POINTER(cr_pair_tuning), C.POINTER(C.c_double)]),
    "cr_an_pairs_noalloc": (C.c_int, [C.c_int, C.c_uint64, C.c_uint64, C.c_void_p, C.c_uint32, C.c_void_p]),
    "cr_an_pairs_spectral": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, u64p, C.c_uint32,
                                       C.POINTER(C.c_double)]),
    "cr_an_grid_ref": (C.c_int, [C.c_void_p, C.c_int, C.c_uint32, C.c_uint64, C.c_uint64, C.c_uint64, u64p,
                                 C.c_uint32, C.POINTER(C.c_double)]),
    "cr_an_grid_report": (C.c_int, [C.c_int, C.c_uint32, C.c_uint64, C.c_uint64, u64p, C.c_uint32, u64p]),
    "cr_hamming_parity_bits": (C.c_uint32, [C.c_uint32]),
    "cr_hamming_weights": (C.c_int, [C.c_void_p, C.POINTER(cr_ham_request), u64p, C.c_uint32,
                                     C.POINTER(C.c_double)]),
    "cr_hamming_weights_batch": (C.c_int, [C.c_void_p, C.POINTER(cr_ham_request), C.c_uint32, u64p, C.c_uint32,
                                           C.POINTER(C.c_double)]),
    "cr_hamming_report": (C.c_int, [C.c_uint32, u64p, C.c_int, u64p]),
    "cr_hamming_sdc_table": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int, u64p, C.c_uint32, C.POINTER(C.c_void_p)]),
    "cr_an_batch": (C.c_int, [C.c_void_p, C.POINTER(cr_an_code), C.c_uint32, u64p, C.c_uint32,
                              C.POINTER(cr_pair_tuning), C.POINTER(C.c_double)]),
    "cr_batch_plan_create": (C.c_int, [C.c_void_p, C.POINTER(cr_an_code), C.c_uint32, C.c_uint32,
                                       C.POINTER(cr_pair_tuning), C.POINTER(C.c_void_p)]),
    "cr_batch_plan_run": (C.c_int, [C.c_void_p, u64p, C.POINTER(C.c_double)]),
    "cr_batch_plan_destroy": (C.c_int, [C.c_void_p]),
    "cr_autotune_batch": (C.c_int, [C.c_void_p, C.POINTER(cr_an_code), C.c_uint32, C.c_uint32,
                                    C.POINTER(cr_pair_tuning), C.POINTER(C.c_double)]),
    "cr_server_create": (C.c_int, [C.c_void_p, C.POINTER(cr_server_config), C.POINTER(C.c_void_p)]),
    "cr_server_submit": (C.c_int, [C.c_void_p, C.POINTER(cr_job), C.POINTER(C.c_uint64)]),
    "cr_server_result": (C.c_int, [C.c_void_p, C.c_uint64, u64p, C.c_uint32, C.POINTER(cr_job_result)]),
    "cr_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.c_double]),
    "cr_server_cancel": (C.c_int, [C.c_void_p, C.c_uint64]),
    "cr_server_stats_get": (C.c_int, [C.c_void_p, C.POINTER(cr_server_stats)]),
    "cr_server_now": (C.c_double, [C.c_void_p]),
    "cr_server_destroy": (C.c_int, [C.c_void_p]),
    "cr_supera_search": (C.c_int, [C.c_void_p, C.POINTER(cr_supera_request), C.POINTER(C.c_void_p),
                                   C.POINTER(C.c_double)]),
    "cr_free": (None, [C.c_void_p]),
    "cr_an_candidates": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, u64p, u64p, C.c_uint64, u64p,
                                   C.POINTER(C.c_double)]),
    "cr_an_mc": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint64, u64p, C.c_uint32,
                           C.POINTER(C.c_double)]),
    "cr_an_mc_ranges": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, u64p, C.c_uint32, C.c_uint64, u64p,
                                  C.c_uint32, C.POINTER(C.c_double)]),
}


def _one_hip_runtime() -> None:
    """PyTorch wheels for ROCm carry their own HIP and HSA runtimes
    (torch/lib/libamdhip64.so, soname libamdhip64.so.7), and a process can
    bring up the GPU through one runtime instance only: whichever starts
    second finds no device.  libcodrel.so needs libamdhip64.so.7, so when
    torch is installed its runtime is loaded first (without importing torch);
    libcodrel.so then binds to it by soname, and a later `import torch`
    reuses the same file.  Device pointers and streams from torch and from
    this library are then valid in each other's calls.  A process that has a
    HIP runtime mapped already keeps it: libcodrel.so binds to that one."""
    import importlib.util
    try:
        with open("/proc/self/maps") as maps:
            if any("libamdhip64" in line for line in maps):
                return
    except OSError:
        pass
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        return
    if spec is None or not spec.origin:
        return
    hip = os.path.join(os.path.dirname(spec.origin), "lib", "libamdhip64.so")
    if os.path.exists(hip):
        C.CDLL(hip, mode=C.RTLD_GLOBAL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CRError(3, f"{LIB_PATH} is missing; build it with "
                             "`python -m coding_reliability_amd.build` (no CPU fallback exists)")
        _one_hip_runtime()
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != CR_OK:
        raise CRError(status, lib().cr_last_error().decode(errors="replace"))
