"""ctypes binding of the C ABI in ``include/pyhea_hcs.h`` (libpyhea_hcs.so).

The library is built in-tree by ``pyhea_amd.build`` (or
``__graft_entry__.build()``).  There is no fallback: if the HIP library is
missing, every entry point raises, so no result can silently come from
anything but the GPU kernels.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

HEA_OK, HEA_EINVAL, HEA_ENODEV, HEA_EHIP, HEA_EUNSUPPORTED, HEA_ENOMEM = range(6)
PHASE_INIT, PHASE_NEW, PHASE_LS1, PHASE_LS2 = range(4)

LIB_PATH = os.environ.get(
    "PYHEA_AMD_LIB", os.path.join(os.path.dirname(os.path.abspath(__file__)), "libpyhea_hcs.so"))

# Every symbol declared by include/pyhea_hcs.h.
EXPORTS = (
    "hea_device_count",
    "hea_last_error",
    "hea_problem_create",
    "hea_problem_destroy",
    "hea_hcs_run",
    "hea_local_mc",
    "hea_local_mc_device",
    "hea_local_mc_device_noalloc",
    "hea_fitness",
    "hea_pair_counts",
    "hea_generate",
    "hea_session_create",
    "hea_session_init",
    "hea_session_tile",
    "hea_session_guide",
    "hea_guided_scores",
    "hea_session_step",
    "hea_session_step_timed",
    "hea_session_sync",
    "hea_session_read",
    "hea_session_best",
    "hea_session_destroy",
    "hea_batch_create",
    "hea_batch_grid",
    "hea_batch_tile",
    "hea_batch_guide",
    "hea_batch_init",
    "hea_batch_step",
    "hea_batch_step_timed",
    "hea_batch_sync",
    "hea_batch_read",
    "hea_batch_destroy",
    "hea_relax_forward",
    "hea_relax_backward",
    "hea_relax_cols",
    "hea_site_gain",
)


class HeaProblemDesc(C.Structure):
    _fields_ = [
        ("n_sites", C.c_int32),
        ("n_species", C.c_int32),
        ("n_shells", C.c_int32),
        ("shell_size", C.c_int32 * 3),
        ("nbr", C.POINTER(C.c_int32)),
        ("counts", C.POINTER(C.c_int32)),
        ("weight", C.POINTER(C.c_double)),
        ("target", C.POINTER(C.c_double)),
    ]


class HeaStats(C.Structure):
    _fields_ = [
        ("evals", C.c_int64),
        ("rounds", C.c_int64),
        ("accepts", C.c_int64),
        ("draws", C.c_int64),
        ("first_hit_s", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class HeaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[hea error {code}] {msg}")
        self.code = code


_lib = None
_lock = threading.Lock()


def _declare(lib: C.CDLL) -> None:
    P, I32, U32, U64, D = C.c_void_p, C.c_int32, C.c_uint32, C.c_uint64, C.c_double
    u8p, dp, i32p = C.POINTER(C.c_uint8), C.POINTER(C.c_double), C.POINTER(C.c_int32)
    sig = {
        "hea_device_count": (C.c_int, []),
        "hea_last_error": (C.c_char_p, []),
        "hea_problem_create": (C.c_int, [C.POINTER(HeaProblemDesc), C.c_int, C.POINTER(P)]),
        "hea_problem_destroy": (C.c_int, [P]),
        "hea_hcs_run": (C.c_int, [P, I32, I32, I32, I32, D, U64, u8p, dp, C.POINTER(HeaStats)]),
        "hea_local_mc": (
            C.c_int,
            [P, I32, u8p, dp, I32, I32, U64, U32, U32, I32, C.POINTER(HeaStats)],
        ),
        "hea_local_mc_device": (C.c_int, [P, I32, P, P, I32, I32, U64, U32, U32, P]),
        "hea_local_mc_device_noalloc": (C.c_int, [P, I32, P, P, I32, I32, U64, U32, U32, P]),
        "hea_fitness": (C.c_int, [P, I32, u8p, dp]),
        "hea_pair_counts": (C.c_int, [P, I32, u8p, i32p]),
        "hea_generate": (C.c_int, [P, I32, U64, U32, U32, u8p, dp]),
        "hea_session_create": (C.c_int, [P, I32, I32, I32, D, U64, C.POINTER(P)]),
        "hea_session_init": (C.c_int, [P]),
        "hea_session_tile": (C.c_int, [P, I32]),
        "hea_session_guide": (C.c_int, [P, I32, D, I32]),
        "hea_guided_scores": (C.c_int, [P, I32, u8p, I32, I32, U64, U32, U32, i32p,
                                        C.POINTER(C.c_float)]),
        "hea_session_step": (C.c_int, [P, I32]),
        "hea_session_step_timed": (C.c_int, [P, I32, C.POINTER(C.c_float),
                                             C.POINTER(C.c_int64)]),
        "hea_session_sync": (C.c_int, [P]),
        "hea_session_read": (C.c_int, [P, u8p, dp, C.POINTER(HeaStats)]),
        "hea_session_best": (C.c_int, [P, dp]),
        "hea_session_destroy": (C.c_int, [P]),
        "hea_batch_create": (C.c_int, [C.POINTER(P), I32, I32, I32, I32, D, U64, C.POINTER(P)]),
        "hea_batch_grid": (C.c_int, [P, I32]),
        "hea_batch_tile": (C.c_int, [P, I32]),
        "hea_batch_guide": (C.c_int, [P, I32, D, I32]),
        "hea_batch_init": (C.c_int, [P]),
        "hea_batch_step": (C.c_int, [P, I32]),
        "hea_batch_step_timed": (C.c_int, [P, I32, C.POINTER(C.c_float),
                                           C.POINTER(HeaStats)]),
        "hea_batch_sync": (C.c_int, [P]),
        "hea_batch_read": (C.c_int, [P, I32, u8p, dp, C.POINTER(HeaStats)]),
        "hea_batch_destroy": (C.c_int, [P]),
        "hea_relax_forward": (C.c_int, [P, I32, P, P, P, P, P]),
        "hea_relax_backward": (C.c_int, [P, I32, P, P, P, P, P]),
        "hea_relax_cols": (I32, [I32]),
        "hea_site_gain": (C.c_int, [P, I32, P, P, P, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def lib() -> C.CDLL:
    """The loaded HIP library; raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(
                    f"{LIB_PATH} is missing: build the HIP extension first "
                    "(python -m pyhea_amd.build or __graft_entry__.build())"
                )
            handle = C.CDLL(LIB_PATH)
            _declare(handle)
            _lib = handle
    return _lib


def check(rc: int) -> None:
    if rc != HEA_OK:
        msg = lib().hea_last_error().decode(errors="replace")
        if rc in (HEA_EINVAL, HEA_EUNSUPPORTED):
            raise ValueError(f"[hea error {rc}] {msg}")
        raise HeaError(rc, msg)


def device_count() -> int:
    return int(lib().hea_device_count())
