"""ctypes binding of ``include/ldsc_l2.h`` (libldsc_l2.so).

The library is loaded from the package directory only; when it is missing the
import of any compute entry point fails loudly -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libldsc_l2.so")

# error codes (include/ldsc_l2.h)
LDSC_OK = 0
ERR_NAMES = {
    1: "INVALID_ARG", 2: "BAD_BED", 3: "EMPTY", 4: "SKETCH_RANGE", 5: "ANNOT",
    6: "DEVICE", 7: "OOM", 8: "BUSY", 9: "NOT_FOUND", 10: "PENDING", 11: "CANCELLED",
    12: "TIMEOUT",
}
ERR_PENDING = 10
ERR_CANCELLED = 11
ERR_TIMEOUT = 12
WINDOW_CM, WINDOW_KB, WINDOW_SNP = 0, 1, 2
PREC_EXACT, PREC_FP32, PREC_EXACT_I8 = 0, 1, 2
PREC_INT8_FP64 = PREC_EXACT

EXPORTED = [
    "ldsc_l2_open_panel", "ldsc_l2_close_panel", "ldsc_l2_panel_n_indiv", "ldsc_l2_compute",
    "ldsc_l2_block_left", "ldsc_l2_pair_bounds", "ldsc_l2_decode", "ldsc_l2_countsketch_hash",
    "ldsc_l2_sketch_buckets", "ldsc_server_create", "ldsc_server_destroy", "ldsc_l2_submit",
    "ldsc_l2_poll", "ldsc_l2_wait", "ldsc_server_stats_get", "ldsc_l2_cancel",
    "ldsc_l2_compute_noalloc",
]


class LdscError(RuntimeError):
    """An LDSC_ERR_* return of the C ABI; ``code`` holds the number."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{ERR_NAMES.get(code, code)}] {msg}")
        self.code = code


# numpy mirror of ldsc_snp (32 bytes)
SNP_DTYPE = np.dtype([("chr", "u1"), ("pad_", "u1", (7,)), ("cm", "<f8"), ("bp", "<u4"),
                      ("pad2_", "<u4"), ("bed_idx", "<u8")])


class Config(C.Structure):
    _fields_ = [("window_mode", C.c_int32), ("precision", C.c_int32),
                ("window_value", C.c_double), ("chunk_size", C.c_uint64),
                ("sketch_d", C.c_uint64), ("sketch_maf_aware", C.c_int32),
                ("snp_level_masking", C.c_int32), ("pq_exp", C.c_double),
                ("yes_really", C.c_int32), ("per_chromosome", C.c_int32)]


class Perf(C.Structure):
    _fields_ = [("m", C.c_uint64), ("n_indiv", C.c_uint64), ("bed_read_secs", C.c_double),
                ("norm_secs", C.c_double), ("sketch_secs", C.c_double),
                ("bb_dot_secs", C.c_double), ("ab_dot_secs", C.c_double),
                ("ring_store_secs", C.c_double), ("total_secs", C.c_double),
                ("pairs_included", C.c_uint64), ("pairs_evaluated", C.c_uint64),
                ("tiles", C.c_uint64), ("any_full_chr_window", C.c_int32), ("sketch_int8", C.c_int32)]


class Progress(C.Structure):
    _fields_ = [("chunks_done", C.c_uint64), ("chunks_total", C.c_uint64),
                ("snps_done", C.c_uint64), ("snps_total", C.c_uint64),
                ("chunk_wall_ms", C.c_double)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_inflight", C.c_uint32),
                ("queue_capacity", C.c_uint32), ("batch_window_us", C.c_uint32),
                ("max_queue_ms", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [("requests_done", C.c_uint64), ("band_launches", C.c_uint64),
                ("batched_requests", C.c_uint64), ("max_batch", C.c_uint32),
                ("reserved", C.c_uint32), ("requests_cancelled", C.c_uint64),
                ("requests_timed_out", C.c_uint64), ("hbm_charged", C.c_uint64),
                ("hbm_charged_max", C.c_uint64), ("workspace_trims", C.c_uint64)]


PROGRESS_CB = C.CFUNCTYPE(None, C.POINTER(Progress), C.c_void_p)

_lib = None


def lib():
    """Load libldsc_l2.so (built in-tree by ``python -m ldsc_amd.build``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -m ldsc_amd.build` (no CPU fallback exists)")
    L = C.CDLL(LIB_PATH)
    vp, u64, i32, dbl, cp = C.c_void_p, C.c_uint64, C.c_int32, C.c_double, C.c_char_p
    L.ldsc_l2_open_panel.argtypes = [vp, u64, u64, u64, vp, u64, C.POINTER(vp), cp, u64]
    L.ldsc_l2_close_panel.argtypes = [vp]
    L.ldsc_l2_panel_n_indiv.argtypes = [vp]
    L.ldsc_l2_panel_n_indiv.restype = u64
    L.ldsc_l2_compute.argtypes = [vp, vp, u64, C.POINTER(Config), vp, u64, vp, vp,
                                  C.POINTER(Perf), PROGRESS_CB, vp, cp, u64]
    L.ldsc_l2_block_left.argtypes = [vp, u64, i32, dbl, vp, C.POINTER(i32), cp, u64]
    L.ldsc_l2_pair_bounds.argtypes = [vp, u64, C.POINTER(Config), vp, cp, u64]
    L.ldsc_l2_decode.argtypes = [vp, vp, u64, vp, vp, vp, vp, cp, u64]
    L.ldsc_l2_countsketch_hash.argtypes = [u64, u64, u64, vp, vp, cp, u64]
    L.ldsc_l2_sketch_buckets.argtypes = [vp, vp, u64, u64, vp, vp, cp, u64]
    L.ldsc_server_create.argtypes = [C.POINTER(ServerConfig), C.POINTER(vp), cp, u64]
    L.ldsc_server_destroy.argtypes = [vp]
    L.ldsc_l2_submit.argtypes = [vp, vp, vp, u64, C.POINTER(Config), vp, u64, vp, vp,
                                 C.POINTER(u64), cp, u64]
    L.ldsc_l2_poll.argtypes = [vp, u64, C.POINTER(dbl), cp, u64]
    L.ldsc_l2_wait.argtypes = [vp, u64, C.POINTER(dbl), cp, u64]
    L.ldsc_server_stats_get.argtypes = [vp, C.POINTER(ServerStats)]
    L.ldsc_l2_cancel.argtypes = [vp, u64, cp, u64]
    L.ldsc_l2_compute_noalloc.argtypes = [vp, vp, u64, C.POINTER(Config), vp, u64, vp, vp,
                                          C.POINTER(Perf), PROGRESS_CB, vp, cp, u64]
    for name in EXPORTED:
        getattr(L, name)  # every declared symbol must resolve
    _lib = L
    return L


def errbuf():
    return C.create_string_buffer(512)


def check(rc: int, err) -> None:
    if rc != LDSC_OK:
        raise LdscError(rc, err.value.decode(errors="replace"))


def ptr(a) -> C.c_void_p | None:
    return None if a is None else C.c_void_p(a.ctypes.data)
