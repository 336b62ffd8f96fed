"""ctypes front end of the CPU parity checker (``oracle/ldsc_oracle.c``).

Test infrastructure only: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` to compute expected answers.
The product package ``ldsc_amd`` never imports this module.

Every wrapper names the reference function it restates (paths are relative
to the sharifhsn/ldsc source tree).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

WINDOW_CM, WINDOW_KB, WINDOW_SNP = 0, 1, 2

_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C")
_i8p = np.ctypeslib.ndpointer(np.int8, flags="C")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C")


def build() -> str:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp = C.c_void_p
        L.orc_bed_validate.argtypes = [_u8p, C.c_uint64, C.c_uint64, C.c_uint64]
        L.orc_bed_validate.restype = C.c_int
        L.orc_decode.argtypes = [vp, C.c_uint64, C.c_uint64, vp, C.c_uint64, _u64p,
                                 C.c_uint64, _i8p]
        L.orc_stats.argtypes = [vp, C.c_uint64, C.c_uint64, vp, C.c_uint64, _u64p,
                                C.c_uint64, _u32p]
        L.orc_block_left.argtypes = [_u8p, _f64p, _u32p, C.c_uint64, C.c_int, C.c_double,
                                     _u64p]
        L.orc_block_left.restype = C.c_int
        L.orc_countsketch_hash.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, _u32p, _i8p]
        L.orc_sketch_int.argtypes = [vp, C.c_uint64, C.c_uint64, vp, C.c_uint64, _u64p,
                                     C.c_uint64, C.c_uint64, _u32p, _i8p, _i32p, _i32p]
        L.orc_std_f32.argtypes = [vp, C.c_uint64, C.c_uint64, vp, C.c_uint64, _u64p,
                                  C.c_uint64, _f32p]
        L.orc_l2.argtypes = [vp, C.c_uint64, C.c_uint64, vp, C.c_uint64, _u8p, _f64p, _u32p,
                             _u64p, C.c_uint64, C.c_int, C.c_double, C.c_uint64, C.c_int,
                             C.c_uint64, C.c_int, C.c_int, C.c_double, vp, C.c_uint64,
                             _f64p, _f64p, C.c_int]
        L.orc_l2.restype = C.c_int
        L.orc_num_threads.restype = C.c_int
        L.orc_set_projection.argtypes = [vp, vp, C.c_uint64, C.c_uint64]
        L.orc_kappa.argtypes = [vp, C.c_uint64, C.c_uint64, vp, C.c_uint64, _u64p,
                                C.c_uint64, _f64p]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _body(bed: np.ndarray) -> np.ndarray:
    """The .bed body after the 3-byte header."""
    return np.ascontiguousarray(bed[3:])


def bed_validate(bed: np.ndarray, n_indiv: int, m: int) -> int:
    """Bed::from_source validation (src/bed.rs:76-109). 0 = ok."""
    b = np.ascontiguousarray(bed, dtype=np.uint8)
    if b.size == 0:
        b = np.zeros(1, np.uint8)
        return lib().orc_bed_validate(b, 0, n_indiv, m)
    return lib().orc_bed_validate(b, b.size, n_indiv, m)


def _keep(keep):
    if keep is None:
        return None, 0
    k = np.ascontiguousarray(keep, dtype=np.int64)
    return k, k.size


def decode(bed, n_fam, bed_idx, keep=None) -> np.ndarray:
    """decode_column with the count-A1 LUT [2, NaN, 1, 0] (src/bed.rs:470-521);
    returns N x m int8 with -1 for missing."""
    body = _body(bed)
    bps = (n_fam + 3) // 4
    k, nk = _keep(keep)
    n = nk if k is not None else n_fam
    idx = np.ascontiguousarray(bed_idx, dtype=np.uint64)
    out = np.zeros(idx.size * n, np.int8)
    lib().orc_decode(_ptr(body), bps, n_fam, _ptr(k), nk, idx, idx.size, out)
    return out.reshape(idx.size, n).T


def stats(bed, n_fam, bed_idx, keep=None) -> np.ndarray:
    """snp_stats_from_bytes (src/l2/compute.rs:363-427) + het count; m x 4 uint32."""
    body = _body(bed)
    bps = (n_fam + 3) // 4
    k, nk = _keep(keep)
    idx = np.ascontiguousarray(bed_idx, dtype=np.uint64)
    out = np.zeros(idx.size * 4, np.uint32)
    lib().orc_stats(_ptr(body), bps, n_fam, _ptr(k), nk, idx, idx.size, out)
    return out.reshape(idx.size, 4)


def std_f32(bed, n_fam, bed_idx, keep=None) -> np.ndarray:
    """normalize_col_f32_with_stats (src/l2/normalize.rs:115-154); N x m float32."""
    body = _body(bed)
    bps = (n_fam + 3) // 4
    k, nk = _keep(keep)
    n = nk if k is not None else n_fam
    idx = np.ascontiguousarray(bed_idx, dtype=np.uint64)
    out = np.zeros(idx.size * n, np.float32)
    lib().orc_std_f32(_ptr(body), bps, n_fam, _ptr(k), nk, idx, idx.size, out)
    return out.reshape(idx.size, n).T


def block_left(chr_, cm, bp, mode: int, value: float):
    """get_block_lefts_by_chr (src/l2/window.rs:32-90) -> (bl, any_full)."""
    c = np.ascontiguousarray(chr_, dtype=np.uint8)
    cmv = np.ascontiguousarray(cm, dtype=np.float64)
    bpv = np.ascontiguousarray(bp, dtype=np.uint32)
    out = np.zeros(c.size, np.uint64)
    full = lib().orc_block_left(c, cmv, bpv, c.size, mode, float(value), out)
    return out, bool(full)


def countsketch_hash(n: int, d: int, seed: int = 42):
    """CountSketchProj::new (src/l2/compute.rs:179-192), fastrand 2.3.0 stream."""
    b = np.zeros(n, np.uint32)
    s = np.zeros(n, np.int8)
    lib().orc_countsketch_hash(n, d, seed, b, s)
    return b, s


def sketch_int(bed, n_fam, bed_idx, d, keep=None, seed=42):
    """Exact integer CountSketch buckets (sg, sm), each d x m int32."""
    body = _body(bed)
    bps = (n_fam + 3) // 4
    k, nk = _keep(keep)
    n = nk if k is not None else n_fam
    idx = np.ascontiguousarray(bed_idx, dtype=np.uint64)
    b, s = countsketch_hash(n, d, seed)
    sg = np.zeros(d * idx.size, np.int32)
    sm = np.zeros(d * idx.size, np.int32)
    lib().orc_sketch_int(_ptr(body), bps, n_fam, _ptr(k), nk, idx, idx.size, d, b, s, sg, sm)
    return sg.reshape(idx.size, d).T, sm.reshape(idx.size, d).T


def l2(bed, n_fam, chr_, cm, bp, bed_idx=None, *, mode=WINDOW_CM, window=1.0, chunk=200,
       use_f32=False, sketch=0, maf_aware=False, masking=False, pq_exp=None, annot=None,
       keep=None, per_chromosome=False, threads=0):
    """compute_ldscore_global (src/l2/compute.rs:1009-2447) over the given SNP list.

    per_chromosome=True splits the list into contiguous chromosome runs and
    calls the restatement once per run, like run() (src/l2/mod.rs:478-553).
    Returns (l2 [m x K], maf [m]).
    """
    chr_ = np.ascontiguousarray(chr_, dtype=np.uint8)
    m = chr_.size
    cm = np.ascontiguousarray(cm, dtype=np.float64)
    bp = np.ascontiguousarray(bp, dtype=np.uint32)
    if bed_idx is None:
        bed_idx = np.arange(m, dtype=np.uint64)
    bed_idx = np.ascontiguousarray(bed_idx, dtype=np.uint64)
    if annot is not None:
        annot = np.asarray(annot, dtype=np.float64)
        if annot.ndim == 1:
            annot = annot[:, None]
    K = 1 if annot is None else annot.shape[1]
    if per_chromosome:
        outs, mafs = [], []
        s = 0
        while s < m:
            e = s + 1
            while e < m and chr_[e] == chr_[s]:
                e += 1
            o, f = l2(bed, n_fam, chr_[s:e], cm[s:e], bp[s:e], bed_idx[s:e], mode=mode,
                      window=window, chunk=chunk, use_f32=use_f32, sketch=sketch,
                      maf_aware=maf_aware, masking=masking, pq_exp=pq_exp,
                      annot=None if annot is None else annot[s:e], keep=keep,
                      per_chromosome=False, threads=threads)
            outs.append(o)
            mafs.append(f)
            s = e
        return np.concatenate(outs, axis=0), np.concatenate(mafs)
    body = _body(bed)
    bps = (n_fam + 3) // 4
    k, nk = _keep(keep)
    a = None if annot is None else np.asfortranarray(annot).ravel(order="F")
    out = np.zeros(m * K, np.float64)
    maf = np.zeros(m, np.float64)
    rc = lib().orc_l2(_ptr(body), bps, n_fam, _ptr(k), nk, chr_, cm, bp, bed_idx, m, mode,
                      float(window), chunk, int(use_f32), sketch, int(maf_aware),
                      int(masking), float("nan") if pq_exp is None else float(pq_exp),
                      _ptr(a), K, out, maf, threads)
    if rc == 1:
        raise ValueError("compute_ldscore_global: chunk_size must be >= 1")
    if rc == 2:
        raise ValueError(f"--sketch {sketch} is out of bounds")
    return out.reshape(K, m).T, maf


def kappa(bed, n_fam, bed_idx, keep=None) -> np.ndarray:
    """per_snp_kappa (src/l2/compute.rs:525-538) of the listed SNPs, from the
    integer moments of the decoded column."""
    body = _body(bed)
    bps = (n_fam + 3) // 4
    k, nk = _keep(keep)
    idx = np.ascontiguousarray(bed_idx, dtype=np.uint64)
    out = np.zeros(idx.size, np.float64)
    lib().orc_kappa(_ptr(body), bps, n_fam, _ptr(k), nk, idx, idx.size, out)
    return out


class projection:
    """Context manager: l2(..., sketch=d) inside it uses the given bucket /
    sign arrays (length n, panel order) instead of the seed-42 hash, and
    allows d = n.  Test hook for checking the sketch chain apart from the
    hash."""

    def __init__(self, bucket, sign, d):
        self.b = np.ascontiguousarray(bucket, dtype=np.uint32)
        self.s = np.ascontiguousarray(sign, dtype=np.int8)
        self.d = int(d)

    def __enter__(self):
        lib().orc_set_projection(_ptr(self.b), _ptr(self.s), self.b.size, self.d)
        return self

    def __exit__(self, *exc):
        lib().orc_set_projection(None, None, 0, 0)
        return False


def num_threads() -> int:
    return lib().orc_num_threads()
