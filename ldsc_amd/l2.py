"""Host mirror of the reference's l2 library surface over the C ABI.

Names, argument meaning and error behaviour follow sharifhsn/ldsc:
``WindowMode`` (src/l2/window.rs:5-12), ``L2Config`` (src/l2/mod.rs:951-988),
``L2Output`` / ``L2Perf`` / ``L2Progress`` (mod.rs:993-1034, 22-39),
``compute_l2_from_bytes[_with_progress]`` (mod.rs:1049-1085),
``compute_l2_from_bed[_with_progress]`` (mod.rs:1099-1169),
``parse_bim_str`` (src/parse.rs:520-592) and ``count_fam_str``
(src/l2/io.rs:16-22).  ``Panel`` replaces ``Bed``: the packed genotypes stay
resident in HBM and every computation runs in libldsc_l2.so.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field
from typing import Callable, Optional, Sequence

import numpy as np

from . import _native as N


# ----------------------------------------------------------------------------
# Configuration
# ----------------------------------------------------------------------------
@dataclass(frozen=True)
class WindowMode:
    """Window definition: ``WindowMode.Cm(x)``, ``.Kb(x)`` or ``.Snp(n)``."""

    kind: int
    value: float

    @classmethod
    def Cm(cls, x: float) -> "WindowMode":
        return cls(N.WINDOW_CM, float(x))

    @classmethod
    def Kb(cls, x: float) -> "WindowMode":
        return cls(N.WINDOW_KB, float(x))

    @classmethod
    def Snp(cls, n: int) -> "WindowMode":
        return cls(N.WINDOW_SNP, float(int(n)))


@dataclass
class L2Config:
    """Knobs of one LD-score computation (mirrors ``L2Config``).

    ``use_f32`` selects the FP32 standardized-column form (``--fast-f32``);
    otherwise the exact form runs as exact integer products with an FP64
    finish (the reference's default f64 path): on the e2m1 MFMA for up to
    1,800,000 individuals, on the int8 MFMA beyond or when
    ``exact_kernel="i8"``.  ``per_chromosome`` restarts
    the chunk grid at every chromosome run, as ``run()`` does without
    ``--global-pass``; the reference's library API is one global pass.
    """

    mode: WindowMode = field(default_factory=lambda: WindowMode.Kb(1000.0))
    chunk_size: int = 200
    use_f32: bool = False
    sketch: Optional[int] = None
    sketch_maf_aware: bool = False
    snp_level_masking: bool = False
    yes_really: bool = True
    pq_exp: Optional[float] = None
    verbose_timing: bool = False
    per_chromosome: bool = False
    exact_kernel: str = "auto"

    def to_c(self) -> N.Config:
        c = N.Config()
        c.window_mode = self.mode.kind
        c.window_value = self.mode.value
        c.chunk_size = int(self.chunk_size)
        if self.exact_kernel not in ("auto", "i8"):
            raise ValueError(f"exact_kernel must be 'auto' or 'i8', got {self.exact_kernel!r}")
        c.precision = (N.PREC_FP32 if self.use_f32 else
                       N.PREC_EXACT_I8 if self.exact_kernel == "i8" else N.PREC_EXACT)
        c.sketch_d = int(self.sketch or 0)
        c.sketch_maf_aware = int(bool(self.sketch_maf_aware))
        c.snp_level_masking = int(bool(self.snp_level_masking))
        c.pq_exp = float("nan") if self.pq_exp is None else float(self.pq_exp)
        c.yes_really = int(bool(self.yes_really))
        c.per_chromosome = int(bool(self.per_chromosome))
        return c


@dataclass
class BimRecord:
    chr: int
    snp: str
    cm: float
    bp: int
    bed_idx: int


@dataclass
class L2Perf:
    m: int = 0
    n_indiv: int = 0
    bed_read_secs: float = 0.0
    norm_secs: float = 0.0
    sketch_secs: Optional[float] = None
    bb_dot_secs: float = 0.0
    ab_dot_secs: float = 0.0
    ring_store_secs: float = 0.0
    total_secs: float = 0.0
    pairs_included: int = 0
    pairs_evaluated: int = 0
    tiles: int = 0
    sketch_int8: bool = False  # the sketch contraction ran on exact int8 operands

    @classmethod
    def from_c(cls, p: N.Perf, sketch: bool) -> "L2Perf":
        return cls(p.m, p.n_indiv, p.bed_read_secs, p.norm_secs,
                   p.sketch_secs if sketch else None, p.bb_dot_secs, p.ab_dot_secs,
                   p.ring_store_secs, p.total_secs, p.pairs_included, p.pairs_evaluated,
                   p.tiles, bool(p.sketch_int8))


@dataclass
class L2Progress:
    chunks_done: int
    chunks_total: int
    snps_done: int
    snps_total: int
    chunk_wall_ms: float


@dataclass
class L2Output:
    snps: list
    l2: np.ndarray
    maf: np.ndarray
    wall_seconds: float
    perf: L2Perf


# ----------------------------------------------------------------------------
# Text inputs
# ----------------------------------------------------------------------------
def _lines(s: str):
    """``BufRead::lines`` semantics: split on \\n, strip one \\r before a
    \\n, no trailing empty record (a final line without \\n keeps its \\r)."""
    if not s:
        return []
    parts = s.split("\n")
    last = parts.pop()
    out = [p[:-1] if p.endswith("\r") else p for p in parts]
    if last:
        out.append(last)
    return out


def _parse_u(s: str, bits: int) -> Optional[int]:
    t = s[1:] if s.startswith("+") else s
    if not t or not t.isascii() or not t.isdigit():
        return None
    v = int(t)
    return v if v < (1 << bits) else None


def _parse_f64(s: str) -> Optional[float]:
    """Rust ``f64::from_str``: ASCII only, no surrounding space, no ``_``."""
    if not s or not s.isascii() or s != s.strip() or "_" in s:
        return None
    try:
        return float(s)
    except ValueError:
        return None


def parse_bim_str(s: str) -> list:
    """parse_bim_reader_filtered (src/parse.rs:550-592): tab-separated,
    chr as u8 or 0, cM as f64 or 0, bp as u32 or 0; A1/A2 must exist."""
    out = []
    for line_no, line in enumerate(_lines(s)):
        cols = line.split("\t")
        if len(cols) < 2:
            raise ValueError(f"BIM line {line_no + 1}: expected SNP column")
        if len(cols) < 5:
            raise ValueError(f"BIM line {line_no + 1}: missing A1 col")
        if len(cols) < 6:
            raise ValueError(f"BIM line {line_no + 1}: missing A2 col")
        chr_ = _parse_u(cols[0], 8) or 0
        cm = _parse_f64(cols[2])
        bp = _parse_u(cols[3], 32)
        out.append(BimRecord(chr_, cols[1], 0.0 if cm is None else cm, bp or 0, line_no))
    return out


def count_fam_str(s: str) -> int:
    """count_fam_str (src/l2/io.rs:16-22)."""
    return len(_lines(s))


def snps_to_array(snps: Sequence[BimRecord]) -> np.ndarray:
    a = np.zeros(len(snps), dtype=N.SNP_DTYPE)
    a["chr"] = [s.chr for s in snps]
    a["cm"] = [s.cm for s in snps]
    a["bp"] = [s.bp for s in snps]
    a["bed_idx"] = [s.bed_idx for s in snps]
    return a


def make_snp_array(chr_, cm, bp, bed_idx=None) -> np.ndarray:
    chr_ = np.asarray(chr_)
    a = np.zeros(chr_.size, dtype=N.SNP_DTYPE)
    a["chr"] = chr_
    a["cm"] = np.asarray(cm, dtype=np.float64)
    a["bp"] = np.asarray(bp, dtype=np.uint32)
    a["bed_idx"] = np.arange(chr_.size) if bed_idx is None else np.asarray(bed_idx)
    return a


# ----------------------------------------------------------------------------
# Panels
# ----------------------------------------------------------------------------
class Panel:
    """A validated PLINK .bed resident on the current HIP device."""

    def __init__(self, bed_bytes, n_indiv: int, m_snps: int, keep=None):
        L = N.lib()
        buf = np.frombuffer(bed_bytes, dtype=np.uint8) if isinstance(
            bed_bytes, (bytes, bytearray, memoryview)) else np.ascontiguousarray(bed_bytes,
                                                                                 dtype=np.uint8)
        k = None if keep is None else np.ascontiguousarray(keep, dtype=np.int64)
        h = C.c_void_p()
        err = N.errbuf()
        rc = L.ldsc_l2_open_panel(N.ptr(buf), buf.size, int(n_indiv), int(m_snps), N.ptr(k),
                                  0 if k is None else k.size, C.byref(h), err, 512)
        N.check(rc, err)
        self._h = h
        self.n_fam = int(n_indiv)
        self.m_bim = int(m_snps)
        self.n_indiv = int(L.ldsc_l2_panel_n_indiv(h))

    @classmethod
    def from_body(cls, body: np.ndarray, n_indiv: int, m_snps: int, keep=None) -> "Panel":
        """Open from the SNP slabs without the 3-byte header."""
        full = np.empty(body.size + 3, np.uint8)
        full[:3] = (0x6C, 0x1B, 0x01)
        full[3:] = body.reshape(-1)
        return cls(full, n_indiv, m_snps, keep)

    @classmethod
    def from_file(cls, bed_path: str, n_indiv: int, m_snps: int, keep=None) -> "Panel":
        data = np.fromfile(bed_path, dtype=np.uint8)
        return cls(data, n_indiv, m_snps, keep)

    @property
    def handle(self) -> C.c_void_p:
        if self._h is None:
            raise ValueError("panel is closed")
        return self._h

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            N.lib().ldsc_l2_close_panel(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ----------------------------------------------------------------------------
# Compute
# ----------------------------------------------------------------------------
def compute_ldscore(panel: Panel, snps: np.ndarray, config: L2Config, annot=None,
                    on_progress: Optional[Callable[[L2Progress], None]] = None, out=None):
    """compute_ldscore_global (src/l2/compute.rs:1009-2447) on the device.

    ``snps`` is an array of ``SNP_DTYPE``; ``annot`` an optional m x K matrix.
    ``out`` (optional) is an m x K float64 array in column-major order that
    receives the LD scores, e.g. the result of an earlier call reused by a
    loop: large results then land in memory that is already mapped.
    Returns (l2 [m x K], maf [m], L2Perf).
    """
    L = N.lib()
    snps = np.ascontiguousarray(snps, dtype=N.SNP_DTYPE)
    m = snps.size
    a = None
    K = 1
    if annot is not None:
        a = np.asarray(annot, dtype=np.float64)
        if a.ndim == 1:
            a = a[:, None]
        if a.shape[0] != m:
            raise N.LdscError(5, f"Annotation matrix has {a.shape[0]} rows but BIM has {m} SNPs")
        K = a.shape[1]
        a = np.asfortranarray(a)
    if out is None:
        l2 = np.zeros((K, m), dtype=np.float64)
    else:
        if (not isinstance(out, np.ndarray) or out.dtype != np.float64 or out.shape != (m, K)
                or not out.T.flags.c_contiguous or not out.flags.writeable):
            raise ValueError(f"out must be a writeable column-major float64 array of shape "
                             f"({m}, {K})")
        l2 = out.T
    maf = np.zeros(m, dtype=np.float64)
    perf = N.Perf()
    err = N.errbuf()
    cfg = config.to_c()

    if on_progress is not None:
        def _cb(p, _ctx):
            q = p.contents
            on_progress(L2Progress(q.chunks_done, q.chunks_total, q.snps_done, q.snps_total,
                                   q.chunk_wall_ms))
        cb = N.PROGRESS_CB(_cb)
    else:
        cb = N.PROGRESS_CB()
    rc = L.ldsc_l2_compute(panel.handle, N.ptr(snps), m, C.byref(cfg),
                           None if a is None else C.c_void_p(a.ctypes.data), K, N.ptr(l2),
                           N.ptr(maf), C.byref(perf), cb, None, err, 512)
    N.check(rc, err)
    if config.verbose_timing:
        import sys
        print(f"[perf] compute_ldscore: bed_read(stall)={perf.bed_read_secs:.3f}s "
              f"norm={perf.norm_secs:.3f}s bb_dot={perf.bb_dot_secs:.3f}s "
              f"ab_dot={perf.ab_dot_secs:.3f}s ring_store={perf.ring_store_secs:.3f}s",
              file=sys.stderr)
    return (l2.T if out is None else out), maf, L2Perf.from_c(perf, bool(config.sketch))


def compute_l2_from_bed_with_progress(panel: Panel, snps: Sequence[BimRecord], n_indiv: int,
                                      config: L2Config, on_progress) -> L2Output:
    """compute_l2_from_bed_with_progress (src/l2/mod.rs:1123-1169): K = 1."""
    if len(snps) == 0:
        raise ValueError("compute_l2_from_bed: BIM has zero SNPs")
    if n_indiv == 0:
        raise ValueError("compute_l2_from_bed: FAM has zero individuals")
    t0 = time.perf_counter()
    l2, maf, perf = compute_ldscore(panel, snps_to_array(snps), config, None, on_progress)
    return L2Output(list(snps), l2[:, 0].copy(), maf, time.perf_counter() - t0, perf)


def compute_l2_from_bed(panel: Panel, snps: Sequence[BimRecord], n_indiv: int,
                        config: L2Config) -> L2Output:
    """compute_l2_from_bed (src/l2/mod.rs:1099-1106)."""
    return compute_l2_from_bed_with_progress(panel, snps, n_indiv, config, None)


def compute_l2_from_bytes_with_progress(bed_bytes, bim_text: str, fam_text: str,
                                        config: L2Config, on_progress) -> L2Output:
    """compute_l2_from_bytes_with_progress (src/l2/mod.rs:1058-1085)."""
    snps = parse_bim_str(bim_text)
    if not snps:
        raise ValueError("compute_l2_from_bytes: BIM has zero SNPs")
    n_indiv = count_fam_str(fam_text)
    if n_indiv == 0:
        raise ValueError("compute_l2_from_bytes: FAM has zero individuals")
    with Panel(bed_bytes, n_indiv, len(snps)) as panel:
        return compute_l2_from_bed_with_progress(panel, snps, n_indiv, config, on_progress)


def compute_l2_from_bytes(bed_bytes, bim_text: str, fam_text: str,
                          config: L2Config) -> L2Output:
    """compute_l2_from_bytes (src/l2/mod.rs:1049-1056)."""
    return compute_l2_from_bytes_with_progress(bed_bytes, bim_text, fam_text, config, None)


# ----------------------------------------------------------------------------
# Sub-steps
# ----------------------------------------------------------------------------
def block_left(snps: np.ndarray, mode: WindowMode):
    """get_block_lefts_by_chr (src/l2/window.rs:32-90) -> (block_left, any_full)."""
    L = N.lib()
    snps = np.ascontiguousarray(snps, dtype=N.SNP_DTYPE)
    out = np.zeros(snps.size, np.uint64)
    full = C.c_int32()
    err = N.errbuf()
    N.check(L.ldsc_l2_block_left(N.ptr(snps), snps.size, mode.kind, mode.value, N.ptr(out),
                                 C.byref(full), err, 512), err)
    return out, bool(full.value)


def pair_bounds(snps: np.ndarray, config: L2Config) -> np.ndarray:
    """Per-SNP pair bound lo[j] of the chunk/window/mask inclusion rule."""
    L = N.lib()
    snps = np.ascontiguousarray(snps, dtype=N.SNP_DTYPE)
    out = np.zeros(snps.size, np.uint64)
    cfg = config.to_c()
    err = N.errbuf()
    N.check(L.ldsc_l2_pair_bounds(N.ptr(snps), snps.size, C.byref(cfg), N.ptr(out), err, 512),
            err)
    return out


def countsketch_hash(n: int, d: int, seed: int = 42):
    """CountSketchProj::new buckets and signs (src/l2/compute.rs:179-192)."""
    L = N.lib()
    b = np.zeros(n, np.uint32)
    s = np.zeros(n, np.int8)
    err = N.errbuf()
    N.check(L.ldsc_l2_countsketch_hash(n, d, seed, N.ptr(b), N.ptr(s), err, 512), err)
    return b, s


def decode(panel: Panel, bed_idx, dosage=True, miss=True, stats=True, std_f32=True):
    """Device decode of SNP slabs: dosage (N x m int8, -1 missing), missing
    bit masks (ceil(N/32) x m uint32), stats (m x 4: sum, count, sum_sq, het)
    and f32 standardized columns (N x m)."""
    L = N.lib()
    idx = np.ascontiguousarray(bed_idx, dtype=np.uint64)
    m, n = idx.size, panel.n_indiv
    words = (n + 31) // 32
    d = np.zeros((m, n), np.int8) if dosage else None
    mk = np.zeros((m, words), np.uint32) if miss else None
    st = np.zeros((m, 4), np.uint32) if stats else None
    sf = np.zeros((m, n), np.float32) if std_f32 else None
    err = N.errbuf()
    N.check(L.ldsc_l2_decode(panel.handle, N.ptr(idx), m, N.ptr(d), N.ptr(mk), N.ptr(st),
                             N.ptr(sf), err, 512), err)
    return (None if d is None else d.T, None if mk is None else mk.T, st,
            None if sf is None else sf.T)


def sketch_buckets(panel: Panel, bed_idx, d: int):
    """Exact integer CountSketch buckets (sg, sm), each d x m int32."""
    L = N.lib()
    idx = np.ascontiguousarray(bed_idx, dtype=np.uint64)
    m = idx.size
    sg = np.zeros((m, d), np.int32)
    sm = np.zeros((m, d), np.int32)
    err = N.errbuf()
    N.check(L.ldsc_l2_sketch_buckets(panel.handle, N.ptr(idx), m, d, N.ptr(sg), N.ptr(sm), err,
                                     512), err)
    return sg.T, sm.T


def unbiased_r2_constants(n: int):
    """r2_unbiased constants (src/l2/compute.rs:1067-1077)."""
    denom = n - 2.0 if n > 2 else float(n)
    return 1.0 + 1.0 / denom, -1.0 / denom


__all__ = [
    "WindowMode", "L2Config", "BimRecord", "L2Perf", "L2Progress", "L2Output", "Panel",
    "parse_bim_str", "count_fam_str", "snps_to_array", "make_snp_array", "compute_ldscore",
    "compute_l2_from_bed", "compute_l2_from_bed_with_progress", "compute_l2_from_bytes",
    "compute_l2_from_bytes_with_progress", "block_left", "pair_bounds", "countsketch_hash",
    "decode", "sketch_buckets", "unbiased_r2_constants",
]
