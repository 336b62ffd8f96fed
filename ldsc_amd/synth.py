"""Seeded synthetic PLINK panels for the LD-score workloads (libldsc_synth.so).

Founder-haplotype Markov generator: 2 haplotypes per individual copy one of
20 founders and switch with probability 0.05 per SNP; founder alleles are
Bernoulli(MAF_j), MAF_j ~ U(0.01, 0.5); iid missingness; bp gaps ~ Exp(2 kb),
cM = bp * 1e-6, one chromosome.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libldsc_synth.so")
_lib = None

_f64p = np.ctypeslib.ndpointer(np.float64, flags="C")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C")


def _L():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise ImportError(f"{_LIB} is not built; run `python -m ldsc_amd.build`")
        L = C.CDLL(_LIB)
        L.synth_snps.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, _f64p, _u32p, _f64p, _u32p]
        L.synth_bed.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_uint32, C.c_double,
                                C.c_double, _u32p, _u8p]
        L.synth_annot.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, C.c_double,
                                  _f64p]
        _lib = L
    return _lib


@dataclass
class SynthPanel:
    n: int
    m: int
    buf: np.ndarray      # the whole .bed file: 3-byte header + body
    chr: np.ndarray
    bp: np.ndarray
    cm: np.ndarray
    maf_true: np.ndarray

    @property
    def body(self) -> np.ndarray:
        """m x ceil(n/4) packed SNP rows (a view into ``buf``)."""
        return self.buf[3:].reshape(self.m, (self.n + 3) // 4)

    def bed_bytes(self) -> np.ndarray:
        return self.buf


def make_panel(n: int, m: int, seed: int, miss_rate: float = 0.01, n_founders: int = 20,
               switch_p: float = 0.05) -> SynthPanel:
    L = _L()
    maf = np.zeros(m, np.float64)
    bp = np.zeros(m, np.uint32)
    cm = np.zeros(m, np.float64)
    fnd = np.zeros(m, np.uint32)
    L.synth_snps(seed, m, n_founders, maf, bp, cm, fnd)
    bps = (n + 3) // 4
    buf = np.empty(m * bps + 3, np.uint8)
    buf[:3] = (0x6C, 0x1B, 0x01)
    body = buf[3:]
    body[:] = 0
    L.synth_bed(seed, n, m, n_founders, switch_p, miss_rate, fnd, body)
    return SynthPanel(n, m, buf, np.ones(m, np.uint8), bp, cm, maf)


def make_annot(m: int, seed: int, n_bin: int = 40, n_cont: int = 10,
               run_mean: float = 50.0) -> np.ndarray:
    """m x (n_bin + n_cont): binary columns in contiguous runs (density
    ~ U(0.005, 0.2)) followed by LogNormal(0, 1) columns."""
    out = np.zeros((n_bin + n_cont) * m, np.float64)
    _L().synth_annot(seed, m, n_bin, n_cont, run_mean, out)
    return out.reshape(n_bin + n_cont, m).T


def write_plink(prefix: str, panel: SynthPanel) -> None:
    """Write {prefix}.bed/.bim/.fam (tab-separated BIM)."""
    panel.bed_bytes().tofile(prefix + ".bed")
    with open(prefix + ".bim", "w") as f:
        for j in range(panel.m):
            f.write(f"{int(panel.chr[j])}\trs{j + 1}\t{float(panel.cm[j])!r}\t{int(panel.bp[j])}\tA\tG\n")
    with open(prefix + ".fam", "w") as f:
        for i in range(panel.n):
            f.write(f"F{i}\tI{i}\t0\t0\t1\t-9\n")
