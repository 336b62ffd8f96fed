"""ctypes binding of ``include/ldsc_io.h`` (libldsc_io.so): the BIM table,
the ``.annot`` reader and the ``.l2.ldscore.gz`` writer of the ``l2``
command line.

``BimTable`` holds the parsed columns of a ``.bim`` as numpy arrays, with the
SNP IDs left as byte spans of the file's text; ``str`` objects are made only
for the rows that need one (``--extract``, ``--print-snps``).  The library is
loaded from the package directory; when it is missing the first use raises.
"""
from __future__ import annotations

import bz2
import ctypes as C
import gzip
import os
import re
from typing import Iterable, List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libldsc_io.so")
EXPORTED = ["ldsc_io_count_lines", "ldsc_io_parse_bim", "ldsc_io_write_ldscore",
            "ldsc_io_format_fixed3", "ldsc_io_count_nonempty_lines", "ldsc_io_parse_annot"]

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m ldsc_amd.build`")
        L = C.CDLL(LIB_PATH)
        P, U64, I = C.c_void_p, C.c_uint64, C.c_int
        L.ldsc_io_count_lines.argtypes = [P, U64]
        L.ldsc_io_count_lines.restype = U64
        L.ldsc_io_parse_bim.argtypes = [P, U64, U64, P, P, P, P, P, I, C.c_char_p, U64]
        L.ldsc_io_parse_bim.restype = I
        L.ldsc_io_write_ldscore.argtypes = [C.c_char_p, C.c_char_p, U64, P, P, P, U64, P, P, P,
                                            U64, P, U64, I, C.c_char_p, U64]
        L.ldsc_io_write_ldscore.restype = I
        L.ldsc_io_format_fixed3.argtypes = [C.c_double, C.c_char_p, U64]
        L.ldsc_io_format_fixed3.restype = U64
        L.ldsc_io_count_nonempty_lines.argtypes = [P, U64]
        L.ldsc_io_count_nonempty_lines.restype = U64
        L.ldsc_io_parse_annot.argtypes = [P, U64, I, U64, U64, U64, P, I, C.c_char_p, U64]
        L.ldsc_io_parse_annot.restype = I
        _lib = L
    return _lib


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def format_fixed3(v: float) -> str:
    """Rust ``{:.3}`` of one f64 (the value format of write_ldscore_refs)."""
    buf = C.create_string_buffer(400)
    k = lib().ldsc_io_format_fixed3(float(v), buf, 400)
    return buf.raw[:k].decode()


class BimTable:
    """parse_bim_reader_filtered (src/parse.rs:550-592) over a whole file,
    as columns: ``chr`` u8 (0 when not a u8), ``cm`` f64 (0 when not an f64),
    ``bp`` u32 (0 when not a u32), the SNP ID spans, and ``bed_idx`` = line
    index.  Raises ValueError with the reference's message for a line that
    lacks the SNP, A1 or A2 column."""

    def __init__(self, data: bytes, threads: int = 0):
        if isinstance(data, str):
            data = data.encode()
        self.data = bytes(data)
        L = lib()
        buf = C.c_char_p(self.data)
        n = int(L.ldsc_io_count_lines(buf, len(self.data)))
        self.n = n
        self.chr = np.zeros(n, np.uint8)
        self.cm = np.zeros(n, np.float64)
        self.bp = np.zeros(n, np.uint32)
        self.name_off = np.zeros(n, np.uint64)
        self.name_len = np.zeros(n, np.uint32)
        err = C.create_string_buffer(512)
        rc = L.ldsc_io_parse_bim(buf, len(self.data), n, _ptr(self.chr), _ptr(self.cm),
                                 _ptr(self.bp), _ptr(self.name_off), _ptr(self.name_len),
                                 int(threads), err, 512)
        if rc:
            raise ValueError(err.value.decode(errors="replace"))
        self.bed_idx = np.arange(n, dtype=np.int64)

    @classmethod
    def from_file(cls, path: str, threads: int = 0) -> "BimTable":
        with open(path, "rb") as f:
            return cls(f.read(), threads)

    def __len__(self) -> int:
        return self.n

    def name(self, i: int) -> str:
        o = int(self.name_off[i])
        return self.data[o:o + int(self.name_len[i])].decode()

    def names(self, idx: Optional[Iterable[int]] = None) -> List[str]:
        d, off, ln = self.data, self.name_off.tolist(), self.name_len.tolist()
        rows = range(self.n) if idx is None else (int(i) for i in idx)
        return [d[off[i]:off[i] + ln[i]].decode() for i in rows]

    def name_in(self, ids: set, idx: Optional[Sequence[int]] = None) -> np.ndarray:
        """Boolean mask: is the SNP ID of each row (of ``idx``) in ``ids``."""
        enc = {s.encode() for s in ids}
        d, off, ln = self.data, self.name_off.tolist(), self.name_len.tolist()
        rows = range(self.n) if idx is None else [int(i) for i in idx]
        return np.fromiter((d[off[i]:off[i] + ln[i]] in enc for i in rows), dtype=bool,
                           count=len(rows))


def write_ldscore(path: str, table: BimTable, bed_idx: np.ndarray, l2: np.ndarray,
                  rows: np.ndarray, col_names: Sequence[str], threads: int = 0) -> None:
    """write_ldscore_refs (src/l2/io.rs:94-125) through libldsc_io.so.

    ``bed_idx[i]`` is the BIM row of LD-score row ``i`` of ``l2`` (shape
    m x K); ``rows`` lists the LD-score rows to write, in order."""
    l2 = np.asarray(l2, dtype=np.float64)
    if l2.ndim == 1:
        l2 = l2[:, None]
    bed_idx = np.asarray(bed_idx, dtype=np.int64)
    if l2.shape[0] != bed_idx.size:
        raise ValueError(f"write_ldscore: {l2.shape[0]} LD-score rows for {bed_idx.size} SNPs")
    rows = np.asarray(rows, dtype=np.int64)
    if rows.size and (rows.min() < 0 or rows.max() >= bed_idx.size):
        raise ValueError("write_ldscore: row index out of range")
    # only the written rows are gathered (row-major), so a file of one
    # chromosome costs its own rows, not the whole table
    sub = bed_idx[rows]
    vals = np.ascontiguousarray(l2[rows])
    chr_ = np.ascontiguousarray(table.chr[sub])
    bp = np.ascontiguousarray(table.bp[sub])
    off = np.ascontiguousarray(table.name_off[sub])
    ln = np.ascontiguousarray(table.name_len[sub])
    local = np.arange(rows.size, dtype=np.int64)
    tail = "".join("\t" + c for c in col_names).encode()
    err = C.create_string_buffer(512)
    rc = lib().ldsc_io_write_ldscore(os.fsencode(path), tail, rows.size, _ptr(chr_),
                                     _ptr(bp), C.c_char_p(table.data), len(table.data),
                                     _ptr(off), _ptr(ln), _ptr(vals), l2.shape[1], _ptr(local),
                                     rows.size, int(threads), err, 512)
    if rc:
        raise OSError(err.value.decode(errors="replace"))


# `char::is_whitespace`, the separator set of `split_whitespace` and `trim`
_WS = "\t\n\x0b\x0c\r \x85\xa0\u1680\u2000-\u200a\u2028\u2029\u202f\u205f\u3000"
_WS_RUN = re.compile(f"[{_WS}]+")


def read_text_bytes(path: str) -> bytes:
    """open_text_reader (src/frame.rs:420-432): gzip (every member) / bzip2
    by suffix, else the plain file."""
    with open(path, "rb") as f:
        raw = f.read()
    if path.endswith(".gz"):
        return gzip.decompress(raw)
    if path.endswith(".bz2"):
        return bz2.decompress(raw)
    return raw


def parse_annot_text(data: bytes, thin: bool, path: str = "<annot>"):
    """read_tsv (src/frame.rs:319-413) + read_annot_path (src/parse.rs:191-229)
    over decompressed text: the first non-empty line is the header, split on
    tabs when it holds one, else on whitespace runs; the full format skips
    CHR SNP BP CM.  Returns (matrix rows x annotations, names); NA, NaN and
    cells that are not numbers are 0."""
    pos, header = 0, None
    while pos < len(data):
        nl = data.find(b"\n", pos)
        end = len(data) if nl < 0 else nl
        line = data[pos:end].rstrip(b"\r\n")
        pos = end + 1
        if line:
            header = line
            break
    if header is None:
        raise ValueError(f"'{path}' is empty (no header row)")
    try:
        htext = header.decode()
    except UnicodeDecodeError:
        raise ValueError(f"reading line of '{path}': stream did not contain valid UTF-8")
    use_tab = "\t" in htext
    cols = htext.split("\t") if use_tab else [c for c in _WS_RUN.split(htext) if c]
    skip = 0 if thin else 4
    if len(cols) <= skip:
        kind = "thin format" if thin else "full format: CHR SNP BP CM + annotations"
        raise ValueError(f"Annot file '{path}' has {len(cols)} columns; expected > {skip} "
                         f"({kind})")
    L = lib()
    n_rows = int(L.ldsc_io_count_nonempty_lines(data, len(data))) - 1
    out = np.zeros((n_rows, len(cols) - skip), np.float64)
    err = C.create_string_buffer(512)
    rc = L.ldsc_io_parse_annot(data, len(data), int(use_tab), len(cols), skip, n_rows,
                               _ptr(out), 0, err, 512)
    if rc:
        raise ValueError(f"reading line of '{path}': {err.value.decode(errors='replace')}")
    return out, cols[skip:]
