"""libldsc_io.so (include/ldsc_io.h), no GPU: the BIM table agrees with the
line-by-line Python reader on the reference's parse rules, the `{:.3}`
formatter agrees with exact rational rounding, and the gzip writer yields
one gzip member whose text is the reference's table."""
import gzip
import math
import os
import re
import zlib
from fractions import Fraction

import numpy as np
import pytest

import ldsc_amd as L
from ldsc_amd import formats as F

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "ldsc_io.h")


def test_io_library_exports_every_declared_symbol():
    decl = sorted(set(re.findall(r"\b(ldsc_io_\w+)\s*\(", open(HEADER).read())))
    lib = F.lib()
    for name in decl:
        assert hasattr(lib, name), name
    assert set(F.EXPORTED) == set(decl)


def _fixed3_exact(v: float) -> str:
    """{:.3} from the exact binary value: round(v * 1000) half to even."""
    if math.isnan(v):
        return "NaN"
    if math.isinf(v):
        return "inf" if v > 0 else "-inf"
    x = abs(Fraction(v)) * 1000
    q, r = divmod(x.numerator, x.denominator)
    if 2 * r > x.denominator or (2 * r == x.denominator and q % 2 == 1):
        q += 1
    sign = "-" if math.copysign(1.0, v) < 0 else ""
    return f"{sign}{q // 1000}.{q % 1000:03d}"


def test_fixed3_matches_exact_rounding():
    rng = np.random.default_rng(5)
    vals = [0.0, -0.0, 0.0625, 0.1875, 2.0625, -2.0625, 0.0005, -0.0005, 0.0004999, 1e-320,
            -1e-320, 5e-324, 0.9995, 9.9995, 123.4565, 1.0, 1e15, 9.2e15, 9.3e15, 1e17,
            2.0 ** 60, 1.7976931348623157e308, -1.7976931348623157e308, float("nan"),
            float("inf"), float("-inf")]
    # exact ties k/2000 (odd k) that are representable: k * 2^-j with small j
    vals += [k / 16 for k in range(-40, 41)] + [k / 1024 for k in range(0, 2048, 7)]
    vals += list(rng.random(2000) * 10.0 ** rng.integers(-6, 7, 2000))
    vals += list(-rng.random(500) * 10.0 ** rng.integers(-6, 7, 500))
    vals += list(rng.integers(0, 2 ** 40, 200) / 8000.0)  # many sit on or next to ties
    for v in vals:
        assert F.format_fixed3(v) == _fixed3_exact(v), v
        if not math.isnan(v):
            assert F.format_fixed3(v) == "%.3f" % v, v


def _bim_fuzz(rng, n):
    """Lines over the reference's lenient fields and its error cases."""
    chrs = ["1", "22", "+7", "255", "256", "X", "", "-1", "01", "+", "1 ", "٣"]
    cms = ["0", "0.5", "1e-3", "-2.5E+2", ".5", "5.", ".", "nan", "-inf", "Infinity", "1_0",
           " 1", "1e", "abc", "0x10", "NaN", "+.5e-3", "١"]
    bps = ["100", "+5", "0", "4294967295", "4294967296", "-5", "", "1.0", "007", "²"]
    out = []
    for i in range(n):
        cols = [chrs[rng.integers(len(chrs))], f"rs{i}é" if i % 7 == 0 else f"rs{i}",
                cms[rng.integers(len(cms))], bps[rng.integers(len(bps))], "A", "G"]
        if i % 5 == 0:
            cols.append("extra")
        out.append("\t".join(cols))
    return out


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_bim_table_matches_line_reader(threads):
    rng = np.random.default_rng(11)
    lines = _bim_fuzz(rng, 3000)
    for ending in ("\n", "\r\n"):
        text = ending.join(lines) + (ending if threads != 3 else "")
        recs = L.parse_bim_str(text)
        t = F.BimTable(text.encode(), threads=threads)
        assert t.n == len(recs)
        assert t.chr.tolist() == [r.chr for r in recs]
        assert np.array_equal(t.cm, np.array([r.cm for r in recs]), equal_nan=True)
        assert t.bp.tolist() == [r.bp for r in recs]
        assert t.names() == [r.snp for r in recs]
        assert t.bed_idx.tolist() == [r.bed_idx for r in recs]


def test_bim_table_reports_the_first_bad_line():
    good = "1\trs1\t0\t100\tA\tG"
    lines = [good] * 40000
    lines[30001] = "1\trs2\t0\t100\tA"       # no A2
    lines[20000] = "1"                       # no SNP column (reported first)
    text = "\n".join(lines) + "\n"
    for threads in (1, 8):
        with pytest.raises(ValueError, match=r"^BIM line 20001: expected SNP column$"):
            F.BimTable(text, threads=threads)
    with pytest.raises(ValueError, match=r"^BIM line 2: missing A1 col$"):
        F.BimTable(good + "\n1\trs\t0\t1\n")
    with pytest.raises(ValueError, match=r"^BIM line 1: missing A2 col$"):
        F.BimTable("1\trs\t0\t1\tA\n")
    with pytest.raises(ValueError, match=r"^reading BIM line 2: "):
        F.BimTable(good.encode() + b"\n1\trs\xff\t0\t1\tA\tG\n")
    assert F.BimTable(b"").n == 0
    # a final line without '\n' counts, a final '\r' without '\n' stays in the line
    t = F.BimTable(good + "\n2\trs9\t0\t7\tA\tG\r")
    assert t.n == 2 and t.chr.tolist() == [1, 2] and t.bp.tolist() == [100, 7]
    assert F.lib().ldsc_io_count_lines(b"a\n\nb", 4) == 3


def test_name_membership_and_lookup():
    t = F.BimTable("1\trs1\t0\t1\tA\tG\n1\trsé\t0\t2\tA\tG\n2\trs3\t0\t3\tA\tG\n")
    assert t.name(1) == "rsé"
    assert t.name_in({"rsé", "rs3"}).tolist() == [False, True, True]
    assert t.name_in({"rs1"}, [2, 0]).tolist() == [False, True]


def _expected_table(t, bed_idx, l2, rows, cols):
    head = "CHR\tSNP\tBP" + "".join("\t" + c for c in cols) + "\n"
    body = []
    for r in rows:
        b = bed_idx[r]
        body.append(f"{t.chr[b]}\t{t.name(b)}\t{t.bp[b]}" +
                    "".join("\t" + _fixed3_exact(float(v)) for v in l2[r]) + "\n")
    return head + "".join(body)


def _one_member(path):
    raw = open(path, "rb").read()
    d = zlib.decompressobj(wbits=31)
    text = d.decompress(raw) + d.flush()
    assert d.eof and d.unused_data == b"", "not a single gzip member"
    return text.decode()


@pytest.mark.parametrize("m,K,threads", [(10, 1, 1), (300_000, 1, 8), (20_000, 3, 4),
                                         (1_600_000, 1, 8)])
def test_write_ldscore_round_trip(tmp_path, m, K, threads):
    rng = np.random.default_rng(m + K)
    text = "".join(f"{1 + i * 3 // m}\trs{i}\t0\t{i * 10}\tA\tG\n" for i in range(m))
    t = F.BimTable(text)
    bed_idx = np.sort(rng.choice(m, size=m * 3 // 4, replace=False))
    l2 = rng.random((bed_idx.size, K)) * 10.0 ** rng.integers(-4, 4, (bed_idx.size, K))
    l2[::97] = -l2[::97]
    l2[5 % bed_idx.size, 0] = float("nan")
    rows = np.flatnonzero(rng.random(bed_idx.size) < 0.9)
    cols = [f"A{k}L2" for k in range(K)]
    path = str(tmp_path / "o.l2.ldscore.gz")
    F.write_ldscore(path, t, bed_idx, l2, rows, cols, threads=threads)
    got = _one_member(path)
    assert got == _expected_table(t, bed_idx, l2, rows, cols)
    assert gzip.open(path, "rt").read() == got


def test_write_ldscore_empty_and_errors(tmp_path):
    t = F.BimTable("1\trs1\t0\t1\tA\tG\n")
    path = str(tmp_path / "e.gz")
    F.write_ldscore(path, t, np.array([0]), np.array([[1.0]]), np.array([], np.int64), ["L2"])
    assert _one_member(path) == "CHR\tSNP\tBP\tL2\n"
    with pytest.raises(ValueError):
        F.write_ldscore(path, t, np.array([0]), np.array([[1.0]]), np.array([1]), ["L2"])
    with pytest.raises(OSError, match="creating output"):
        F.write_ldscore(str(tmp_path / "no" / "x.gz"), t, np.array([0]), np.array([[1.0]]),
                        np.array([0]), ["L2"])


_RUST_WS = ("\t\n\x0b\x0c\r \x85\xa0\u1680" + "".join(chr(c) for c in range(0x2000, 0x200B)) +
            "\u2028\u2029\u202f\u205f\u3000")


def _rust_f64(s):
    """Rust f64::from_str on an already-trimmed ASCII token, else None."""
    if not s or not s.isascii() or "_" in s or s != s.strip():
        return None
    try:
        return float(s)
    except ValueError:
        return None


def _annot_by_lines(text, thin):
    """read_tsv (src/frame.rs:319-413) + read_annot_path (src/parse.rs:191-229)
    line by line: non-empty lines after removing trailing CR/LF, tab split
    when the header holds a tab, else runs of Rust whitespace; pad / cut to
    the header width; trimmed cells; NA/NaN/./empty/non-numbers -> 0."""
    lines = [ln.rstrip("\r\n") for ln in text.split("\n")]
    lines = [ln for ln in lines if ln]
    use_tab = "\t" in lines[0]

    def split(s):
        if use_tab:
            return s.split("\t")
        return [c for c in re.split(f"[{_RUST_WS}]+", s) if c]

    header = split(lines[0])
    skip = 0 if thin else 4
    rows = []
    for ln in lines[1:]:
        c = split(ln)[:len(header)]
        c += [""] * (len(header) - len(c))
        vals = []
        for cell in c[skip:]:
            t = cell.strip(_RUST_WS)
            if t.lower() in ("", "na", "nan", "."):
                vals.append(0.0)
                continue
            v = _rust_f64(t)
            vals.append(0.0 if v is None or v != v else v)
        rows.append(vals)
    return np.array(rows, dtype=np.float64).reshape(len(rows), len(header) - skip), header[skip:]


def _annot_fuzz(rng, n, use_tab, width):
    toks = ["0", "1", "0.25", "-3e-2", "NA", "na", "nA", "NaN", "nan", "NAN", ".", "", "1_0",
            "inf", "-Infinity", " 2.5 ", "\xa03\u3000", "x", "1e400", "+.5", "5.", "١"]
    sep = "\t" if use_tab else " "
    head = sep.join((["CHR", "SNP", "BP", "CM"] if width > 4 else []) +
                    [f"A{k}" for k in range(width - (4 if width > 4 else 0))])
    out = [head]
    for i in range(n):
        k = width + int(rng.integers(-2, 3))
        cells = [toks[rng.integers(len(toks))] for _ in range(max(k, 0))]
        if not use_tab:
            cells = [c.strip(_RUST_WS) or "NA" for c in cells]
            seps = [[" ", "  ", "\t", "\xa0", "  "][rng.integers(5)] for _ in cells]
            out.append("".join(c + s for c, s in zip(cells, seps)).rstrip(_RUST_WS))
        else:
            out.append(sep.join(cells))
        if i % 37 == 0:
            out.append("")
    return "\r\n".join(out) + "\n\n"


@pytest.mark.parametrize("use_tab,thin", [(True, True), (True, False), (False, True),
                                          (False, False)])
def test_annot_parser_matches_line_reader(use_tab, thin):
    rng = np.random.default_rng(7 + 2 * use_tab + thin)
    text = _annot_fuzz(rng, 4000, use_tab, 7)
    got, names = F.parse_annot_text(text.encode(), thin)
    exp, enames = _annot_by_lines(text, thin)
    assert names == enames
    assert got.shape == exp.shape
    assert np.array_equal(got, exp)


def test_annot_known_answers(tmp_path):
    text = ("CHR\tSNP\tBP\tCM\tA\tB\n"
            "1\trs1\t1\t0\t1\tNA\n"
            "\r\n"
            "1\trs2\t2\t0\t 0.5 \t1_0\n"
            "1\trs3\t3\t0\tinf\n"
            "1\trs4\t4\t0\tnan\t2\textra\n")
    m, names = F.parse_annot_text(text.encode(), thin=False)
    assert names == ["A", "B"]
    assert m.tolist() == [[1.0, 0.0], [0.5, 0.0], [float("inf"), 0.0], [0.0, 2.0]]
    m, names = F.parse_annot_text(b"X Y\n1\xc2\xa0 2\n", thin=True)
    assert names == ["X", "Y"] and m.tolist() == [[1.0, 2.0]]
    m, names = F.parse_annot_text(b"\n\nX\tY\n", thin=True)
    assert m.shape == (0, 2)
    with pytest.raises(ValueError, match="empty"):
        F.parse_annot_text(b"\r\n\n", thin=True)
    with pytest.raises(ValueError, match="has 2 columns"):
        F.parse_annot_text(b"A\tB\n1\t2\n", thin=False)
    with pytest.raises(ValueError, match="UTF-8"):
        F.parse_annot_text(b"A\tB\n1\t\xff\n", thin=True)
    # the compressed forms read through the same path
    import bz2 as _bz2
    (tmp_path / "a.annot.gz").write_bytes(gzip.compress(text.encode()))
    (tmp_path / "a.annot.bz2").write_bytes(_bz2.compress(text.encode()))
    for f in ("a.annot.gz", "a.annot.bz2"):
        raw = F.read_text_bytes(str(tmp_path / f))
        assert raw == text.encode()


def test_write_ldscore_c_abi_checks_its_arrays(tmp_path):
    """The C entry point rejects row indices and SNP ID spans outside the
    arrays it is given, before it creates the file."""
    import ctypes as C

    chr_ = np.array([1, 2], np.uint8)
    bp = np.array([10, 20], np.uint32)
    names = b"rs1rs2"
    off = np.array([0, 3], np.uint64)
    ln = np.array([3, 3], np.uint32)
    l2 = np.array([1.0, 2.0])
    err = C.create_string_buffer(256)
    path = tmp_path / "x.gz"

    def call(rows, names_len=len(names)):
        rows = np.array(rows, np.int64)
        return F.lib().ldsc_io_write_ldscore(
            os.fsencode(str(path)), b"\tL2", 2, chr_.ctypes.data, bp.ctypes.data, names,
            names_len, off.ctypes.data, ln.ctypes.data, l2.ctypes.data, 1, rows.ctypes.data,
            rows.size, 1, err, 256)

    assert call([2]) == 1 and b"row index out of range" in err.value and not path.exists()
    assert call([-1]) == 1 and not path.exists()
    assert call([1], names_len=5) == 1 and b"span" in err.value and not path.exists()
    assert call([1, 0]) == 0
    assert _one_member(str(path)) == "CHR\tSNP\tBP\tL2\n2\trs2\t20\t2.000\n1\trs1\t10\t1.000\n"
