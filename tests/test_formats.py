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


@pytest.mark.parametrize("m,K,threads", [(10, 1, 1), (300_000, 1, 8), (20_000, 3, 4)])
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
