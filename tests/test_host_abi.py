"""C-ABI library checks that need no GPU: every symbol of include/ldsc_l2.h
resolves, and the host-side index work (windows, pair bounds, CountSketch
hash, BED validation, BIM/FAM parsing) agrees with the oracle / reference."""
import os
import re

import numpy as np
import pytest

import ldsc_amd as L
from ldsc_amd import _native as N

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "ldsc_l2.h")


def test_library_exports_every_declared_symbol():
    decl = re.findall(r"\b(ldsc_\w+)\s*\(", open(HEADER).read())
    decl = sorted(set(d for d in decl if not d.endswith("_cb")))
    lib = N.lib()
    for name in decl:
        assert hasattr(lib, name), name
    assert set(N.EXPORTED) <= set(decl)


def _rand_snps(rng, m, nchr):
    chr_ = np.sort(rng.integers(1, nchr + 1, size=m)).astype(np.uint8)
    bp = np.cumsum(rng.integers(1, 5000, size=m)).astype(np.uint32)
    return chr_, bp * 1e-6, bp


@pytest.mark.parametrize("mode,w", [(0, 0.004), (1, 3.0), (2, 5), (0, 0.0), (1, 1e9), (2, 0)])
def test_block_left_matches_oracle(oracle, mode, w):
    rng = np.random.default_rng(int(w * 1000) + mode)
    for nchr in (1, 3):
        chr_, cm, bp = _rand_snps(rng, 500, nchr)
        snps = L.make_snp_array(chr_, cm, bp)
        kinds = [L.WindowMode.Cm, L.WindowMode.Kb, L.WindowMode.Snp]
        got, full = L.block_left(snps, kinds[mode](w))
        exp, efull = oracle.block_left(chr_, cm, bp, mode, w)
        assert np.array_equal(got, exp) and full == efull


def test_block_left_unsorted_coordinates(oracle):
    """get_block_lefts_f64 uses |c_i - c_j| with a forward-only pointer; it
    is defined (and must match) for non-monotone coordinates too."""
    rng = np.random.default_rng(3)
    bp = rng.integers(1, 100000, size=300).astype(np.uint32)
    chr_ = np.ones(300, np.uint8)
    got, _ = L.block_left(L.make_snp_array(chr_, bp * 1e-6, bp), L.WindowMode.Kb(10.0))
    exp, _ = oracle.block_left(chr_, bp * 1e-6, bp, 1, 10.0)
    assert np.array_equal(got, exp)


def test_block_left_large_split_scan(oracle):
    """Lists long enough for the host to split the scan into tasks: runs
    of several chromosomes, tied coordinates, one run with a position out of
    order and one NaN cM entry (those runs scan from their start), against
    the single forward scan of the restatement."""
    rng = np.random.default_rng(11)
    m = 200_000
    chr_ = np.repeat(np.arange(1, 6, dtype=np.uint8), m // 5)
    bp = np.cumsum(rng.integers(0, 3000, size=m)).astype(np.uint32)  # gaps of 0 tie
    cm = bp * 1e-6
    bp_bad = bp.copy()
    bp_bad[50_000] = bp_bad[49_000]  # chromosome 2 out of order (kb mode)
    cm_bad = cm.copy()
    cm_bad[130_000] = np.nan         # chromosome 4 with a NaN (cM mode)
    cases = [(0, 1.0, cm, bp), (1, 40.0, cm, bp_bad), (0, 0.5, cm_bad, bp), (2, 700, cm, bp)]
    kinds = [L.WindowMode.Cm, L.WindowMode.Kb, L.WindowMode.Snp]
    for mode, w, c, b in cases:
        snps = L.make_snp_array(chr_, c, b)
        got, full = L.block_left(snps, kinds[mode](w))
        exp, efull = oracle.block_left(chr_, c, b, mode, w)
        assert np.array_equal(got, exp) and full == efull, (mode, w)
        for per_chr in (False, True):
            cfg = L.L2Config(mode=kinds[mode](w), chunk_size=200, per_chromosome=per_chr)
            assert np.array_equal(L.pair_bounds(snps, cfg),
                                  _a_left_rule(exp, m, chr_, 200, per_chr)), (mode, per_chr)


@pytest.mark.parametrize("seed", [3, 4, 5])
def test_block_left_split_scan_edges(oracle, seed):
    """The split scan's edges against the restatement: hundreds of runs of
    1-8 SNPs (shorter than the four-wide test), chromosome codes that come
    back (a new run each time), gaps far wider than the window (the scan
    jumps many SNPs at once), heavy ties at W = 0, an infinite coordinate
    closing a run, and out-of-order positions scattered over the list."""
    rng = np.random.default_rng(seed)
    lens = list(rng.integers(1, 9, size=400)) + [60_000, 7, 90_000, 3, 80_000]
    chr_ = np.concatenate([np.full(n, 1 + (k % 3), np.uint8) for k, n in enumerate(lens)])
    m = chr_.size
    gaps = np.where(rng.random(m) < 0.02, rng.integers(50_000, 400_000, size=m),
                    rng.integers(0, 4, size=m) * rng.integers(0, 900, size=m))
    bp = np.cumsum(gaps).astype(np.uint32)
    cm = bp * 1e-6
    cm_inf = cm.copy()
    ends = np.cumsum(lens) - 1
    cm_inf[ends[400]] = np.inf
    bp_bad = bp.copy()
    for k in rng.choice(np.arange(1, m), size=25, replace=False):
        bp_bad[k] = bp_bad[k - 1] // 2
    kinds = [L.WindowMode.Cm, L.WindowMode.Kb, L.WindowMode.Snp]
    cases = [(1, 5.0, cm, bp), (1, 0.0, cm, bp), (0, 0.2, cm_inf, bp), (1, 30.0, cm, bp_bad),
             (0, 0.0, cm, bp)]
    for mode, w, c, b in cases:
        snps = L.make_snp_array(chr_, c, b)
        got, full = L.block_left(snps, kinds[mode](w))
        exp, efull = oracle.block_left(chr_, c, b, mode, w)
        assert np.array_equal(got, exp) and full == efull, (mode, w)
        cfg = L.L2Config(mode=kinds[mode](w), chunk_size=200, per_chromosome=True)
        assert np.array_equal(L.pair_bounds(snps, cfg), _a_left_rule(exp, m, chr_, 200, True))


def _a_left_rule(bl, m, chr_, chunk, per_chr):
    lo = np.zeros(m, np.uint64)
    run0 = 0
    while run0 < m:
        run1 = m
        if per_chr:
            run1 = run0 + 1
            while run1 < m and chr_[run1] == chr_[run0]:
                run1 += 1
        for s in range(run0, run1, chunk):
            dist = s - int(bl[s])
            bs = 0 if dist == 0 else -(-dist // chunk) * chunk
            a_left = s - min(bs, s - run0)
            lo[s:min(s + chunk, run1)] = a_left
        run0 = run1
    return lo


@pytest.mark.parametrize("chunk", [1, 3, 50, 200])
@pytest.mark.parametrize("per_chr", [False, True])
def test_pair_bounds_rule(oracle, chunk, per_chr):
    rng = np.random.default_rng(chunk)
    chr_, cm, bp = _rand_snps(rng, 700, 3)
    snps = L.make_snp_array(chr_, cm, bp)
    bl, _ = oracle.block_left(chr_, cm, bp, 1, 40.0)
    cfg = L.L2Config(mode=L.WindowMode.Kb(40.0), chunk_size=chunk, per_chromosome=per_chr)
    assert np.array_equal(L.pair_bounds(snps, cfg), _a_left_rule(bl, 700, chr_, chunk, per_chr))
    cfg.snp_level_masking = True
    assert np.array_equal(L.pair_bounds(snps, cfg), bl)


def test_pair_bounds_errors():
    snps = L.make_snp_array([1, 1], [0, 0], [1, 2])
    with pytest.raises(L.LdscError) as e:
        L.pair_bounds(snps, L.L2Config(chunk_size=0))
    assert "chunk_size must be >= 1" in str(e.value)


def test_countsketch_hash_matches_oracle(oracle):
    for n, d in ((1000, 200), (4097, 5000), (50, 3), (100000, 1000)):
        b, s = L.countsketch_hash(n, d)
        eb, es = oracle.countsketch_hash(n, d)
        assert np.array_equal(b, eb) and np.array_equal(s, es)


def _bed(n, m, fill=0):
    bps = (n + 3) // 4
    return bytes([0x6C, 0x1B, 0x01]) + bytes([fill]) * (bps * m)


def test_open_panel_rejects_bad_bed():
    """Bed::from_source (src/bed.rs:76-109), compute_l2_from_bytes
    (src/l2/mod.rs:1229-1246): header and length are checked before any
    device work."""
    with pytest.raises(L.LdscError) as e:
        L.Panel(b"\xff" + _bed(4, 2)[1:], 4, 2)
    assert e.value.code == 2 and "magic" in str(e.value)
    with pytest.raises(L.LdscError) as e:
        L.Panel(_bed(4, 2)[:-1], 4, 2)
    assert e.value.code == 2 and "invalid length" in str(e.value)
    with pytest.raises(L.LdscError) as e:
        L.Panel(bytes([0x6C, 0x1B, 0x01, 0xF0]), 4, 2)
    assert "invalid length" in str(e.value)
    with pytest.raises(L.LdscError) as e:
        L.Panel(bytes([0x6C, 0x1B, 0x00, 0xF0, 0xF0]), 4, 2)
    assert "SNP-major" in str(e.value)
    with pytest.raises(L.LdscError) as e:
        L.Panel(b"\x6c\x1b", 4, 0)
    assert "too short" in str(e.value)
    with pytest.raises(L.LdscError) as e:
        L.Panel(_bed(4, 0), 4, 0)
    assert e.value.code == 3
    with pytest.raises(L.LdscError) as e:
        L.Panel(_bed(4, 2), 4, 2, keep=[0, 7])
    assert e.value.code == 1


def test_truncated_bed_through_bytes_api():
    """compute_l2_from_bytes_rejects_truncated_bed (src/l2/mod.rs:1229-1246)."""
    bim = "1\trs1\t0\t100\tA\tG\n1\trs2\t0\t200\tA\tG\n"
    fam = "".join(f"F{i}\tI{i}\t0\t0\t1\t-9\n" for i in range(1, 5))
    with pytest.raises(L.LdscError):
        L.compute_l2_from_bytes(bytes([0x6C, 0x1B, 0x01, 0xF0]), bim, fam,
                                L.L2Config(yes_really=True))


def test_parse_bim_str_reference_case():
    """parse_bim_str_matches_file (src/parse.rs:602-615)."""
    bim = "1\trs1\t0\t100\tA\tG\n1\trs2\t0.1\t200\tC\tT\n2\trs3\t0\t300\tG\tA\n"
    recs = L.parse_bim_str(bim)
    assert len(recs) == 3
    assert recs[0].snp == "rs1" and recs[0].chr == 1 and recs[0].bp == 100
    assert recs[1].cm == 0.1 and recs[2].chr == 2 and recs[2].bed_idx == 2


def test_parse_bim_lenient_fields_and_errors():
    recs = L.parse_bim_str("X\trs1\tnan?\t-5\tA\tG\r\n300\trs2\t1e-3\t7\tA\tG")
    assert recs[0].chr == 0 and recs[0].cm == 0.0 and recs[0].bp == 0
    assert recs[1].chr == 0 and recs[1].cm == 1e-3 and recs[1].bp == 7
    with pytest.raises(ValueError):
        L.parse_bim_str("1\trs1\t0\t100\tA\n")
    assert L.count_fam_str("a\nb\n") == 2 and L.count_fam_str("a\nb") == 2
    assert L.count_fam_str("") == 0
