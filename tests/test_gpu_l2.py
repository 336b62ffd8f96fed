"""Device LD scores against the CPU restatement (both arithmetic forms)."""
import numpy as np
import pytest

import ldsc_amd as L
from ldsc_amd import synth

pytestmark = pytest.mark.gpu

# The exact form's integer products (e2m1 or int8 MFMA) are exact and
# finished in FP64: agreement with the FP64 restatement is to rounding.  The
# FP32 form is compared with the f32 restatement of --fast-f32 (different
# summation order).
TOL_EXACT = 1e-9
TOL_F32 = 2e-3


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b))))


def test_reference_smoke_case():
    """compute_l2_from_bytes_smoke (src/l2/mod.rs:1191-1225): l2 = [2, 2]."""
    bed = bytes([0x6C, 0x1B, 0x01, 0xF0, 0xF0])
    bim = "1\trs1\t0\t100\tA\tG\n1\trs2\t0\t200\tA\tG\n"
    fam = "".join(f"F{i}\tI{i}\t0\t0\t1\t-9\n" for i in range(1, 5))
    for use_f32 in (False, True):
        cfg = L.L2Config(mode=L.WindowMode.Kb(100.0), chunk_size=2, yes_really=True,
                         use_f32=use_f32)
        out = L.compute_l2_from_bytes(bed, bim, fam, cfg)
        assert np.allclose(out.l2, [2.0, 2.0], atol=1e-9)
        assert np.allclose(out.maf, [0.5, 0.5], atol=1e-9)


@pytest.mark.parametrize("mode,window,masking,chunk", [
    ("cm", 0.3, True, 50), ("cm", 0.3, False, 50), ("kb", 250.0, False, 37),
    ("snp", 90, True, 64), ("snp", 90, False, 200),
])
@pytest.mark.parametrize("form", ["exact", "exact_i8", "f32"])
def test_l2_matches_oracle(oracle, mode, window, masking, chunk, form):
    p = synth.make_panel(n=517, m=1500, seed=11, miss_rate=0.02)
    md = {"cm": L.WindowMode.Cm, "kb": L.WindowMode.Kb, "snp": L.WindowMode.Snp}[mode](window)
    om = {"cm": oracle.WINDOW_CM, "kb": oracle.WINDOW_KB, "snp": oracle.WINDOW_SNP}[mode]
    snps = L.make_snp_array(p.chr, p.cm, p.bp)
    use_f32 = form == "f32"
    cfg = L.L2Config(mode=md, chunk_size=chunk, use_f32=use_f32, snp_level_masking=masking,
                     exact_kernel="i8" if form == "exact_i8" else "auto")
    print(f"case mode={mode} masking={masking} chunk={chunk} form={form}", flush=True)
    with L.Panel(p.bed_bytes(), p.n, p.m) as panel:
        got, maf, perf = L.compute_ldscore(panel, snps, cfg)
    exp, emaf = oracle.l2(p.bed_bytes(), p.n, p.chr, p.cm, p.bp, mode=om, window=window,
                          chunk=chunk, use_f32=use_f32, masking=masking)
    assert _rel(got, exp) < (TOL_F32 if use_f32 else TOL_EXACT)
    assert np.array_equal(maf, emaf)


@pytest.mark.parametrize("form", ["exact", "exact_i8", "f32"])
def test_cpu_runnable_config(oracle, form):
    """The reference's CPU-runnable case: founder-haplotype panel N=2,490 x
    M=20,000, 0.5 % missing, --ld-wind-cm 1, chunk 50 (--python-compat), no
    SNP-level masking, against the restatement of the f64 path (and of
    --fast-f32 for the FP32 form)."""
    p = synth.make_panel(2490, 20000, seed=20261204, miss_rate=0.005)
    snps = L.make_snp_array(p.chr, p.cm, p.bp)
    use_f32 = form == "f32"
    cfg = L.L2Config(mode=L.WindowMode.Cm(1.0), chunk_size=50, use_f32=use_f32,
                     exact_kernel="i8" if form == "exact_i8" else "auto")
    print(f"case N=2490 M=20000 form={form}", flush=True)
    with L.Panel(p.bed_bytes(), p.n, p.m) as panel:
        got, maf, perf = L.compute_ldscore(panel, snps, cfg)
    exp, emaf = oracle.l2(p.bed_bytes(), p.n, p.chr, p.cm, p.bp, mode=oracle.WINDOW_CM,
                          window=1.0, chunk=50, use_f32=use_f32, masking=False, threads=16)
    assert _rel(got, exp) < (TOL_F32 if use_f32 else TOL_EXACT)
    assert np.array_equal(maf, emaf)
    assert perf.pairs_included > 9_000_000


def _chunk_total(chr_, chunk, per_chr):
    if not per_chr:
        return -(-len(chr_) // chunk)
    tot, s = 0, 0
    while s < len(chr_):
        e = s + 1
        while e < len(chr_) and chr_[e] == chr_[s]:
            e += 1
        tot += -(-(e - s) // chunk)
        s = e
    return tot


@pytest.mark.parametrize("per_chr,K", [(False, 1), (True, 1), (False, 220)])
def test_progress_callbacks(per_chr, K):
    """L2Progress (mod.rs:22-39, called per chunk at compute.rs:2391-2398):
    progress arrives while the band runs, chunks_done is monotone over the
    reference's chunk grid (restarted per chromosome when per_chromosome),
    the last call reports every chunk and every SNP, and the results are
    bit-identical to a call without a callback.  K = 220 takes the segmented
    annotation upload (m x K x 8 bytes >= 32 MiB)."""
    m, chunk = 20000, 50
    p = synth.make_panel(n=700, m=m, seed=3, miss_rate=0.01)
    if per_chr:
        p.chr[:] = 1 + (np.arange(m) * 3) // m
    snps = L.make_snp_array(p.chr, p.cm, p.bp)
    rng = np.random.default_rng(1)
    annot = rng.random((m, K)) if K > 1 else None
    cfg = L.L2Config(mode=L.WindowMode.Cm(1.0), chunk_size=chunk, per_chromosome=per_chr)
    print(f"progress per_chr={per_chr} K={K}", flush=True)
    calls = []
    with L.Panel(p.bed_bytes(), p.n, p.m) as panel:
        ref, _, _ = L.compute_ldscore(panel, snps, cfg, annot)
        got, _, _ = L.compute_ldscore(panel, snps, cfg, annot, on_progress=calls.append)
    assert np.array_equal(got, ref)
    total = _chunk_total(p.chr, chunk, per_chr)
    assert 2 <= len(calls) <= 9
    assert all(c.chunks_total == total and c.snps_total == m for c in calls)
    done = [c.chunks_done for c in calls]
    assert all(b > a for a, b in zip(done, done[1:])) and done[0] >= 1
    assert calls[-1].chunks_done == total and calls[-1].snps_done == m
    assert all(0 < c.snps_done < m for c in calls[:-1])
    assert all(c.chunk_wall_ms >= 0 for c in calls)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 63])
@pytest.mark.parametrize("m", [1, 2, 40])
@pytest.mark.parametrize("form", ["exact", "exact_i8", "f32"])
def test_l2_tiny_panels(oracle, n, m, form):
    """The degenerate sizes: N = 1 (every column monomorphic, so every pair
    adds r2u = -1/D with D = N), N = 2 (D = N) and N = 3 (D = 1), a single
    SNP and two SNPs, against the restatement (compute.rs:1067-1077)."""
    p = synth.make_panel(n=n, m=m, seed=100 * n + m, miss_rate=0.05)
    snps = L.make_snp_array(p.chr, p.cm, p.bp)
    use_f32 = form == "f32"
    print(f"tiny n={n} m={m} form={form}", flush=True)
    with L.Panel(p.bed_bytes(), p.n, p.m) as panel:
        for md, om, window, masking, chunk in ((L.WindowMode.Snp(3), oracle.WINDOW_SNP, 3, False, 2),
                                               (L.WindowMode.Cm(0.05), oracle.WINDOW_CM, 0.05, True, 7)):
            cfg = L.L2Config(mode=md, chunk_size=chunk, use_f32=use_f32, snp_level_masking=masking,
                             exact_kernel="i8" if form == "exact_i8" else "auto", yes_really=True)
            got, maf, _ = L.compute_ldscore(panel, snps, cfg)
            exp, emaf = oracle.l2(p.bed_bytes(), p.n, p.chr, p.cm, p.bp, mode=om, window=window,
                                  chunk=chunk, use_f32=use_f32, masking=masking)
            assert _rel(got, exp) < (TOL_F32 if use_f32 else TOL_EXACT)
            assert np.array_equal(maf, emaf)


@pytest.mark.parametrize("n,m,d", [(4, 30, 3), (300, 60, 299), (2000, 120, 1999)])
def test_l2_sketch_extreme_d(oracle, n, m, d):
    """--sketch at its bounds 3 <= d < N (compute.rs:1083-1089): the smallest
    panel that admits a sketch, and d = N - 1 (buckets of about one member),
    against the f32 sequential-scatter restatement."""
    p = synth.make_panel(n=n, m=m, seed=d, miss_rate=0.02)
    snps = L.make_snp_array(p.chr, p.cm, p.bp)
    cfg = L.L2Config(mode=L.WindowMode.Cm(0.05), chunk_size=7, sketch=d, snp_level_masking=True)
    print(f"sketch n={n} d={d}", flush=True)
    with L.Panel(p.bed_bytes(), p.n, p.m) as panel:
        got, maf, _ = L.compute_ldscore(panel, snps, cfg)
    exp, emaf = oracle.l2(p.bed_bytes(), p.n, p.chr, p.cm, p.bp, mode=oracle.WINDOW_CM,
                          window=0.05, chunk=7, sketch=d, use_f32=True, masking=True)
    assert _rel(got, exp) < 2e-5
    assert np.array_equal(maf, emaf)
