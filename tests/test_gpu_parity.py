"""Device results against the CPU restatement beyond the basic window modes:
bit-exact decode / stats / f32 standardization / integer sketch buckets, and
l2 with --keep, --extract, annotations, pq, per-chromosome chunking, special
columns and the CountSketch path."""
import numpy as np
import pytest

import ldsc_amd as L
from ldsc_amd import synth

pytestmark = pytest.mark.gpu

TOL_EXACT = 1e-9     # exact integer products + FP64 finish vs FP64 restatement (relative)
TOL_F32 = 2e-3       # FP32 form vs f32 restatement (relative, floor 1)
TOL_SKETCH = 2e-5    # exact-integer sketch vs f32 sequential scatter (relative, floor 1)


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b))))


def _panel(n, m, seed, miss=0.02, nchr=1):
    p = synth.make_panel(n=n, m=m, seed=seed, miss_rate=miss)
    if nchr > 1:
        p.chr[:] = 1 + (np.arange(m) * nchr) // m
    return p


@pytest.mark.parametrize("n", [1, 3, 4, 5, 127, 128, 129, 1000])
def test_decode_bit_exact(oracle, n):
    print(f"decode n={n}", flush=True)
    p = _panel(n, 70, seed=n)
    idx = np.array([5, 0, 69, 33, 33, 12], np.uint64)
    keep_cases = [None]
    if n >= 4:
        keep_cases.append(np.array(sorted(set(range(n)) - {1, n // 2}))[::-1].copy())
    for keep in keep_cases:
        with L.Panel(p.bed_bytes(), n, p.m, keep=keep) as panel:
            dos, miss, st, sf = L.decode(panel, idx)
        edos = oracle.decode(p.bed_bytes(), n, idx, keep=keep)
        assert np.array_equal(dos, edos)
        assert np.array_equal(st, oracle.stats(p.bed_bytes(), n, idx, keep=keep))
        assert np.array_equal(sf.view(np.uint32),
                              oracle.std_f32(p.bed_bytes(), n, idx, keep=keep).view(np.uint32))
        nn = edos.shape[0]
        bits = (edos == -1)
        words = (nn + 31) // 32
        exp = np.zeros((words, len(idx)), np.uint32)
        for i in range(nn):
            exp[i // 32] |= (bits[i].astype(np.uint32) << np.uint32(i % 32))
        assert np.array_equal(miss, exp)


@pytest.mark.parametrize("n,m", [(2490, 450_000), (1024, 1_100_000)])
def test_panel_upload_across_pieces(oracle, n, m):
    """A BED body above the 256 MiB upload piece goes up in several
    contiguous pieces; rows of 623 bytes (restrided to 624 on the device) and
    rows of 256 bytes (already aligned) decode exactly on both sides of every
    piece boundary."""
    bps = (n + 3) // 4
    piece = (256 << 20) // bps
    print(f"panel upload n={n} m={m} piece_rows={piece}", flush=True)
    rng = np.random.default_rng(n)
    bed = np.empty(3 + bps * m, np.uint8)
    bed[:3] = (0x6C, 0x1B, 0x01)
    bed[3:] = rng.integers(0, 256, bps * m, dtype=np.uint8)
    idx = np.array(sorted({0, 1, piece - 1, piece, piece + 1, m - 2, m - 1}), np.uint64)
    with L.Panel(bed, n, m) as panel:
        dos, _, st, _ = L.decode(panel, idx)
    assert np.array_equal(st, oracle.stats(bed, n, idx))
    assert np.array_equal(dos, oracle.decode(bed, n, idx))


@pytest.mark.parametrize("n,d", [(1000, 50), (2049, 200), (3000, 3000), (777, 5)])
def test_sketch_buckets_bit_exact(oracle, n, d):
    print(f"sketch buckets n={n} d={d}", flush=True)
    p = _panel(n, 96, seed=d)
    idx = np.arange(96, dtype=np.uint64)[::-1].copy()
    with L.Panel(p.bed_bytes(), n, p.m) as panel:
        sg, sm = L.sketch_buckets(panel, idx, d)
    esg, esm = oracle.sketch_int(p.bed_bytes(), n, idx, d)
    assert np.array_equal(sg, esg) and np.array_equal(sm, esm)


@pytest.mark.parametrize("n,d", [(16385, 7), (20001, 128), (16385, 150), (20001, 190),
                                 (40000, 200), (16500, 224), (20001, 256),
                                 (40000, 3000), (16884, 2000), (33000, 5000)])
def test_sketch_buckets_chunked_order(oracle, n, d):
    """More members than one 16,384-member chunk.  Up to 7 bucket blocks
    (d <= 224) the sums come from the dense one-pass kernel (natural order,
    the bucket matrix built in LDS, the blocks of a SNP group split over two
    waves; n = 16,385, 16,500 and 20,001 end in a partial pair of K-steps,
    and in a phase of two pairs that is partly past the last individual).
    Beyond, each chunk is put in bucket order on its own (the last one
    ragged) and the one-pass gather kernel sums it; a short last chunk (500
    or 232 members) has K-steps that reach several bucket blocks.  The
    integer sums do not depend on either order."""
    print(f"sketch buckets chunked n={n} d={d}", flush=True)
    p = _panel(n, 70, seed=n + d)
    idx = np.arange(70, dtype=np.uint64)[::-1].copy()
    with L.Panel(p.bed_bytes(), n, p.m) as panel:
        sg, sm = L.sketch_buckets(panel, idx, d)
    esg, esm = oracle.sketch_int(p.bed_bytes(), n, idx, d)
    assert np.array_equal(sg, esg) and np.array_equal(sm, esm)


def test_sketch_buckets_with_keep(oracle):
    n, d = 1200, 97
    p = _panel(n, 40, seed=8)
    keep = np.random.default_rng(1).permutation(n)[:900]
    with L.Panel(p.bed_bytes(), n, p.m, keep=keep) as panel:
        sg, sm = L.sketch_buckets(panel, np.arange(40), d)
    esg, esm = oracle.sketch_int(p.bed_bytes(), n, np.arange(40), d, keep=keep)
    assert np.array_equal(sg, esg) and np.array_equal(sm, esm)


@pytest.mark.parametrize("use_f32", [False, True])
def test_l2_keep_annot_pq(oracle, use_f32):
    n, m = 613, 900
    p = _panel(n, m, seed=21)
    rng = np.random.default_rng(2)
    keep = np.sort(rng.choice(n, 500, replace=False))
    annot = synth.make_annot(m, seed=5, n_bin=3, n_cont=2)
    annot[rng.random(m) < 0.01, 0] = np.nan
    snps = L.make_snp_array(p.chr, p.cm, p.bp)
    cfg = L.L2Config(mode=L.WindowMode.Kb(120.0), chunk_size=50, use_f32=use_f32, pq_exp=0.75)
    with L.Panel(p.bed_bytes(), n, m, keep=keep) as panel:
        got, maf, _ = L.compute_ldscore(panel, snps, cfg, annot=annot)
    exp, emaf = oracle.l2(p.bed_bytes(), n, p.chr, p.cm, p.bp, mode=1, window=120.0, chunk=50,
                          use_f32=use_f32, pq_exp=0.75, annot=np.nan_to_num(annot), keep=keep)
    assert got.shape == (m, 5)
    assert _rel(got, exp) < (TOL_F32 if use_f32 else TOL_EXACT)
    assert np.array_equal(maf, emaf)


@pytest.mark.parametrize("K", [2, 16, 17, 37])
def test_l2_many_annotations(oracle, K):
    """Annotation counts on and across the 16-column passes of the FP64
    contraction (one, exactly one full, one plus one column, two plus a
    ragged third), NaN entries, pq scaling; both exact kernels and FP32."""
    n, m = 331, 700
    p = _panel(n, m, seed=K)
    rng = np.random.default_rng(K)
    annot = rng.lognormal(size=(m, K))
    annot[rng.random((m, K)) < 0.5] = 0.0
    annot[rng.random((m, K)) < 0.01] = np.nan
    snps = L.make_snp_array(p.chr, p.cm, p.bp)
    with L.Panel(p.bed_bytes(), n, m) as panel:
        for use_f32, kern in ((False, "auto"), (False, "i8"), (True, "auto")):
            print(f"many annotations K={K} f32={use_f32} kernel={kern}", flush=True)
            cfg = L.L2Config(mode=L.WindowMode.Kb(150.0), chunk_size=40, use_f32=use_f32,
                             pq_exp=0.5, exact_kernel=kern)
            got, maf, _ = L.compute_ldscore(panel, snps, cfg, annot=annot)
            exp, emaf = oracle.l2(p.bed_bytes(), n, p.chr, p.cm, p.bp, mode=1, window=150.0,
                                  chunk=40, use_f32=use_f32, pq_exp=0.5,
                                  annot=np.nan_to_num(annot))
            assert got.shape == (m, K)
            assert _rel(got, exp) < (TOL_F32 if use_f32 else TOL_EXACT), (use_f32, kern)
            assert np.array_equal(maf, emaf)


def test_l2_large_annotation_matrix(oracle):
    """A 36 MB annotation matrix and result: both cross the host copies'
    pinned ring several times, with a ragged last piece; repeated calls
    reuse the ring."""
    n, m, K = 100, 100_003, 45
    p = _panel(n, m, seed=7)
    annot = np.random.default_rng(1).random((m, K))
    snps = L.make_snp_array(p.chr, p.cm, p.bp)
    exp, _ = oracle.l2(p.bed_bytes(), n, p.chr, p.cm, p.bp, mode=2, window=10, chunk=50,
                       annot=annot)
    cfg = L.L2Config(mode=L.WindowMode.Snp(10), chunk_size=50)
    with L.Panel(p.bed_bytes(), n, m) as panel:
        got, _, _ = L.compute_ldscore(panel, snps, cfg, annot=annot)
        dest = np.full((K, m), np.nan).T
        again, _, _ = L.compute_ldscore(panel, snps, cfg, annot=np.asfortranarray(annot), out=dest)
        with pytest.raises(ValueError):
            L.compute_ldscore(panel, snps, cfg, annot=annot, out=np.empty((m, K)))
    assert _rel(got, exp) < TOL_EXACT
    assert again is dest and np.array_equal(got, again)


@pytest.mark.parametrize("m,K,f32,pq", [(100_000, 48, False, None), (200_000, 96, False, None),
                                        (100_000, 48, True, None), (100_000, 48, False, 0.75)])
def test_segmented_annotation_upload_matches_one_piece(m, K, f32, pq):
    """A large annotation goes up in row segments interleaved with the band
    (two segments at 38 MB, four at 154 MB) on the one-call path; the
    request server uploads it in one piece.  The tiles and every reduction
    are the same, so the results are bit-identical, for the FP32 form and
    with pq scaling too."""
    from ldsc_amd.serve import Server

    n = 600
    print(f"segmented annotation m={m} K={K} f32={f32} pq={pq}", flush=True)
    p = _panel(n, m, seed=K, miss=0.01, nchr=2)
    snps = L.make_snp_array(p.chr, p.cm, p.bp)
    annot = np.random.default_rng(K).random((m, K))
    annot[::97, 3] = np.nan
    cfg = L.L2Config(mode=L.WindowMode.Kb(300.0), chunk_size=200, use_f32=f32, pq_exp=pq)
    with L.Panel(p.bed_bytes(), n, m) as panel:
        got, maf, _ = L.compute_ldscore(panel, snps, cfg, annot=annot)
        with Server(hbm_budget_bytes=16 << 30, max_inflight=1) as srv:
            t = srv.submit(panel, snps, cfg, annot)
            srv.wait(t)
            ref, rmaf = t.result()
    assert np.array_equal(got, ref)
    assert np.array_equal(maf, rmaf)
    assert np.all(np.isfinite(got))


def test_l2_extract_per_chromosome(oracle):
    n, m = 400, 1200
    p = _panel(n, m, seed=31, nchr=3)
    sel = np.sort(np.random.default_rng(4).choice(m, 800, replace=False))
    snps = L.make_snp_array(p.chr[sel], p.cm[sel], p.bp[sel], bed_idx=sel)
    for per_chr in (False, True):
        for masking in (False, True):
            cfg = L.L2Config(mode=L.WindowMode.Cm(0.1), chunk_size=40, per_chromosome=per_chr,
                             snp_level_masking=masking)
            with L.Panel(p.bed_bytes(), n, m) as panel:
                got, maf, _ = L.compute_ldscore(panel, snps, cfg)
            exp, emaf = oracle.l2(p.bed_bytes(), n, p.chr[sel], p.cm[sel], p.bp[sel],
                                  bed_idx=sel, mode=0, window=0.1, chunk=40, masking=masking,
                                  per_chromosome=per_chr)
            assert _rel(got, exp) < TOL_EXACT, (per_chr, masking)
            assert np.array_equal(maf, emaf)


def test_l2_special_columns(oracle):
    """Identical, monomorphic and all-missing SNPs; N not a multiple of 4."""
    n, m = 203, 130
    p = _panel(n, m, seed=41)
    body = p.body
    body[10] = body[9]
    body[20] = 0x00
    body[30] = 0x55
    body[40] = 0xFF
    snps = L.make_snp_array(p.chr, p.cm, p.bp)
    for use_f32 in (False, True):
        cfg = L.L2Config(mode=L.WindowMode.Snp(20), chunk_size=16, use_f32=use_f32)
        with L.Panel(p.bed_bytes(), n, m) as panel:
            got, maf, _ = L.compute_ldscore(panel, snps, cfg)
        exp, emaf = oracle.l2(p.bed_bytes(), n, p.chr, p.cm, p.bp, mode=2, window=20, chunk=16,
                              use_f32=use_f32)
        assert _rel(got, exp) < (TOL_F32 if use_f32 else TOL_EXACT)
        assert np.array_equal(maf, emaf)
        assert np.all(np.isfinite(got))


@pytest.mark.parametrize("n,d,maf_aware", [(3000, 64, False), (3000, 300, False),
                                           (3000, 300, True), (3000, 1000, False),
                                           (20000, 500, True), (20000, 3, False),
                                           (20000, 3, True), (16884, 2000, True),
                                           (20000, 200, True), (33001, 256, False),
                                           (33001, 128, False), (20000, 100, True)])
def test_l2_sketch_matches_oracle(oracle, n, d, maf_aware):
    """Sketch LD scores against the f32 sequential-scatter restatement.  The
    contraction runs on exact int8 operands (w = SG - k SM, e = T - SM)
    while they fit; at d = 3 over 20,000 individuals (buckets of ~6,700
    members) they do not, and the f32 form runs instead.  Over 16,884
    individuals the last member chunk has 500, so a K-step of its bucket
    order reaches up to four blocks of d = 2000.  Over 16,384 individuals
    the bucket sums come from a one-pass kernel: the dense one up to d = 224
    (d = 3, 100, 128, 200 here), the gather one up to d = 256 (256); above,
    the chunk gather into bucket-order tiles and the tile sums (500, 2000)."""
    m = 700
    p = _panel(n, m, seed=d)
    snps = L.make_snp_array(p.chr, p.cm, p.bp)
    cfg = L.L2Config(mode=L.WindowMode.Cm(0.15), chunk_size=60, sketch=d,
                     sketch_maf_aware=maf_aware, snp_level_masking=True)
    print(f"sketch n={n} d={d} maf_aware={maf_aware}", flush=True)
    with L.Panel(p.bed_bytes(), n, m) as panel:
        got, maf, perf = L.compute_ldscore(panel, snps, cfg)
    exp, emaf = oracle.l2(p.bed_bytes(), n, p.chr, p.cm, p.bp, mode=0, window=0.15, chunk=60,
                          sketch=d, maf_aware=maf_aware, use_f32=True, masking=True)
    print(f"  max rel {_rel(got, exp):.3e}, int8 operands: {perf.sketch_int8}", flush=True)
    assert _rel(got, exp) < TOL_SKETCH
    assert np.array_equal(maf, emaf)
    assert perf.sketch_int8 == (d > 3)


def test_sketch_errors_and_unmasked(oracle):
    n, m = 500, 300
    p = _panel(n, m, seed=77)
    snps = L.make_snp_array(p.chr, p.cm, p.bp)
    with L.Panel(p.bed_bytes(), n, m) as panel:
        for bad in (2, n):
            with pytest.raises(L.LdscError) as e:
                L.compute_ldscore(panel, snps, L.L2Config(sketch=bad))
            assert e.value.code == 4
        cfg = L.L2Config(mode=L.WindowMode.Snp(40), chunk_size=25, sketch=100)
        got, _, _ = L.compute_ldscore(panel, snps, cfg)
    exp, _ = oracle.l2(p.bed_bytes(), n, p.chr, p.cm, p.bp, mode=2, window=40, chunk=25,
                       sketch=100, use_f32=True)
    assert _rel(got, exp) < TOL_SKETCH


def test_l2_deterministic_and_chunk_invariant_when_masked():
    n, m = 2000, 3000
    p = _panel(n, m, seed=5)
    snps = L.make_snp_array(p.chr, p.cm, p.bp)
    with L.Panel(p.bed_bytes(), n, m) as panel:
        a, _, _ = L.compute_ldscore(panel, snps, L.L2Config(mode=L.WindowMode.Cm(0.2),
                                                             chunk_size=200,
                                                             snp_level_masking=True))
        b, _, _ = L.compute_ldscore(panel, snps, L.L2Config(mode=L.WindowMode.Cm(0.2),
                                                             chunk_size=7,
                                                             snp_level_masking=True))
        c, _, _ = L.compute_ldscore(panel, snps, L.L2Config(mode=L.WindowMode.Cm(0.2),
                                                             chunk_size=200,
                                                             snp_level_masking=True))
    assert np.array_equal(a, b) and np.array_equal(a, c)


# The e2m1 and int8 kernels form the same integer products and the same
# per-pair FP64 r2u; only the order in which a tile's r2u are summed into its
# partial sums differs, so their LD scores agree to a few ulps.  A wrong
# integer anywhere moves an LD score by far more than this.
TOL_KERNELS = 1e-12


def test_exact_kernels_agree():
    """e2m1 vs int8 MFMA exact kernels: N around the K-block edges, M off the
    tile edges, --keep, several annotations with pq, --extract across
    chromosome runs with and without masking, special columns."""
    def both(panel, snps, annot=None, **kw):
        a, ma, _ = L.compute_ldscore(panel, snps, L.L2Config(**kw), annot=annot)
        b, mb, _ = L.compute_ldscore(panel, snps, L.L2Config(**kw, exact_kernel="i8"), annot=annot)
        return a, b, ma, mb

    for n, m in ((1, 40), (3, 65), (127, 130), (128, 64), (129, 200), (1000, 777)):
        print(f"exact kernels n={n} m={m}", flush=True)
        p = _panel(n, m, seed=n + m)
        snps = L.make_snp_array(p.chr, p.cm, p.bp)
        with L.Panel(p.bed_bytes(), n, m) as panel:
            a, b, ma, mb = both(panel, snps, mode=L.WindowMode.Snp(50), chunk_size=30)
        assert _rel(a, b) < TOL_KERNELS and np.array_equal(ma, mb), (n, m, _rel(a, b))
    n, m = 613, 900
    p = _panel(n, m, seed=21, nchr=3)
    body = p.body
    body[10] = body[9]
    body[20] = 0x00
    body[30] = 0x55
    body[40] = 0xFF
    rng = np.random.default_rng(2)
    keep = np.sort(rng.choice(n, 500, replace=False))
    annot = synth.make_annot(m, seed=5, n_bin=3, n_cont=2)
    snps = L.make_snp_array(p.chr, p.cm, p.bp)
    sel = np.sort(rng.choice(m, 700, replace=False))
    snps_sel = L.make_snp_array(p.chr[sel], p.cm[sel], p.bp[sel], bed_idx=sel)
    with L.Panel(p.bed_bytes(), n, m, keep=keep) as panel:
        a, b, _, _ = both(panel, snps, annot=annot, mode=L.WindowMode.Kb(120.0), chunk_size=50,
                          pq_exp=0.75)
        assert _rel(a, b) < TOL_KERNELS
        for per_chr in (False, True):
            for masking in (False, True):
                a, b, _, _ = both(panel, snps_sel, mode=L.WindowMode.Cm(0.1), chunk_size=40,
                                  per_chromosome=per_chr, snp_level_masking=masking)
                assert _rel(a, b) < TOL_KERNELS, (per_chr, masking)


@pytest.mark.parametrize("n", [1_800_000, 1_800_001])
def test_exact_kernels_at_the_e2m1_limit(n):
    """The largest N of the e2m1 kernel (f32 sums of products of v/2, q/2
    exact while 9 N_pad <= 2^24) and one past it (the exact form switches to
    the int8 kernel).  Planted rows make the sums as large as they get: two
    rows all missing (v/2 = 1.5 in every cell, 2.25 N_pad per pair), two all
    dosage 2, one all dosage 1.  Either way the result must equal the int8
    kernel's (int32 sums) to a few ulps."""
    m = 96
    p = _panel(n, m, seed=17, miss=0.02)
    bps = (n + 3) // 4
    for row, byte in ((3, 0x55), (40, 0x55), (7, 0x00), (50, 0x00), (60, 0xAA)):
        p.body[row, :] = byte
        if n % 4:  # padding bits of the last byte stay code 00
            p.body[row, bps - 1] &= (1 << (2 * (n % 4))) - 1
    snps = L.make_snp_array(p.chr, p.cm, p.bp)
    print(f"e2m1 limit n={n} m={m}", flush=True)
    with L.Panel(p.bed_bytes(), n, m) as panel:
        cfg = dict(mode=L.WindowMode.Snp(60), chunk_size=32)
        a, ma, _ = L.compute_ldscore(panel, snps, L.L2Config(**cfg))
        b, mb, _ = L.compute_ldscore(panel, snps, L.L2Config(**cfg, exact_kernel="i8"))
    assert np.array_equal(ma, mb)
    assert _rel(a, b) < TOL_KERNELS, _rel(a, b)
    assert np.all(np.isfinite(a))


def test_exact_and_fp32_at_panel_scale():
    """N=50,000 x M=8,000 slice of the panel-scale workload: the e2m1 and
    int8 exact kernels agree to a few ulps, the FP32 form agrees to its
    tolerance, and the result is reproducible bit for bit."""
    n, m = 50_000, 8_000
    p = _panel(n, m, seed=20261205, miss=0.01)
    snps = L.make_snp_array(p.chr, p.cm, p.bp)
    base = dict(mode=L.WindowMode.Cm(1.0), chunk_size=200, snp_level_masking=True)
    with L.Panel(p.bed_bytes(), n, m) as panel:
        a, _, perf = L.compute_ldscore(panel, snps, L.L2Config(**base))
        b, _, _ = L.compute_ldscore(panel, snps, L.L2Config(**base, use_f32=True))
        a2, _, _ = L.compute_ldscore(panel, snps, L.L2Config(**base))
        c, _, _ = L.compute_ldscore(panel, snps, L.L2Config(**base, exact_kernel="i8"))
    assert np.array_equal(a, a2)
    assert _rel(c, a) < TOL_KERNELS
    assert _rel(b, a) < TOL_F32
    # l2 >= self term minus the bias of pairs: every SNP has its own 1
    assert np.all(a > 0.5)
