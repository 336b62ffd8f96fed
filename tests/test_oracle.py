"""Pins the CPU restatement (oracle/) before it is trusted as the checker.

* Known answers held by the reference's own tests (cited per test).
* A brute-force definition written here from the mathematics: Pearson r of
  mean-imputed genotypes, the unbiased r^2, and the inclusion rules computed
  directly from coordinates (masked: per-SNP window; unmasked: the chunk
  grid) -- independent of the oracle's chunk/ring loop.
* Invariants: identical SNPs, monomorphic and all-missing SNPs, empty
  windows, chunk-size invariance under masking, the sketch bias inversion.
"""
import numpy as np
import pytest

from ldsc_amd import synth

WIN = {"cm": 0, "kb": 1, "snp": 2}


# ---------------------------------------------------------------------------
# brute force
# ---------------------------------------------------------------------------
def random_panel(rng, n, m, miss=0.05, nchr=1):
    codes = rng.integers(0, 4, size=(m, n))
    codes[rng.random((m, n)) < miss] = 1
    codes[(codes == 1) & (rng.random((m, n)) >= miss * 2)] = 3
    bps = (n + 3) // 4
    body = np.zeros((m, bps), np.uint8)
    for i in range(n):
        body[:, i // 4] |= (codes[:, i] << (2 * (i % 4))).astype(np.uint8)
    bed = np.concatenate([np.array([0x6C, 0x1B, 0x01], np.uint8), body.reshape(-1)])
    chr_ = np.sort(rng.integers(1, nchr + 1, size=m)).astype(np.uint8)
    bp = np.cumsum(rng.integers(1, 4000, size=m)).astype(np.uint32)
    cm = bp * 1e-6
    return bed, codes, chr_, cm, bp


def dosages(codes):
    return np.array([2, -1, 1, 0])[codes]  # m x n, -1 missing


def std_cols(dos):
    """Mean-imputed standardized columns, variance over all N individuals."""
    m, n = dos.shape
    X = np.zeros((n, m))
    for j in range(m):
        pres = dos[j] >= 0
        c = pres.sum()
        mean = dos[j, pres].sum() / c if c else 0.0
        css = ((dos[j, pres] - mean) ** 2).sum()
        sd = np.sqrt(css / n)
        if sd > 0:
            X[pres, j] = (dos[j, pres] - mean) / sd
    return X


def coords(mode, cm, bp):
    return np.asarray(cm, float) if mode == "cm" else np.asarray(bp, float) / 1000.0


def window_ok(mode, w, chr_, cm, bp, i, j):
    """i < j share a window: same chromosome run and distance within w."""
    if np.any(chr_[i:j + 1] != chr_[j]):
        return False
    if mode == "snp":
        return j - i <= int(w)
    c = coords(mode, cm, bp)
    return abs(c[j] - c[i]) <= w


def chunk_ok(mode, w, chr_, cm, bp, chunk, per_chr, i, j):
    """Unmasked rule: i >= a_left(chunk_start(j)), chunk grid from the run
    start (per chromosome) or from 0 (global pass)."""
    m = len(chr_)
    run0 = 0
    if per_chr:
        run0 = j
        while run0 > 0 and chr_[run0 - 1] == chr_[j]:
            run0 -= 1
        if chr_[i] != chr_[j] or i < run0:
            return False
    s = run0 + ((j - run0) // chunk) * chunk
    # block_left of s: first index in s's chromosome run within the window
    r0 = s
    while r0 > 0 and chr_[r0 - 1] == chr_[s]:
        r0 -= 1
    bl = s
    for q in range(r0, s + 1):
        if window_ok(mode, w, chr_, cm, bp, q, s) or q == s:
            bl = q
            break
    dist = s - bl
    bs = 0 if dist == 0 else -(-dist // chunk) * chunk
    a_left = s - min(bs, s - run0)
    return i >= a_left


def brute_l2(bed, n, chr_, cm, bp, mode, w, chunk, masking, per_chr=False, annot=None,
             pq_exp=None, keep=None):
    m = len(chr_)
    body = bed[3:].reshape(m, -1)
    codes = np.zeros((m, n if keep is None else len(keep)), int)
    rows = np.arange(n) if keep is None else np.asarray(keep)
    for k, r in enumerate(rows):
        codes[:, k] = (body[:, r // 4] >> (2 * (r % 4))) & 3
    dos = dosages(codes)
    nn = codes.shape[1]
    X = std_cols(dos)
    r2 = (X.T @ X / nn) ** 2
    D = nn - 2.0 if nn > 2 else float(nn)
    r2u = r2 * (1 + 1 / D) - 1 / D
    A = np.ones((m, 1)) if annot is None else np.asarray(annot, float).reshape(m, -1)
    if pq_exp is not None:
        maf = np.zeros(m)
        for j in range(m):
            pres = dos[j] >= 0
            mean = dos[j, pres].mean() if pres.any() else 0.0
            f = min(max(mean / 2, 0), 1)
            maf[j] = min(f, 1 - f)
        A = A * ((maf * (1 - maf)) ** pq_exp)[:, None]
    L = A.copy()
    for j in range(m):
        for i in range(j):
            ok = (window_ok(mode, w, chr_, cm, bp, i, j) if masking else
                  chunk_ok(mode, w, chr_, cm, bp, chunk, per_chr, i, j))
            if ok:
                L[j] += r2u[i, j] * A[i]
                L[i] += r2u[i, j] * A[j]
    return L


# ---------------------------------------------------------------------------
# known answers from the reference's tests
# ---------------------------------------------------------------------------
def test_reference_smoke_known_answer(oracle):
    """src/l2/mod.rs:1191-1225: 4 individuals x 2 identical SNPs (bytes F0 F0),
    kb window 100, chunk 2 -> l2 = [2, 2], MAF = [0.5, 0.5] (tol 1e-9)."""
    bed = np.array([0x6C, 0x1B, 0x01, 0xF0, 0xF0], np.uint8)
    for use_f32 in (False, True):
        l2, maf = oracle.l2(bed, 4, [1, 1], [0.0, 0.0], [100, 200], mode=oracle.WINDOW_KB,
                            window=100.0, chunk=2, use_f32=use_f32)
        assert np.allclose(l2[:, 0], [2.0, 2.0], atol=1e-9)
        assert np.allclose(maf, [0.5, 0.5], atol=1e-9)


def test_reference_bed_validation(oracle):
    """src/bed.rs:786-823 and src/l2/mod.rs:1229-1246."""
    good = np.array([0x6C, 0x1B, 0x01, 0, 0], np.uint8)  # 4 indiv x 2 SNPs
    assert oracle.bed_validate(good, 4, 2) == 0
    bad = good.copy()
    bad[0] = 0xFF
    assert oracle.bed_validate(bad, 4, 2) == 2           # invalid magic
    assert oracle.bed_validate(good[:-1], 4, 2) == 4      # invalid length
    assert oracle.bed_validate(np.array([0x6C, 0x1B, 0x01, 0xF0], np.uint8), 4, 2) == 4
    mode = good.copy()
    mode[2] = 0
    assert oracle.bed_validate(mode, 4, 2) == 3           # not SNP-major


def test_reference_decode_lut(oracle):
    """build_lut count_a1 = [2, NaN, 1, 0] (src/bed.rs:470-476), low bits first."""
    # byte 0b11_10_01_00 -> individuals 0..3 have codes 0,1,2,3
    bed = np.array([0x6C, 0x1B, 0x01, 0b11100100], np.uint8)
    d = oracle.decode(bed, 4, [0])
    assert d[:, 0].tolist() == [2, -1, 1, 0]
    st = oracle.stats(bed, 4, [0])
    assert st[0].tolist() == [3, 3, 5, 1]  # sum, count, sum_sq, het


def _lcg_bed_fixture(seed, c, n):
    """make_bed_fixture (src/wasm_simd.rs:2818-2843)."""
    bps = (n + 3) // 4
    state = seed
    raw = np.zeros(c * bps, np.uint8)
    M = (1 << 64) - 1
    for j in range(c):
        for b in range(bps):
            byte = 0
            for slot in range(4):
                state = (state * 0x9E3779B97F4A7C15 + 1) & M
                r = (state >> 32) % 100
                code = 0b01 if r <= 5 else (0b00 if r <= 37 else (0b10 if r <= 69 else 0b11))
                byte |= code << (2 * slot)
            raw[j * bps + b] = byte
    return raw


def _lcg_bucket_sign(n, d, seed):
    """build_cs_bucket_sign (src/wasm_simd.rs:2900-2916)."""
    M = (1 << 64) - 1
    state = seed
    bucket = np.zeros(n, np.uint32)
    sign = np.zeros(n, np.float32)
    for i in range(n):
        state = (state * 0x9E3779B97F4A7C15 + 1) & M
        bucket[i] = ((state >> 32) & 0xFFFFFFFF) % d
    for i in range(n):
        state = (state * 0x9E3779B97F4A7C15 + 1) & M
        sign[i] = 1.0 if (state >> 63) & 1 == 0 else -1.0
    return bucket, sign


def test_reference_scatter_fixture(oracle):
    """scatter_one_column_f32_matches_ref_all_iids (src/wasm_simd.rs:2918-2998):
    c=8, n=503, d=50, seeds 42 / 7; the LUT scatter equals the direct
    decode->normalize->scatter->renorm reference (bit-identical by
    construction; the reference asserts 1e-4)."""
    import ctypes as C
    c, n, d = 8, 503, 50
    raw = _lcg_bed_fixture(42, c, n)
    bps = (n + 3) // 4
    bucket, sign = _lcg_bucket_sign(n, d, 7)
    L = oracle.lib()
    L.orc_scatter_col_f32.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_float,
                                      C.c_float, C.c_double, C.c_uint64, C.c_void_p]
    for j in range(c):
        slab = np.ascontiguousarray(raw[j * bps:(j + 1) * bps])
        codes = np.array([(slab[i // 4] >> (2 * (i % 4))) & 3 for i in range(n)])
        g = np.array([2.0, np.nan, 1.0, 0.0])[codes]
        pres = ~np.isnan(g)
        mean = g[pres].sum() / pres.sum()
        var = (np.sum(g[pres] ** 2) - pres.sum() * mean * mean) / n
        mean32, inv32 = np.float32(mean), np.float32(1.0 / np.sqrt(var))
        # direct reference, f32 ops in order
        col = np.zeros(d, np.float32)
        for i in range(n):
            val = np.float32(0.0) if not pres[i] else (np.float32(g[i]) - mean32) * inv32
            col[bucket[i]] = np.float32(col[bucket[i]] + sign[i] * val)
        nrm = float(np.sum(col.astype(np.float64) ** 2))
        col = col * np.float32(np.sqrt(n / nrm))
        out = np.zeros(d, np.float32)
        L.orc_scatter_col_f32(slab.ctypes.data, n, bucket.ctypes.data, sign.ctypes.data,
                              float(mean32), float(inv32), float(n), d, out.ctypes.data)
        assert np.array_equal(out, col), j


def test_quadratic_correction_inverts_bias(oracle):
    """compute.rs:928-937: feeding t = r^2 + (1-r^2)(1-2r^2)/d returns
    r^2 (1 + 1/D) - 1/D exactly (up to rounding)."""
    import ctypes as C
    L = oracle.lib()
    L.orc_quad_corr.argtypes = [C.c_double] * 7
    L.orc_quad_corr.restype = C.c_double
    n, d = 1000.0, 200.0
    D = n - 2
    for r2 in np.linspace(0.0, 1.0, 21):
        t = r2 + (1 - r2) * (1 - 2 * r2) / d
        val = n * np.sqrt(t)
        got = L.orc_quad_corr(val, 1 / n ** 2, d, 1 + 1 / D, -1 / D, 0.0, 0.0)
        assert abs(got - (r2 * (1 + 1 / D) - 1 / D)) < 1e-12
    # negative discriminant -> r^2_s = 0
    assert L.orc_quad_corr(0.0, 1e-6, 3.0, 1.0, 0.0, 0.0, 0.0) in (0.0,)


def test_wyrand_stream_shape(oracle):
    b, s = oracle.countsketch_hash(10000, 37)
    assert b.max() < 37 and set(np.unique(s)) <= {-1, 1}
    # Lemire reduction is near-uniform; signs balanced
    assert abs(np.bincount(b, minlength=37).std() / (10000 / 37)) < 0.2
    assert abs(s.mean()) < 0.05


# ---------------------------------------------------------------------------
# oracle == brute force
# ---------------------------------------------------------------------------
CASES = [
    # mode, window, chunk, masking, per_chr, nchr
    ("kb", 9.0, 3, True, False, 1),
    ("kb", 9.0, 3, False, False, 1),
    ("kb", 9.0, 4, False, False, 2),
    ("kb", 9.0, 4, False, True, 3),
    ("cm", 0.007, 5, True, False, 3),
    ("cm", 0.007, 7, False, False, 3),
    ("snp", 4, 3, True, True, 2),
    ("snp", 6, 1, False, False, 2),
    ("snp", 3, 50, False, False, 1),
    ("kb", 1e9, 6, False, False, 1),
]


@pytest.mark.parametrize("mode,w,chunk,masking,per_chr,nchr", CASES)
def test_oracle_equals_brute_force(oracle, mode, w, chunk, masking, per_chr, nchr):
    rng = np.random.default_rng(abs(hash((mode, w, chunk, masking, per_chr, nchr))) % 2**32)
    n, m = 37, 29
    bed, codes, chr_, cm, bp = random_panel(rng, n, m, nchr=nchr)
    annot = rng.random((m, 3))
    annot[rng.random((m, 3)) < 0.3] = 0.0
    exp = brute_l2(bed, n, chr_, cm, bp, mode, w, chunk, masking, per_chr, annot, pq_exp=0.5)
    got, _ = oracle.l2(bed, n, chr_, cm, bp, mode=WIN[mode], window=w, chunk=chunk,
                       masking=masking, per_chromosome=per_chr, annot=annot, pq_exp=0.5)
    assert np.allclose(got, exp, rtol=1e-12, atol=1e-12)
    got32, _ = oracle.l2(bed, n, chr_, cm, bp, mode=WIN[mode], window=w, chunk=chunk,
                         masking=masking, per_chromosome=per_chr, annot=annot, pq_exp=0.5,
                         use_f32=True)
    assert np.allclose(got32, exp, rtol=1e-4, atol=1e-4)


def test_oracle_keep_and_extract(oracle):
    rng = np.random.default_rng(5)
    n, m = 41, 33
    bed, codes, chr_, cm, bp = random_panel(rng, n, m)
    keep = np.sort(rng.choice(n, 29, replace=False))
    exp = brute_l2(bed, n, chr_, cm, bp, "kb", 20.0, 4, False, keep=keep)
    got, _ = oracle.l2(bed, n, chr_, cm, bp, mode=1, window=20.0, chunk=4, keep=keep)
    assert np.allclose(got, exp, rtol=1e-12, atol=1e-12)
    # --extract: a non-contiguous subset of BED rows keeps its own coordinates
    sel = np.sort(rng.choice(m, 20, replace=False))
    sub = np.concatenate([bed[:3], bed[3:].reshape(m, -1)[sel].reshape(-1)])
    exp2, _ = oracle.l2(sub, n, chr_[sel], cm[sel], bp[sel], mode=1, window=20.0, chunk=4)
    got2, _ = oracle.l2(bed, n, chr_[sel], cm[sel], bp[sel], bed_idx=sel, mode=1, window=20.0,
                        chunk=4)
    assert np.allclose(got2, exp2, rtol=1e-13, atol=1e-13)


def test_oracle_special_columns(oracle):
    """Identical SNPs give r^2 = 1; monomorphic / all-missing SNPs give
    r = 0 -> r2u = -1/D with partners and self term 1."""
    rng = np.random.default_rng(9)
    n, m = 24, 6
    bed, codes, chr_, cm, bp = random_panel(rng, n, m)
    body = bed[3:].reshape(m, -1)
    body[1] = body[0]                 # identical to SNP 0
    body[3] = 0x00                    # monomorphic (all dosage 2)
    body[4] = 0x55                    # all missing
    bed = np.concatenate([bed[:3], body.reshape(-1)])
    l2, maf = oracle.l2(bed, n, np.ones(m), np.zeros(m), np.arange(m) * 10 + 1, mode=2,
                        window=m, chunk=2)
    exp = brute_l2(bed, n, np.ones(m, np.uint8), np.zeros(m), np.arange(m) * 10 + 1, "snp", m,
                   2, False)
    assert np.allclose(l2, exp, atol=1e-12)
    D = n - 2.0
    # SNP 3 (monomorphic): every partner contributes -1/D
    assert abs(l2[3, 0] - (1.0 - (m - 1) / D)) < 1e-12
    assert abs(l2[4, 0] - (1.0 - (m - 1) / D)) < 1e-12
    assert maf[3] == 0.0 and maf[4] == 0.0


def test_oracle_chunk_size_invariance_under_masking(oracle):
    p = synth.make_panel(n=150, m=400, seed=3, miss_rate=0.02)
    ref, _ = oracle.l2(p.bed_bytes(), p.n, p.chr, p.cm, p.bp, mode=0, window=0.05, chunk=200,
                       masking=True)
    for c in (1, 7, 64):
        got, _ = oracle.l2(p.bed_bytes(), p.n, p.chr, p.cm, p.bp, mode=0, window=0.05, chunk=c,
                           masking=True)
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_oracle_empty_window(oracle):
    rng = np.random.default_rng(2)
    bed, codes, chr_, cm, bp = random_panel(rng, 30, 12)
    annot = rng.random((12, 2))
    l2, _ = oracle.l2(bed, 30, chr_, cm, bp, mode=2, window=0, chunk=1, masking=True,
                      annot=annot)
    assert np.allclose(l2, annot)


def test_oracle_sketch_errors(oracle):
    bed = np.array([0x6C, 0x1B, 0x01, 0xF0, 0xF0], np.uint8)
    with pytest.raises(ValueError):
        oracle.l2(bed, 4, [1, 1], [0, 0], [1, 2], sketch=4)
    with pytest.raises(ValueError):
        oracle.l2(bed, 4, [1, 1], [0, 0], [1, 2], sketch=2)
    with pytest.raises(ValueError):
        oracle.l2(bed, 4, [1, 1], [0, 0], [1, 2], chunk=0)


def test_oracle_sketch_tracks_exact(oracle):
    """The quadratic-corrected sketch is an estimator of the exact l2: at
    d = N/2 its error is far below the spread of l2."""
    p = synth.make_panel(n=2000, m=300, seed=4, miss_rate=0.01)
    ex, _ = oracle.l2(p.bed_bytes(), p.n, p.chr, p.cm, p.bp, mode=0, window=0.1, chunk=50,
                      masking=True)
    sk, maf = oracle.l2(p.bed_bytes(), p.n, p.chr, p.cm, p.bp, mode=0, window=0.1, chunk=50,
                        masking=True, sketch=1000, use_f32=True)
    assert np.corrcoef(ex[:, 0], sk[:, 0])[0, 1] > 0.98


# ---------------------------------------------------------------------------
# sketch chain apart from the hash, and the MAF-aware kurtosis
# ---------------------------------------------------------------------------
def brute_kappa(dos):
    """kappa_j = sum (g - mu)^4 / css^2 over the present individuals, the
    definition per_snp_kappa (compute.rs:525-538) computes from raw moments;
    0 for a column with no variance or no present individual."""
    out = np.zeros(dos.shape[0])
    for j in range(dos.shape[0]):
        g = dos[j][dos[j] >= 0].astype(float)
        if g.size == 0:
            continue
        c = g - g.mean()
        css = (c * c).sum()
        if css > 0:
            out[j] = (c ** 4).sum() / css ** 2
    return out


def test_kappa_matches_definition(oracle):
    """The oracle's raw-moment kurtosis equals the central-moment definition
    on random, missing-heavy, monomorphic and all-missing columns."""
    rng = np.random.default_rng(17)
    n, m = 1503, 60
    bed, codes, _, _, _ = random_panel(rng, n, m, miss=0.1)
    body = bed[3:].reshape(m, -1).copy()
    codes[3] = 3                      # monomorphic dosage 0
    codes[4] = 0                      # monomorphic dosage 2
    codes[5] = 1                      # all missing
    codes[6] = np.where(rng.random(n) < 0.5, 1, 2)  # dosage 1 or missing: no variance
    codes[7] = np.where(rng.random(n) < 0.002, 0, 3)  # rare allele
    body[:] = 0
    for i in range(n):
        body[:, i // 4] |= (codes[:, i] << (2 * (i % 4))).astype(np.uint8)
    bed = np.concatenate([bed[:3], body.reshape(-1)])
    got = oracle.kappa(bed, n, np.arange(m, dtype=np.uint64))
    exp = brute_kappa(dosages(codes))
    assert np.all(got[[3, 4, 5, 6]] == 0.0)
    assert np.max(np.abs(got - exp) / np.maximum(1.0, np.abs(exp))) < 1e-12
    # with --keep: the moments are over the kept individuals only
    keep = rng.permutation(n)[:700]
    got_k = oracle.kappa(bed, n, np.arange(m, dtype=np.uint64), keep=keep)
    exp_k = brute_kappa(dosages(codes[:, keep]))
    assert np.max(np.abs(got_k - exp_k) / np.maximum(1.0, np.abs(exp_k))) < 1e-12


def _quad_inv(t, d, D, kj=0.0, kk=0.0):
    """Inverse of the sketch bias model t = r^2 + (1 - r^2)(1 - 2 r^2)/d (with
    the MAF-aware d - 3 - 2(kj + kk) term), then the unbiased r^2; pinned
    below by feeding it the model's own output."""
    dm3 = d - 3.0 - 2.0 * (kj + kk)
    disc = dm3 * dm3 + 8.0 * (d * t - 1.0)
    rs = np.where(disc >= 0, (-dm3 + np.sqrt(np.maximum(disc, 0))) / 4.0, 0.0)
    return rs * (1 + 1 / D) - 1 / D


@pytest.mark.parametrize("maf_aware", [False, True])
def test_sketch_chain_with_injective_projection(oracle, maf_aware):
    """With an injective projection (each individual alone in its bucket,
    random signs, d > N so that d and N differ) the sketched column is a
    signed copy of the standardized column: no sketch noise.  The sketch
    path must then give exactly the window sum of the quadratic inversion
    applied to the exact r^2 -- which pins the scatter, the renormalization
    to ||x||^2 = N, the product scaling and the per-pair kurtosis wiring apart
    from the hash (whose stream stays unpinned)."""
    rng = np.random.default_rng(5 + maf_aware)
    n, m, d = 96, 48, 113
    bed, codes, chr_, cm, bp = random_panel(rng, n, m, miss=0.04)
    bucket = rng.permutation(d)[:n].astype(np.uint32)
    sign = np.where(rng.random(n) < 0.5, 1, -1).astype(np.int8)
    with oracle.projection(bucket, sign, d):
        got, _ = oracle.l2(bed, n, chr_, cm, bp, mode=1, window=30.0, chunk=7, masking=True,
                           use_f32=True, sketch=d, maf_aware=maf_aware)
    # the same call outside the hook is the hashed projection's range error
    with pytest.raises(ValueError):
        oracle.l2(bed, n, chr_, cm, bp, mode=1, window=30.0, chunk=7, masking=True,
                  use_f32=True, sketch=d)
    D = n - 2.0
    # the inversion used for the expectation is itself pinned to the model
    for r2 in np.linspace(0, 1, 11):
        t = r2 + (1 - r2) * (1 - 2 * r2) / d
        assert abs(_quad_inv(t, d, D) - (r2 * (1 + 1 / D) - 1 / D)) < 1e-12
    dos = dosages(codes)
    X = std_cols(dos)
    r2 = (X.T @ X / n) ** 2
    kap = brute_kappa(dos) if maf_aware else np.zeros(m)
    L = np.ones(m)
    for j in range(m):
        for i in range(j):
            if window_ok("kb", 30.0, chr_, cm, bp, i, j):
                v = _quad_inv(r2[i, j], float(d), D, kap[i], kap[j])
                L[j] += v
                L[i] += v
    assert np.max(np.abs(got[:, 0] - L)) < 2e-5
    # a planted mistake (the inversion run with N in place of d) is visible
    Lbad = np.ones(m)
    for j in range(m):
        for i in range(j):
            if window_ok("kb", 30.0, chr_, cm, bp, i, j):
                v = _quad_inv(r2[i, j], float(n), D, kap[i], kap[j])
                Lbad[j] += v
                Lbad[i] += v
    assert np.max(np.abs(got[:, 0] - Lbad)) > 1e-3
