"""The five benchmark workloads at their full shapes on the device.

* C1 (N=2,490, M=20,000, cM 1, chunk 50, unmasked) is
  test_gpu_l2.py::test_cpu_runnable_config, against the restatement.
* C2 (N=50,000, M=100,000, 1 % missing, cM 1, chunk 200, masked): the whole
  panel against the FP64 restatement, the e2m1 and int8 exact kernels
  against each other, the FP32 form within its tolerance, bit-reproducible.
* C3 (N=100,000, M=200,000, d in {200, 1000, 5000}): integer buckets
  bit-exact, sketch LD scores on a slice against the restatement (also
  with annotations), and at full size with 48 annotations the segmented
  upload equals the one-piece path.
* C4 (N=20,000, M=500,000, K=50, kb 1000): a slice with all 50 annotations
  against the restatement; at full size the two exact kernels agree and the
  segmented upload equals the one-piece path.
* C5 (resident 50,000 x 1,000,000 and 10,000 x 500,000 panels): a mixed
  request stream served concurrently equals the standalone calls, and the
  smaller requests equal the restatement.
Panels come from the seeded founder-haplotype generator (seed 20261203 + k
for config k).
"""
import numpy as np
import pytest

import ldsc_amd as L
from ldsc_amd import synth
from ldsc_amd.serve import Server

pytestmark = pytest.mark.gpu

TOL_EXACT = 1e-9     # exact integer products + FP64 finish vs the FP64 restatement
TOL_KERNELS = 1e-12  # e2m1 vs int8 exact kernels (same integers, summation order)
TOL_F32 = 2e-3       # FP32 form vs the f32 restatement / the exact form
TOL_SKETCH = 2e-5    # exact-integer sketch vs the f32 sequential-scatter restatement


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b))))


def _oracle(oracle, p, sl, **kw):
    idx = np.arange(sl.start, sl.stop, dtype=np.uint64)
    return oracle.l2(p.bed_bytes(), p.n, p.chr[sl], p.cm[sl], p.bp[sl], idx, threads=16, **kw)


@pytest.fixture(scope="class")
def c2():
    print("C2 panel N=50000 M=100000", flush=True)
    p = synth.make_panel(50_000, 100_000, seed=20261205, miss_rate=0.01)
    with L.Panel(p.bed_bytes(), p.n, p.m) as panel:
        yield p, panel, L.make_snp_array(p.chr, p.cm, p.bp)


class TestC2:
    base = dict(mode=L.WindowMode.Cm(1.0), chunk_size=200, snp_level_masking=True)

    def test_c2_exact_whole_panel_vs_fp64_oracle(self, oracle, c2):
        p, panel, snps = c2
        got, maf, perf = L.compute_ldscore(panel, snps, L.L2Config(**self.base))
        bl, _ = L.block_left(snps, L.WindowMode.Cm(1.0))
        assert perf.pairs_included == int(np.sum(np.arange(p.m) - bl.astype(np.int64)))
        print(f"C2 exact done, {perf.pairs_included} pairs; oracle f64 over 100,000 SNPs",
              flush=True)
        exp, emaf = _oracle(oracle, p, slice(0, p.m), mode=oracle.WINDOW_CM, window=1.0,
                            chunk=200, masking=True)
        err = _rel(got, exp)
        print(f"C2 exact vs FP64 restatement: max rel {err:.3e}", flush=True)
        assert err < TOL_EXACT
        assert np.array_equal(maf, emaf)

    def test_c2_kernels_fp32_and_reproducible(self, c2):
        p, panel, snps = c2
        a, _, _ = L.compute_ldscore(panel, snps, L.L2Config(**self.base))
        a2, _, _ = L.compute_ldscore(panel, snps, L.L2Config(**self.base))
        b, _, _ = L.compute_ldscore(panel, snps, L.L2Config(**self.base, exact_kernel="i8"))
        f, _, _ = L.compute_ldscore(panel, snps, L.L2Config(**self.base, use_f32=True))
        assert np.array_equal(a, a2)
        assert _rel(b, a) < TOL_KERNELS
        assert _rel(f, a) < TOL_F32

    def test_c2_fp32_slice_vs_f32_oracle(self, oracle, c2):
        p, panel, snps = c2
        sl = slice(40_000, 44_000)
        got, maf, _ = L.compute_ldscore(panel, snps[sl], L.L2Config(**self.base, use_f32=True))
        exp, emaf = _oracle(oracle, p, sl, mode=oracle.WINDOW_CM, window=1.0, chunk=200,
                            masking=True, use_f32=True)
        assert _rel(got, exp) < TOL_F32
        assert np.array_equal(maf, emaf)


@pytest.fixture(scope="class")
def c3():
    print("C3 panel N=100000 M=200000", flush=True)
    p = synth.make_panel(100_000, 200_000, seed=20261206, miss_rate=0.01)
    with L.Panel(p.bed_bytes(), p.n, p.m) as panel:
        yield p, panel, L.make_snp_array(p.chr, p.cm, p.bp)


class TestC3:
    @pytest.mark.parametrize("d", [200, 1000, 5000])
    def test_c3_buckets_bit_exact(self, oracle, c3, d):
        p, panel, _ = c3
        idx = np.array(sorted(set(np.linspace(0, p.m - 1, 40).astype(int)) | {0, 1, p.m - 1}),
                       np.uint64)
        sg, sm = L.sketch_buckets(panel, idx, d)
        esg, esm = oracle.sketch_int(p.bed_bytes(), p.n, idx, d)
        assert np.array_equal(sg, esg) and np.array_equal(sm, esm)

    @pytest.mark.parametrize("d", [200, 1000, 5000])
    def test_c3_sketch_slice_vs_oracle(self, oracle, c3, d):
        p, panel, snps = c3
        sl = slice(120_000, 120_600)
        cfg = L.L2Config(mode=L.WindowMode.Cm(1.0), sketch=d, snp_level_masking=True)
        got, maf, _ = L.compute_ldscore(panel, snps[sl], cfg)
        exp, emaf = _oracle(oracle, p, sl, mode=oracle.WINDOW_CM, window=1.0, chunk=200,
                            masking=True, use_f32=True, sketch=d)
        err = _rel(got, exp)
        print(f"C3 d={d} slice vs restatement: {err:.3e}", flush=True)
        assert err < TOL_SKETCH
        assert np.array_equal(maf, emaf)

    def test_c3_sketch_annotations_slice_vs_oracle(self, oracle, c3):
        p, panel, snps = c3
        sl = slice(7_000, 7_500)
        annot = synth.make_annot(500, seed=3, n_bin=4, n_cont=2)
        cfg = L.L2Config(mode=L.WindowMode.Cm(1.0), sketch=1000, snp_level_masking=True,
                         sketch_maf_aware=True)
        got, _, _ = L.compute_ldscore(panel, snps[sl], cfg, annot)
        exp, _ = _oracle(oracle, p, sl, mode=oracle.WINDOW_CM, window=1.0, chunk=200,
                         masking=True, use_f32=True, sketch=1000, maf_aware=True, annot=annot)
        assert _rel(got, exp) < TOL_SKETCH

    def test_c3_full_segmented_annotations_equal_one_piece(self, c3):
        """m x K x 8 = 76.8 MB goes up in two row segments on the one-call
        path; the request server uploads it in one piece.  Same tiles, same
        arithmetic, same summation order: bit-identical."""
        p, panel, snps = c3
        annot = synth.make_annot(p.m, seed=9, n_bin=40, n_cont=8)
        cfg = L.L2Config(mode=L.WindowMode.Cm(1.0), sketch=200, snp_level_masking=True)
        seg, maf, _ = L.compute_ldscore(panel, snps, cfg, annot)
        with Server(hbm_budget_bytes=64 << 30, max_inflight=1) as srv:
            t = srv.submit(panel, snps, cfg, annot)
            srv.wait(t)
            one, omaf = t.result()
        assert np.array_equal(seg, one) and np.array_equal(maf, omaf)
        again, _, _ = L.compute_ldscore(panel, snps, cfg, annot)
        assert np.array_equal(again, seg)


@pytest.fixture(scope="class")
def c4():
    print("C4 panel N=20000 M=500000 K=50", flush=True)
    p = synth.make_panel(20_000, 500_000, seed=20261207, miss_rate=0.01)
    annot = synth.make_annot(p.m, seed=20261207, n_bin=40, n_cont=10)
    with L.Panel(p.bed_bytes(), p.n, p.m) as panel:
        yield p, panel, L.make_snp_array(p.chr, p.cm, p.bp), annot


class TestC4:
    base = dict(mode=L.WindowMode.Kb(1000.0), chunk_size=200)

    @pytest.mark.parametrize("use_f32", [False, True])
    def test_c4_slice_k50_vs_oracle(self, oracle, c4, use_f32):
        p, panel, snps, annot = c4
        sl = slice(250_000, 253_000)
        got, maf, _ = L.compute_ldscore(panel, snps[sl], L.L2Config(**self.base, use_f32=use_f32),
                                        annot[sl])
        exp, emaf = _oracle(oracle, p, sl, mode=oracle.WINDOW_KB, window=1000.0, chunk=200,
                            use_f32=use_f32, annot=annot[sl])
        assert got.shape == (3000, 50)
        assert _rel(got, exp) < (TOL_F32 if use_f32 else TOL_EXACT)
        assert np.array_equal(maf, emaf)

    def test_c4_full_kernels_agree_and_segments_equal_one_piece(self, c4):
        p, panel, snps, annot = c4
        a, maf, _ = L.compute_ldscore(panel, snps, L.L2Config(**self.base), annot)
        b, _, _ = L.compute_ldscore(panel, snps, L.L2Config(**self.base, exact_kernel="i8"),
                                    annot)
        assert _rel(b, a) < TOL_KERNELS
        with Server(hbm_budget_bytes=64 << 30, max_inflight=1) as srv:
            t = srv.submit(panel, snps, L.L2Config(**self.base), annot)
            srv.wait(t)
            one, omaf = t.result()
        assert np.array_equal(a, one) and np.array_equal(maf, omaf)


class TestC5:
    def test_c5_mixed_stream_on_resident_panels(self, oracle):
        """Both C5 panels resident; 32 requests of log-uniform size in
        [1,000, 50,000] SNPs, window type uniform over snps 500 / kb 1000 /
        cM 1, 0-10 annotation columns, served four at a time.  Each equals
        its standalone call bit for bit; those of at most 4,000 SNPs also
        equal the restatement."""
        print("C5 panels 50000 x 1000000 and 10000 x 500000", flush=True)
        pa = synth.make_panel(50_000, 1_000_000, seed=20261208, miss_rate=0.01)
        pb = synth.make_panel(10_000, 500_000, seed=20261209, miss_rate=0.01)
        rng = np.random.default_rng(20261208)
        kinds = [(L.WindowMode.Snp(500), oracle.WINDOW_SNP, 500.0),
                 (L.WindowMode.Kb(1000.0), oracle.WINDOW_KB, 1000.0),
                 (L.WindowMode.Cm(1.0), oracle.WINDOW_CM, 1.0)]
        with L.Panel(pa.bed_bytes(), pa.n, pa.m) as A, L.Panel(pb.bed_bytes(), pb.n, pb.m) as B, \
                Server(hbm_budget_bytes=64 << 30, max_inflight=4) as srv:
            jobs = []
            for i in range(32):
                p, panel = (pa, A) if i % 2 == 0 else (pb, B)
                size = int(np.exp(rng.uniform(np.log(1000), np.log(50000))))
                start = int(rng.integers(0, p.m - size))
                kind = kinds[int(rng.integers(0, 3))]
                K = int(rng.integers(0, 11))
                snps = L.make_snp_array(p.chr[start:start + size], p.cm[start:start + size],
                                        p.bp[start:start + size])
                snps["bed_idx"] = np.arange(start, start + size, dtype=np.uint64)
                annot = None if K == 0 else synth.make_annot(size, seed=i, n_bin=max(K - 2, 0),
                                                             n_cont=min(K, 2))
                cfg = L.L2Config(mode=kind[0], chunk_size=200)
                jobs.append((p, panel, start, size, kind, snps, cfg, annot,
                             srv.submit(panel, snps, cfg, annot)))
            for j in jobs:
                srv.wait(j[-1])
            print(f"C5 stream done: {srv.stats()}", flush=True)
            checked = 0
            for p, panel, start, size, kind, snps, cfg, annot, t in jobs:
                got, maf = t.result()
                alone, amaf, _ = L.compute_ldscore(panel, snps, cfg, annot)
                assert np.array_equal(got, alone) and np.array_equal(maf, amaf)
                if size <= 4000:
                    exp, emaf = _oracle(oracle, p, slice(start, start + size), mode=kind[1],
                                        window=kind[2], chunk=200, annot=annot)
                    assert _rel(got, exp) < TOL_EXACT
                    assert np.array_equal(maf, emaf)
                    checked += 1
            print(f"C5: {checked} requests checked against the restatement", flush=True)
            assert checked >= 4
