"""ldsc_l2_compute_noalloc: the allocation-free variant of ldsc_l2_compute.

Its buffers come from one fixed device array; its kernels are those of
ldsc_l2_compute, so its results must be bit-identical to it -- checked on
two calls in a row, the second with a smaller input, in every arithmetic
form, with annotations and with a --keep panel.  The device's free memory
does not change across a call once the entry point is set up, and calls
beyond the fixed size fail with a clear error."""
import ctypes as C

import numpy as np
import pytest

import ldsc_amd as L
from ldsc_amd import _native as N
from ldsc_amd import synth

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 256 << 20


def _noalloc(panel, snps, cfg, annot=None):
    snps = np.ascontiguousarray(snps, dtype=N.SNP_DTYPE)
    m = snps.size
    a, K = None, 1
    if annot is not None:
        a = np.asfortranarray(np.asarray(annot, np.float64).reshape(m, -1))
        K = a.shape[1]
    l2 = np.zeros((K, m), np.float64)
    maf = np.zeros(m, np.float64)
    err = N.errbuf()
    c = cfg.to_c()
    rc = N.lib().ldsc_l2_compute_noalloc(panel.handle, N.ptr(snps), m, C.byref(c),
                                         None if a is None else C.c_void_p(a.ctypes.data), K,
                                         N.ptr(l2), N.ptr(maf), None, N.PROGRESS_CB(), None,
                                         err, 512)
    N.check(rc, err)
    return l2.T, maf


def _free_device_bytes():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


@pytest.mark.parametrize("form", ["exact", "exact_i8", "f32"])
def test_noalloc_matches_compute_two_calls(form):
    p = synth.make_panel(1500, 12000, seed=21, miss_rate=0.01)
    snps = L.make_snp_array(p.chr, p.cm, p.bp)
    annot = synth.make_annot(p.m, seed=4, n_bin=2, n_cont=1)
    mk = dict(use_f32=form == "f32", exact_kernel="i8" if form == "exact_i8" else "auto")
    cfg1 = L.L2Config(mode=L.WindowMode.Cm(1.0), chunk_size=50, **mk)
    cfg2 = L.L2Config(mode=L.WindowMode.Kb(300.0), snp_level_masking=True, **mk)
    print(f"noalloc form={form}", flush=True)
    with L.Panel(p.bed_bytes(), p.n, p.m) as panel:
        # two calls in a row, the second smaller (and with annotations)
        got1, maf1 = _noalloc(panel, snps, cfg1)
        got2, maf2 = _noalloc(panel, snps[2000:6000], cfg2, annot[2000:6000])
        exp1, emaf1, _ = L.compute_ldscore(panel, snps, cfg1)
        exp2, emaf2, _ = L.compute_ldscore(panel, snps[2000:6000], cfg2, annot[2000:6000])
        assert np.array_equal(got1, exp1) and np.array_equal(maf1, emaf1)
        assert np.array_equal(got2, exp2) and np.array_equal(maf2, emaf2)
        # once set up, a call leaves the device's free memory unchanged
        before = _free_device_bytes()
        got3, _ = _noalloc(panel, snps[:9000], cfg1)
        after = _free_device_bytes()
        assert before == after
        exp3, _, _ = L.compute_ldscore(panel, snps[:9000], cfg1)
        assert np.array_equal(got3, exp3)


def test_noalloc_keep_panel():
    p = synth.make_panel(900, 5000, seed=22, miss_rate=0.02)
    keep = np.sort(np.random.default_rng(1).permutation(900)[:611])[::-1].copy()
    snps = L.make_snp_array(p.chr, p.cm, p.bp)
    cfg = L.L2Config(mode=L.WindowMode.Snp(150), chunk_size=64)
    with L.Panel(p.bed_bytes(), p.n, p.m, keep=keep) as panel:
        for sl in (slice(0, 5000), slice(100, 1300)):
            got, maf = _noalloc(panel, snps[sl], cfg)
            exp, emaf, _ = L.compute_ldscore(panel, snps[sl], cfg)
            assert np.array_equal(got, exp) and np.array_equal(maf, emaf)


def test_noalloc_rejects_sketch_and_oversized_calls():
    p = synth.make_panel(40000, 8000, seed=23, miss_rate=0.01)
    snps = L.make_snp_array(p.chr, p.cm, p.bp)
    with L.Panel(p.bed_bytes(), p.n, p.m) as panel:
        with pytest.raises(L.LdscError) as e:
            _noalloc(panel, snps[:500], L.L2Config(mode=L.WindowMode.Cm(1.0), sketch=100))
        assert e.value.code == 1 and "CountSketch" in str(e.value)
        # 8,000 SNPs x 40,000 individuals: an 80 MB slice, well within
        cfg = L.L2Config(mode=L.WindowMode.Cm(1.0), snp_level_masking=True)
        got, _ = _noalloc(panel, snps, cfg)
        exp, _, _ = L.compute_ldscore(panel, snps, cfg)
        assert np.array_equal(got, exp)
        # with 400 annotation columns the tile partials alone pass 256 MiB
        annot = np.ones((p.m, 400))
        with pytest.raises(L.LdscError) as e:
            _noalloc(panel, snps, cfg, annot)
        assert e.value.code == 7 and "scratch" in str(e.value)
        # the entry point still works after the rejection
        got, _ = _noalloc(panel, snps[:3000], cfg)
        exp, _, _ = L.compute_ldscore(panel, snps[:3000], cfg)
        assert np.array_equal(got, exp)
