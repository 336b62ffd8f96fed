"""Request server: host-side errors (CPU) and continuous batching on the
device (GPU): each result is bit-identical alone, in the mix and in a
shared band launch."""
import numpy as np
import pytest

import ldsc_amd as L
from ldsc_amd import _native as N
from ldsc_amd import synth
from ldsc_amd.serve import Server


def test_server_create_errors_and_unknown_ids():
    with pytest.raises(L.LdscError):
        Server(max_inflight=0)
    with pytest.raises(L.LdscError):
        Server(queue_capacity=0)
    with pytest.raises(L.LdscError):
        Server(batch_window_us=2_000_000)
    with Server(max_inflight=2) as s:
        assert s.stats() == {"requests_done": 0, "band_launches": 0, "batched_requests": 0,
                             "max_batch": 0}
    with Server(max_inflight=2) as s:
        import ctypes as C
        err = N.errbuf()
        lat = C.c_double()
        assert N.lib().ldsc_l2_poll(s._h, 12345, C.byref(lat), err, 512) == 9
        assert N.lib().ldsc_l2_wait(s._h, 12345, C.byref(lat), err, 512) == 9


def _requests(rng, m_panel, k_max=4):
    out = []
    for _ in range(14):
        size = int(np.exp(rng.uniform(np.log(200), np.log(3000))))
        start = int(rng.integers(0, m_panel - size))
        kind = int(rng.integers(0, 3))
        mode = [L.WindowMode.Cm(0.3), L.WindowMode.Kb(300.0), L.WindowMode.Snp(150)][kind]
        K = int(rng.integers(0, k_max + 1))
        out.append((start, size, mode, K, bool(rng.integers(0, 2)), bool(rng.integers(0, 2))))
    return out


@pytest.mark.gpu
def test_mixed_stream_equals_standalone():
    rng = np.random.default_rng(7)
    pa = synth.make_panel(3000, 6000, seed=1, miss_rate=0.01)
    pb = synth.make_panel(1200, 5000, seed=2, miss_rate=0.01)
    with L.Panel(pa.bed_bytes(), pa.n, pa.m) as A, L.Panel(pb.bed_bytes(), pb.n, pb.m) as B, \
            Server(hbm_budget_bytes=8 << 30, max_inflight=4) as srv:
        jobs = []
        for panel, p in ((A, pa), (B, pb)):
            snps_all = L.make_snp_array(p.chr, p.cm, p.bp)
            for start, size, mode, K, f32, mask in _requests(rng, p.m):
                snps = snps_all[start:start + size]
                annot = None if K == 0 else synth.make_annot(size, seed=start, n_bin=K, n_cont=0)
                if annot is not None and K > 1:
                    annot[:, -1] = rng.lognormal(size=size)
                cfg = L.L2Config(mode=mode, chunk_size=200, use_f32=f32, snp_level_masking=mask)
                jobs.append((panel, snps, cfg, annot, srv.submit(panel, snps, cfg, annot)))
        print(f"submitted {len(jobs)} requests", flush=True)
        lats = [srv.wait(t) for (_, _, _, _, t) in jobs]
        for panel, snps, cfg, annot, t in jobs:
            alone, maf, _ = L.compute_ldscore(panel, snps, cfg, annot=annot)
            got, gmaf = t.result()
            assert np.array_equal(got, alone) and np.array_equal(gmaf, maf)
        assert all(x > 0 for x in lats)


@pytest.mark.gpu
@pytest.mark.parametrize("window_us", [0, 50_000])
def test_batched_band_launches_equal_standalone(window_us):
    """A burst of small requests on eight workers: with a batch window their
    band tiles share launches (the leader waits for the workers still
    preparing), without one every request has its own launch.  Either way
    each result equals the standalone call bit for bit."""
    rng = np.random.default_rng(11)
    p = synth.make_panel(2000, 8000, seed=5, miss_rate=0.01)
    snps_all = L.make_snp_array(p.chr, p.cm, p.bp)
    with L.Panel(p.bed_bytes(), p.n, p.m) as panel, \
            Server(hbm_budget_bytes=8 << 30, max_inflight=8, batch_window_us=window_us) as srv:
        jobs = []
        for i in range(24):
            size = int(rng.integers(300, 2500))
            start = int(rng.integers(0, p.m - size))
            K = int(rng.integers(0, 3))
            snps = snps_all[start:start + size]
            annot = None if K == 0 else synth.make_annot(size, seed=i, n_bin=K, n_cont=0)
            cfg = L.L2Config(mode=[L.WindowMode.Cm(0.2), L.WindowMode.Snp(120)][i % 2],
                             chunk_size=100, snp_level_masking=bool(i % 3),
                             exact_kernel="i8" if i % 5 == 4 else "auto")
            jobs.append((snps, cfg, annot, srv.submit(panel, snps, cfg, annot)))
        for (_, _, _, t) in jobs:
            srv.wait(t)
        st = srv.stats()
        print(f"window {window_us} us: {st}", flush=True)
        for snps, cfg, annot, t in jobs:
            alone, maf, _ = L.compute_ldscore(panel, snps, cfg, annot=annot)
            got, gmaf = t.result()
            assert np.array_equal(got, alone) and np.array_equal(gmaf, maf)
    assert st["requests_done"] == 24
    if window_us:
        assert st["batched_requests"] > 0 and st["max_batch"] > 1
        assert st["band_launches"] < 24
    else:
        assert st["band_launches"] == 24 and st["batched_requests"] == 0


@pytest.mark.gpu
def test_admission_budget_and_queue_full():
    p = synth.make_panel(4000, 4000, seed=3)
    snps = L.make_snp_array(p.chr, p.cm, p.bp)
    cfg = L.L2Config(mode=L.WindowMode.Cm(0.5), snp_level_masking=True)
    with L.Panel(p.bed_bytes(), p.n, p.m) as panel:
        with Server(hbm_budget_bytes=1 << 20, max_inflight=1) as srv:
            with pytest.raises(L.LdscError) as e:
                srv.submit(panel, snps, cfg)
            assert e.value.code == 7
        with Server(hbm_budget_bytes=8 << 30, max_inflight=1, queue_capacity=2) as srv:
            tickets, busy = [], 0
            for _ in range(12):
                try:
                    tickets.append(srv.submit(panel, snps, cfg))
                except L.LdscError as e:
                    assert e.code == 8
                    busy += 1
            assert busy > 0
            ref, _, _ = L.compute_ldscore(panel, snps, cfg)
            for t in tickets:
                srv.wait(t)
                assert np.array_equal(t.result()[0], ref)
