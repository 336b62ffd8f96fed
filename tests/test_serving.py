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
    with pytest.raises(L.LdscError):
        Server(hbm_budget_bytes=0)
    with Server(max_inflight=2) as s:
        assert s.stats() == {"requests_done": 0, "band_launches": 0, "batched_requests": 0,
                             "max_batch": 0, "requests_cancelled": 0, "requests_timed_out": 0,
                             "hbm_charged": 0, "hbm_charged_max": 0, "workspace_trims": 0}
    with Server(max_inflight=2) as s:
        import ctypes as C
        err = N.errbuf()
        lat = C.c_double()
        assert N.lib().ldsc_l2_poll(s._h, 12345, C.byref(lat), err, 512) == 9
        assert N.lib().ldsc_l2_wait(s._h, 12345, C.byref(lat), err, 512) == 9
        assert N.lib().ldsc_l2_cancel(s._h, 12345, err, 512) == 9


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


TOL_EXACT, TOL_F32 = 1e-9, 2e-3


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b))))


def _oracle_l2(oracle, p, start, size, cfg, annot):
    mode = {N.WINDOW_CM: oracle.WINDOW_CM, N.WINDOW_KB: oracle.WINDOW_KB,
            N.WINDOW_SNP: oracle.WINDOW_SNP}[cfg.mode.kind]
    sl = slice(start, start + size)
    return oracle.l2(p.bed_bytes(), p.n, p.chr[sl], p.cm[sl], p.bp[sl],
                     np.arange(start, start + size, dtype=np.uint64), mode=mode,
                     window=cfg.mode.value, chunk=cfg.chunk_size, use_f32=cfg.use_f32,
                     masking=cfg.snp_level_masking, annot=annot, threads=16)


@pytest.mark.gpu
def test_mixed_stream_equals_standalone(oracle):
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
                jobs.append((panel, p, start, size, snps, cfg, annot,
                             srv.submit(panel, snps, cfg, annot)))
        print(f"submitted {len(jobs)} requests", flush=True)
        lats = [srv.wait(j[-1]) for j in jobs]
        for panel, p, start, size, snps, cfg, annot, t in jobs:
            alone, maf, _ = L.compute_ldscore(panel, snps, cfg, annot=annot)
            got, gmaf = t.result()
            assert np.array_equal(got, alone) and np.array_equal(gmaf, maf)
            # and each served result against the CPU restatement
            exp, emaf = _oracle_l2(oracle, p, start, size, cfg, annot)
            assert _rel(got, exp) < (TOL_F32 if cfg.use_f32 else TOL_EXACT)
            assert np.array_equal(gmaf, emaf)
        assert all(x > 0 for x in lats)
        st = srv.stats()
        assert st["hbm_charged_max"] <= 8 << 30


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


def _big_and_small(seed=9):
    p = synth.make_panel(6000, 40000, seed=seed, miss_rate=0.01)
    return p, L.make_snp_array(p.chr, p.cm, p.bp)


@pytest.mark.gpu
def test_cancel_queued_and_running(oracle):
    """ldsc_l2_cancel: a queued request ends at once as CANCELLED; a running
    one stops at a phase boundary (or finishes, if it was past the last);
    the server keeps serving correct results afterwards."""
    p, snps_all = _big_and_small()
    big = snps_all
    small = snps_all[1000:3000]
    cfg = L.L2Config(mode=L.WindowMode.Cm(1.0), snp_level_masking=True)
    with L.Panel(p.bed_bytes(), p.n, p.m) as panel, \
            Server(hbm_budget_bytes=16 << 30, max_inflight=1) as srv:
        ref_big, _, _ = L.compute_ldscore(panel, big, cfg)
        ref_small, _, _ = L.compute_ldscore(panel, small, cfg)
        # queued: one worker, busy with the first request
        t1 = srv.submit(panel, big, cfg)
        t2 = srv.submit(panel, small, cfg)
        t3 = srv.submit(panel, small, cfg)
        srv.cancel(t2)
        with pytest.raises(L.LdscError) as e:
            srv.wait(t2)
        assert e.value.code == N.ERR_CANCELLED
        srv.wait(t1)
        srv.wait(t3)
        assert np.array_equal(t1.result()[0], ref_big)
        assert np.array_equal(t3.result()[0], ref_small)
        # running: cancelled shortly after an idle worker took it
        import time
        outcomes = []
        for delay in (0.0005, 0.001, 0.002, 0.004, 0.0):
            t = srv.submit(panel, big, cfg)
            time.sleep(delay)
            srv.cancel(t)
            try:
                srv.wait(t)
                assert np.array_equal(t.result()[0], ref_big)
                outcomes.append("done")
            except L.LdscError as e:
                assert e.code == N.ERR_CANCELLED
                outcomes.append("cancelled")
        print(f"running-cancel outcomes: {outcomes}", flush=True)
        assert "cancelled" in outcomes
        # the workspace is sane after the cancellations
        t = srv.submit(panel, small, cfg)
        srv.wait(t)
        assert np.array_equal(t.result()[0], ref_small)
        exp, _ = oracle.l2(p.bed_bytes(), p.n, p.chr[1000:3000], p.cm[1000:3000],
                           p.bp[1000:3000], np.arange(1000, 3000, dtype=np.uint64),
                           mode=oracle.WINDOW_CM, window=1.0, chunk=200, masking=True,
                           threads=16)
        assert _rel(t.result()[0], exp) < TOL_EXACT
        st = srv.stats()
        assert st["requests_cancelled"] == 1 + outcomes.count("cancelled")


@pytest.mark.gpu
def test_queue_deadline():
    """max_queue_ms: a request still queued after the deadline fails with
    TIMEOUT; the one that was running finishes normally."""
    p, snps_all = _big_and_small(seed=10)
    cfg = L.L2Config(mode=L.WindowMode.Cm(1.0), snp_level_masking=True)
    with L.Panel(p.bed_bytes(), p.n, p.m) as panel, \
            Server(hbm_budget_bytes=16 << 30, max_inflight=1, max_queue_ms=1) as srv:
        ref, _, _ = L.compute_ldscore(panel, snps_all, cfg)
        t1 = srv.submit(panel, snps_all, cfg)
        t2 = srv.submit(panel, snps_all[:5000], cfg)
        import time
        time.sleep(0.05)
        with pytest.raises(L.LdscError) as e:
            srv.wait(t2)
        assert e.value.code == N.ERR_TIMEOUT
        srv.wait(t1)
        assert np.array_equal(t1.result()[0], ref)
        assert srv.stats()["requests_timed_out"] == 1


@pytest.mark.gpu
def test_budget_counts_retained_workspaces():
    """Workspaces keep their buffers between requests; the budget counts
    them.  With room for about one large request, a large one, a stream of
    small ones and a second large one on four workers: the charged HBM never
    exceeds the budget, a worker frees its buffers to let the second large
    request in, and every result is the standalone one."""
    p, snps_all = _big_and_small(seed=11)
    cfg = L.L2Config(mode=L.WindowMode.Cm(1.0), snp_level_masking=True)
    with L.Panel(p.bed_bytes(), p.n, p.m) as panel:
        ref_big, _, _ = L.compute_ldscore(panel, snps_all, cfg)
        # the large request's own estimate sets the budget: 1.5 of it
        with Server(hbm_budget_bytes=1 << 40, max_inflight=1) as probe:
            t = probe.submit(panel, snps_all, cfg)
            probe.wait(t)
            est = probe.stats()["hbm_charged_max"]
        budget = int(est * 1.5)
        print(f"large request estimate {est} bytes, budget {budget}", flush=True)
        rng = np.random.default_rng(3)
        with Server(hbm_budget_bytes=budget, max_inflight=4) as srv:
            jobs = [(snps_all, srv.submit(panel, snps_all, cfg))]
            for _ in range(12):
                a = int(rng.integers(0, 38000))
                sn = snps_all[a:a + 1500]
                jobs.append((sn, srv.submit(panel, sn, cfg)))
            jobs.append((snps_all, srv.submit(panel, snps_all, cfg)))
            for _, t in jobs:
                srv.wait(t)
            st = srv.stats()
            print(st, flush=True)
            assert st["hbm_charged_max"] <= budget
            assert st["hbm_charged"] <= budget
            assert st["workspace_trims"] >= 1
            for sn, t in jobs:
                if sn is snps_all:
                    assert np.array_equal(t.result()[0], ref_big)
                else:
                    alone, _, _ = L.compute_ldscore(panel, sn, cfg)
                    assert np.array_equal(t.result()[0], alone)
