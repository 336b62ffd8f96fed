"""Request server over resident panels (``ldsc_server_*`` of the C ABI).

Continuous batching: up to ``max_inflight`` requests run concurrently on
their own HIP streams, admitted first-fit against an HBM budget, and their
kernels overlap on the device.  With ``batch_window_us`` > 0 the band
contractions of requests prepared at the same time also share one kernel
launch (a request waits at most that long for others still being
prepared); on the mixed 1,000-50,000-SNP request stream of bench.py that
measured 10-15 % below one launch per request, so it is off by default.
Results of a request do not depend on what else is in flight or shares its
launch.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _native as N
from .l2 import L2Config, Panel


@dataclass
class Ticket:
    id: int
    l2: np.ndarray          # K x m buffer written by the device engine
    maf: np.ndarray
    snps: np.ndarray        # kept alive until the request is done
    annot: Optional[np.ndarray]

    def result(self):
        return self.l2.T, self.maf


class Server:
    def __init__(self, hbm_budget_bytes: int = 64 << 30, max_inflight: int = 4,
                 queue_capacity: int = 1024, batch_window_us: int = 0, max_queue_ms: int = 0):
        L = N.lib()
        cfg = N.ServerConfig(int(hbm_budget_bytes), int(max_inflight), int(queue_capacity),
                             int(batch_window_us), int(max_queue_ms))
        h = C.c_void_p()
        err = N.errbuf()
        N.check(L.ldsc_server_create(C.byref(cfg), C.byref(h), err, 512), err)
        self._h = h

    def submit(self, panel: Panel, snps: np.ndarray, config: L2Config, annot=None) -> Ticket:
        L = N.lib()
        snps = np.ascontiguousarray(snps, dtype=N.SNP_DTYPE)
        m = snps.size
        a, K = None, 1
        if annot is not None:
            a = np.asarray(annot, dtype=np.float64)
            if a.ndim == 1:
                a = a[:, None]
            K = a.shape[1]
            a = np.asfortranarray(a)
        l2 = np.zeros((K, m), np.float64)
        maf = np.zeros(m, np.float64)
        rid = C.c_uint64()
        err = N.errbuf()
        cfg = config.to_c()
        N.check(L.ldsc_l2_submit(self._h, panel.handle, N.ptr(snps), m, C.byref(cfg),
                                 None if a is None else C.c_void_p(a.ctypes.data), K,
                                 N.ptr(l2), N.ptr(maf), C.byref(rid), err, 512), err)
        return Ticket(rid.value, l2, maf, snps, a)

    def poll(self, t: Ticket) -> Optional[float]:
        """Latency in seconds when done, None while pending; raises on error."""
        lat = C.c_double()
        err = N.errbuf()
        rc = N.lib().ldsc_l2_poll(self._h, t.id, C.byref(lat), err, 512)
        if rc == N.ERR_PENDING:
            return None
        N.check(rc, err)
        return lat.value

    def wait(self, t: Ticket) -> float:
        lat = C.c_double()
        err = N.errbuf()
        N.check(N.lib().ldsc_l2_wait(self._h, t.id, C.byref(lat), err, 512), err)
        return lat.value

    def cancel(self, t: Ticket) -> None:
        """Cancel a queued or running request (ldsc_l2_cancel); its poll / wait
        then raises LdscError with code ERR_CANCELLED unless it had finished."""
        err = N.errbuf()
        N.check(N.lib().ldsc_l2_cancel(self._h, t.id, err, 512), err)

    def stats(self) -> dict:
        """Counters since creation: requests done, band launches, requests
        whose band shared a launch, the largest launch (in requests),
        cancelled and timed-out requests, the HBM charged to the workers
        (now and at most) and how often a worker freed its buffers."""
        st = N.ServerStats()
        N.check(N.lib().ldsc_server_stats_get(self._h, C.byref(st)), N.errbuf())
        return {"requests_done": st.requests_done, "band_launches": st.band_launches,
                "batched_requests": st.batched_requests, "max_batch": st.max_batch,
                "requests_cancelled": st.requests_cancelled,
                "requests_timed_out": st.requests_timed_out, "hbm_charged": st.hbm_charged,
                "hbm_charged_max": st.hbm_charged_max, "workspace_trims": st.workspace_trims}

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            N.lib().ldsc_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
