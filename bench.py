"""LD-score throughput and request latency on one MI355X.

Default run (one JSON line):
  * headline -- exact LD scores at panel scale: synthetic founder-haplotype
    PLINK panel, N=50,000 individuals x M=100,000 SNPs, 1 % missing,
    --ld-wind-cm 1, --chunk-size 200, --snp-level-masking.  A step is one full
    ldsc_l2_compute over the resident panel.  `value` = SNP-pairs/s of the
    exact form (exact integer products G.G, G.M, M.G, M.M + FP64 finish),
    whose products run on the e2m1 MFMA with f32 accumulation; the same form
    on the int8 MFMA (int32 accumulation, identical results) and the plain
    FP32 form run beside it.
  * serving -- resident panels N=50,000 x M=1,000,000 and N=10,000 x
    M=500,000; Poisson arrivals of contiguous SNP-range requests (size
    log-uniform in [1,000, 50,000], window snps/kb/cM uniform, 0-10
    annotation columns); p50/p99 arrival->result latency and aggregate
    SNP-pairs/s.
  * cpu_baseline -- the C restatement of the reference's f64 path on the
    host cores, on a bounded slice of the same panel.

  * sketch -- CountSketch LD scores, N=100,000 x M=200,000, cM 1, masked,
    d in {200, 1000, 5000}: SNP-pairs/s, the preparation (decode, bucket
    sums) against HBM bandwidth and the d-length contraction against its
    MFMA peak, and the restatement of the reference's sketch path on a slice.
  * partitioned -- N=20,000 x M=500,000, K=50 annotations, kb 1000: the
    exact form on both exact kernels and the FP32 form, with the band's
    roofline and the restatement of the f64 path on a slice.

`--workload sketch` / `--workload partitioned` print those legs alone;
`--workload reference_shapes` (the panel shapes of the reference's published
CPU wall times) and `--workload cli` (the `ldsc l2` command end to end, from
the PLINK files to the written LD-score files, at the 1000G shape) print
their own JSON lines.  `--gpus` above 1 is refused: the engine runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

INT8_PEAK_TOPS = 5000.0    # MI355X dense int8 MFMA, TOP/s
FP4_PEAK_TFLOPS = 10100.0  # MI355X dense FP4 (e2m1) MFMA, TFLOP/s
FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 matrix, TFLOP/s
HBM_PEAK_GBS = 8000.0      # MI355X HBM3E, GB/s
METRIC = "SNP-pairs/s int8-MFMA vs FP32 plain (≤1e-6 rel); p50/p99 LD-request latency"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


PANEL_PROFILE = "profiles/k_band_fp4_50k_x_100k.md"


def profiled(key: str, kernel: str, path: str = PANEL_PROFILE):
    """A figure of `kernel` from a committed rocprofv3 summary: its lines
    `<key> <kernel> <value>`, e.g. traffic_bytes_per_launch (FETCH_SIZE and
    WRITE_SIZE scaled by a copy-kernel calibration) or trace_avg_ms (the
    kernel-trace average duration of the same launches), or None."""
    try:
        for line in open(os.path.join(ROOT, path)):
            parts = line.split()
            if len(parts) == 3 and parts[0] == key and parts[1] == kernel:
                return float(parts[2])
    except OSError:
        pass
    return None


def profiled_traffic(kernel: str, path: str = PANEL_PROFILE):
    v = profiled("traffic_bytes_per_launch", kernel, path)
    return None if v is None else int(v)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def time_cpu(fn, runs=3):
    """Median wall of `runs` calls after one warm-up, and the last result."""
    fn()
    walls, out = [], None
    for _ in range(runs):
        t = time.perf_counter()
        out = fn()
        walls.append(time.perf_counter() - t)
    return float(np.median(walls)), out


def measured_peaks():
    """Sustained MFMA rates (e2m1, int8, FP32; one wave per SIMD on every CU)
    and device-to-device copy bandwidth measured on this GPU with the
    calibration kernels of ldsc_amd/csrc/calib.hip."""
    import ctypes as C

    lib = C.CDLL(os.path.join(ROOT, "ldsc_amd", "libldsc_calib.so"))
    dp = C.POINTER(C.c_double)
    lib.ldsc_calib_mfma.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp]
    lib.ldsc_calib_copy.argtypes = [C.c_int, C.c_uint64, C.c_int, dp]
    out = {}
    for name, form, iters in (("fp4", 0, 100_000), ("i8", 1, 100_000), ("f32", 2, 20_000)):
        rate, ghz = C.c_double(), C.c_double()
        rc = lib.ldsc_calib_mfma(0, form, iters, C.byref(rate), C.byref(ghz))
        if rc:
            raise RuntimeError(f"ldsc_calib_mfma({name}) failed: hipError {rc}")
        out[name] = {"tflops": rate.value, "clock_ghz": ghz.value}
    gbs = C.c_double()
    rc = lib.ldsc_calib_copy(0, 2 << 30, 5, C.byref(gbs))
    if rc:
        raise RuntimeError(f"ldsc_calib_copy failed: hipError {rc}")
    out["copy_gbs"] = gbs.value
    log(f"measured peaks: {out}")
    return out


def timed_steps(L, panel, snps, cfg, steps, warmup, tag, annot=None, dest=None):
    """`dest`: an output array reused by every call (a serving loop's
    buffer), or None for a fresh result array per call."""
    for w in range(warmup):
        L.compute_ldscore(panel, snps, cfg, annot=annot, out=dest)
        log(f"{tag} warmup {w + 1}/{warmup} done")
    walls, perfs, out = [], [], None
    for s in range(steps):
        t = time.perf_counter()
        l2, maf, perf = L.compute_ldscore(panel, snps, cfg, annot=annot, out=dest)
        walls.append(time.perf_counter() - t)
        perfs.append(perf)
        out = (l2, maf)
        log(f"{tag} step {s + 1}/{steps}: {walls[-1] * 1e3:.2f} ms (band kernel "
            f"{perf.bb_dot_secs * 1e3:.2f} ms, tile {perf.norm_secs * 1e3:.2f} ms, sketch "
            f"{(perf.sketch_secs or 0) * 1e3:.2f} ms)")
    return walls, perfs, out


# ---------------------------------------------------------------------------
# Headline: exact LD scores at panel scale
# ---------------------------------------------------------------------------
def leg_panel(args, L, synth):
    seed = 20261205
    t0 = time.perf_counter()
    p = synth.make_panel(args.n, args.m, seed, miss_rate=0.01)
    log(f"panel: generated N={p.n} M={p.m} in {time.perf_counter() - t0:.1f}s")
    snps = L.make_snp_array(p.chr, p.cm, p.bp)
    base = dict(mode=L.WindowMode.Cm(1.0), chunk_size=200, snp_level_masking=True)
    cfg8, cfg32 = L.L2Config(**base), L.L2Config(**base, use_f32=True)
    cfg_i8 = L.L2Config(**base, exact_kernel="i8")
    panel = L.Panel(p.bed_bytes(), p.n, p.m)
    w8, pf8, (l2_8, _) = timed_steps(L, panel, snps, cfg8, args.steps, args.warmup, "exact")
    if args.i8_steps > 0:
        wi8, pfi8, (l2_i8, _) = timed_steps(L, panel, snps, cfg_i8, args.i8_steps, 1, "exact-int8")
    if args.fp32_steps > 0:
        w32, pf32, (l2_32, _) = timed_steps(L, panel, snps, cfg32, args.fp32_steps, 1, "fp32")
    panel.close()
    P = pf8[-1].pairs_included
    t_step = float(np.mean(w8))
    k8 = float(np.mean([q.bb_dot_secs for q in pf8]))
    algo = 2.0 * p.n * P
    fp32 = None
    if args.fp32_steps > 0:
        k32 = float(np.mean([q.bb_dot_secs for q in pf32]))
        fp32 = {"value": P / float(np.mean(w32)), "unit": "SNP-pairs/s",
                "ms_per_step": float(np.mean(w32)) * 1e3, "band_kernel_ms": k32 * 1e3,
                "kernel": "k_band_f32_2w", "tflops_algorithmic": algo / k32 / 1e12,
                "frac_of_fp32_peak": algo / k32 / 1e12 / FP32_PEAK_TFLOPS,
                "profile_trace_avg_ms": profiled("trace_avg_ms", "k_band_f32_2w"),
                "max_rel_vs_exact": float(np.max(np.abs(l2_32 - l2_8) /
                                                np.maximum(1.0, np.abs(l2_8))))}
    exact_i8 = None
    if args.i8_steps > 0:
        ki8 = float(np.mean([q.bb_dot_secs for q in pfi8]))
        exact_i8 = {"value": P / float(np.mean(wi8)), "unit": "SNP-pairs/s",
                    "ms_per_step": float(np.mean(wi8)) * 1e3, "band_kernel_ms": ki8 * 1e3,
                    "kernel": "k_band_i8", "achieved": algo / ki8 / 1e12,
                    "peak": INT8_PEAK_TOPS, "unit_peak": "TOP/s",
                    "frac": algo / ki8 / 1e12 / INT8_PEAK_TOPS,
                    "traffic": profiled_traffic("k_band_i8"),
                    "profile_trace_avg_ms": profiled("trace_avg_ms", "k_band_i8"),
                    "max_rel_vs_e2m1_form": float(np.max(np.abs(l2_i8 - l2_8) /
                                                         np.maximum(1.0, np.abs(l2_8))))}
    fp4 = p.n <= 1_800_000  # the e2m1 kernel's range (f32 sums stay exact)
    kname = "k_band_fp4" if fp4 else "k_band_i8"
    peak = FP4_PEAK_TFLOPS if fp4 else INT8_PEAK_TOPS
    tile_s = float(np.mean([q.norm_secs for q in pf8]))
    # tiler bytes: the call's .bed rows read, the tiled slice written
    n_pad, m_pad = -(-p.n // 128) * 128, -(-p.m // 64) * 64
    tile_bytes = p.m * ((p.n + 3) // 4) + m_pad * n_pad // 4
    res = {
        "value": P / t_step, "ms_per_step": t_step * 1e3, "pairs": P,
        "pairs_evaluated": pf8[-1].pairs_evaluated, "tiles": pf8[-1].tiles,
        "fp32": fp32, "exact_int8": exact_i8,
        "tile_ms": float(np.mean([q.norm_secs for q in pf8])) * 1e3,
        "reduce_ms": float(np.mean([q.ring_store_secs for q in pf8])) * 1e3,
        "host_prep_ms": float(np.mean([q.bed_read_secs for q in pf8])) * 1e3,
        "library_ms": float(np.mean([q.total_secs for q in pf8])) * 1e3,
        "roofline": {"bound": "mfma", "kernel": kname, "achieved": algo / k8 / 1e12,
                     "peak": peak, "unit": "TFLOP/s",
                     "frac": algo / k8 / 1e12 / peak,
                     "traffic": profiled_traffic(kname), "band_kernel_ms": k8 * 1e3,
                     "profile_trace_avg_ms": profiled("trace_avg_ms", kname),
                     "profile": PANEL_PROFILE,
                     "ops_per_launch": algo,
                     "note": "algorithmic 2*N*P per launch (one N-long dot product per "
                             "included pair); peak = dense " +
                             ("e2m1 (FP4)" if fp4 else "int8") + " MFMA"},
        "ld_scored_snps_per_s": p.m / t_step,
        "decode_roofline": {"bound": "hbm", "kernel": "k_tile_stats", "bytes": tile_bytes,
                            "achieved": tile_bytes / tile_s / 1e9, "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": tile_bytes / tile_s / 1e9 / HBM_PEAK_GBS,
                            "note": "M*ceil(N/4) .bed bytes read + M_pad*N_pad/4 tile bytes "
                                    "written, over the tiling phase (tiler + per-SNP stats)"},
    }
    if args.peaks:
        pk = args.peaks
        mp = pk["fp4"]["tflops"] if fp4 else pk["i8"]["tflops"]
        res["roofline"]["peak_measured"] = mp
        res["roofline"]["frac_of_measured"] = algo / k8 / 1e12 / mp
        res["roofline"]["clock_ghz_measured"] = pk["fp4" if fp4 else "i8"]["clock_ghz"]
        res["decode_roofline"]["peak_measured"] = pk["copy_gbs"]
        res["decode_roofline"]["frac_of_measured"] = tile_bytes / tile_s / 1e9 / pk["copy_gbs"]
        if fp32:
            fp32["peak_measured"] = pk["f32"]["tflops"]
            fp32["frac_of_measured"] = fp32["tflops_algorithmic"] / pk["f32"]["tflops"]
        if exact_i8:
            exact_i8["peak_measured"] = pk["i8"]["tflops"]
            exact_i8["frac_of_measured"] = exact_i8["achieved"] / pk["i8"]["tflops"]
    if not args.skip_cpu:
        from oracle import oracle as O

        # four slices of cpu_sample_snps / 4 SNPs spread over the panel (its
        # start, quarter, middle and three quarters), each timed and checked
        ms = min(args.cpu_sample_snps, p.m)
        nsl = 4
        per = max(ms // nsl, 1)
        starts = [int(q * (p.m - per) / (nsl - 1)) for q in range(nsl)]
        bed = p.bed_bytes()

        def cpu_runs(sl, use_f32):
            idx = np.arange(sl.start, sl.stop, dtype=np.uint64)
            kw = dict(mode=O.WINDOW_CM, window=1.0, chunk=200, masking=True, use_f32=use_f32,
                      threads=args.cpu_threads)
            O.l2(bed, p.n, p.chr[sl], p.cm[sl], p.bp[sl], idx, **kw)  # warmup
            walls, out = [], None
            for _ in range(3):
                t = time.perf_counter()
                out, _ = O.l2(bed, p.n, p.chr[sl], p.cm[sl], p.bp[sl], idx, **kw)
                walls.append(time.perf_counter() - t)
            return float(np.median(walls)), out

        slices, acc = [], {"exact": [], "exact_rms": [], "f32_abs": [], "f32_rel": [],
                           "f32_vs32": []}
        panel = L.Panel(bed, p.n, p.m)
        for st in starts:
            sl = slice(st, st + per)
            w64, exp = cpu_runs(sl, False)
            w32, exp32 = cpu_runs(sl, True)
            got, _, perf_s = L.compute_ldscore(panel, snps[sl], cfg8)
            got32, _, _ = L.compute_ldscore(panel, snps[sl], cfg32)
            slices.append({"first_snp": st, "snps": per, "pairs": perf_s.pairs_included,
                           "f64_wall_s": w64, "f32_wall_s": w32,
                           "f64_pairs_per_s": perf_s.pairs_included / w64,
                           "f32_pairs_per_s": perf_s.pairs_included / w32})
            acc["exact"].append(float(np.max(np.abs(got - exp) / np.abs(exp))))
            acc["exact_rms"].append(float(np.sqrt(np.mean(((got - exp) / exp) ** 2))))
            acc["f32_abs"].append(float(np.max(np.abs(got32 - exp))))
            acc["f32_rel"].append(float(np.max(np.abs(got32 - exp) / np.abs(exp))))
            acc["f32_vs32"].append(float(np.max(np.abs(got32 - exp32) / np.abs(exp32))))
        panel.close()
        res["sample_accuracy"] = {
            "exact_vs_fp64_max_rel": max(acc["exact"]),
            "exact_vs_fp64_rms_rel": max(acc["exact_rms"]),
            "fp32_vs_fp64_max_abs": max(acc["f32_abs"]),
            "fp32_vs_fp64_max_rel": max(acc["f32_rel"]),
            "fp32_vs_f32_restatement_max_rel": max(acc["f32_vs32"]),
        }
        model = cpu_model()
        P_s = sum(x["pairs"] for x in slices)
        cpu_wall = sum(x["f64_wall_s"] for x in slices)
        cpu_wall32 = sum(x["f32_wall_s"] for x in slices)
        rates = [x["f64_pairs_per_s"] for x in slices]
        res["cpu_baseline"] = {
            "value": P_s / cpu_wall, "unit": "SNP-pairs/s",
            "cores": args.cpu_threads, "kind": "port",
            "sample": f"{nsl} slices of {per} SNPs at SNPs {starts} of the same panel "
                      f"(N={p.n}, cM 1, masked), {P_s} pairs in all; C restatement of the "
                      f"reference's f64 path, per slice the median of 3 runs after 1 warmup, "
                      f"{cpu_wall:.2f}s summed; slice rates {min(rates):.3e}-{max(rates):.3e}",
            "fast_f32_value": P_s / cpu_wall32,
            "fast_f32_sample": f"same slices, f32 restatement (--fast-f32), {cpu_wall32:.2f}s",
            "slices": slices,
            "host": {"cpu_model": model, "nproc": os.cpu_count()},
        }
        log(f"cpu baseline {res['cpu_baseline']['value']:.3e} pairs/s (f32 "
            f"{res['cpu_baseline']['fast_f32_value']:.3e}); {res['sample_accuracy']}")
        bed_s = np.concatenate([p.buf[:3], p.body[:ms].reshape(-1)])
        res["pair_accuracy"] = pair_accuracy(args, L, O, p, snps, ms, bed_s, cfg8, cfg32)
    return res


def pair_accuracy(args, L, O, p, snps, ms, bed_s, cfg8, cfg32, n_probe=24, d=1000):
    """Errors of individual pair values against the FP64 restatement.

    With annotation column k the indicator of probe SNP j_k, l2[i, k] is the
    unbiased r^2 of the pair (i, j_k) for every i whose pair with j_k is
    included (and 1 at i = j_k), so the LD-score API itself yields pair
    values: r2u of the exact and FP32 forms against the f64 path, and the
    sketch-corrected r2u (d buckets) against the restatement of the
    reference's sketch path.  Also the sketch-corrected l2 against the exact
    f64 l2 (statistical, not parity).  On the first 2,000 SNPs of the slice."""
    ms = min(ms, 2000)
    bed_s = bed_s[:3 + ms * ((p.n + 3) // 4)]
    probes = np.linspace(50, ms - 50, n_probe).astype(np.int64)
    A = np.zeros((ms, n_probe))
    A[probes, np.arange(n_probe)] = 1.0
    kw = dict(mode=O.WINDOW_CM, window=1.0, chunk=200, masking=True, threads=args.cpu_threads)
    ref64, _ = O.l2(bed_s, p.n, p.chr[:ms], p.cm[:ms], p.bp[:ms], annot=A, **kw)
    ref32, _ = O.l2(bed_s, p.n, p.chr[:ms], p.cm[:ms], p.bp[:ms], annot=A, use_f32=True, **kw)
    refsk, _ = O.l2(bed_s, p.n, p.chr[:ms], p.cm[:ms], p.bp[:ms], annot=A, sketch=d, **kw)
    lsk_ref, _ = O.l2(bed_s, p.n, p.chr[:ms], p.cm[:ms], p.bp[:ms], sketch=d, **kw)
    l64, _ = O.l2(bed_s, p.n, p.chr[:ms], p.cm[:ms], p.bp[:ms], **kw)
    cfgsk = L.L2Config(mode=L.WindowMode.Cm(1.0), chunk_size=200, snp_level_masking=True,
                       sketch=d)
    with L.Panel(bed_s, p.n, ms) as ps:
        g8, _, _ = L.compute_ldscore(ps, snps[:ms], cfg8, annot=A)
        g32, _, _ = L.compute_ldscore(ps, snps[:ms], cfg32, annot=A)
        gsk, _, _ = L.compute_ldscore(ps, snps[:ms], cfgsk, annot=A)
        lsk, _, _ = L.compute_ldscore(ps, snps[:ms], cfgsk)
    incl = ref64 != 0.0  # included pairs (an exact zero r2u does not occur)
    incl[probes, np.arange(n_probe)] = False  # drop the self terms

    def stats(got, ref):
        e = np.abs(got[incl] - ref[incl])
        big = np.abs(ref[incl]) >= 0.01
        return {"max_abs": float(e.max()), "rms_abs": float(np.sqrt(np.mean(e ** 2))),
                "max_rel_where_abs_ge_0.01": float(np.max(e[big] / np.abs(ref[incl][big])))}

    out = {"pairs": int(incl.sum()), "probes": n_probe,
           "exact_r2u_vs_fp64": stats(g8, ref64),
           "fp32_r2u_vs_fp64": stats(g32, ref64),
           "fp32_r2u_vs_f32_restatement": stats(g32, ref32),
           f"sketch{d}_r2u_vs_sketch_restatement": stats(gsk, refsk),
           f"sketch{d}_l2_vs_sketch_restatement_max_rel":
               float(np.max(np.abs(lsk - lsk_ref) / np.maximum(1.0, np.abs(lsk_ref)))),
           f"sketch{d}_l2_vs_exact_fp64": {
               "max_abs": float(np.max(np.abs(lsk - l64))),
               "rms_abs": float(np.sqrt(np.mean((lsk - l64) ** 2))),
               "pearson_r": float(np.corrcoef(lsk[:, 0], l64[:, 0])[0, 1])}}
    log(f"pair accuracy: {out}")
    return out


# ---------------------------------------------------------------------------
# Serving: mixed request stream on resident panels
# ---------------------------------------------------------------------------
def leg_serving(args, L, synth):
    from ldsc_amd.serve import Server

    t0 = time.perf_counter()
    specs = [(50_000, args.serve_m_big, 20261207), (10_000, args.serve_m_small, 20261208)]
    panels, snp_sets, annots = [], [], []
    for n, m, seed in specs:
        p = synth.make_panel(n, m, seed, miss_rate=0.01)
        panels.append(L.Panel(p.bed_bytes(), n, m))
        snp_sets.append(L.make_snp_array(p.chr, p.cm, p.bp))
        annots.append(synth.make_annot(m, seed, n_bin=8, n_cont=2))
        del p
        log(f"serving: panel N={n} M={m} resident ({time.perf_counter() - t0:.1f}s)")
    rng = np.random.default_rng(20261209)
    modes = [L.WindowMode.Snp(500), L.WindowMode.Kb(1000.0), L.WindowMode.Cm(1.0)]
    reqs = []
    for _ in range(args.serve_requests):
        which = int(rng.integers(0, 2))
        m = snp_sets[which].size
        size = int(np.exp(rng.uniform(np.log(1000), np.log(50_000))))
        start = int(rng.integers(0, m - size))
        K = int(rng.integers(0, 11))
        cfg = L.L2Config(mode=modes[int(rng.integers(0, 3))], chunk_size=200,
                         snp_level_masking=True)
        snps = snp_sets[which][start:start + size]
        annot = None if K == 0 else np.ascontiguousarray(annots[which][start:start + size, :K])
        pairs = int(np.sum(np.arange(size, dtype=np.int64) -
                           L.pair_bounds(snps, cfg).astype(np.int64)))
        reqs.append((which, snps, cfg, annot, pairs))
    gaps = rng.exponential(1.0 / args.qps, size=len(reqs))
    lats, tickets = [], []
    def warm_workers(srv, inflight):
        """Grow every worker's buffers to the largest request shape of the
        stream (60,000 SNPs of the larger panel, 10 annotation columns, each
        window mode): a worker whose buffers grow mid-stream frees its old
        ones, and hipFree synchronizes the whole device, stalling every
        request in flight."""
        for md in modes:
            cfg = L.L2Config(mode=md, chunk_size=200, snp_level_masking=True)
            warm = [srv.submit(panels[0], snp_sets[0][:60_000], cfg,
                               np.ascontiguousarray(annots[0][:60_000, :10]))
                    for _ in range(inflight)]
            for t in warm:
                srv.wait(t)

    with Server(hbm_budget_bytes=96 << 30, max_inflight=args.serve_inflight,
                queue_capacity=4096, batch_window_us=args.serve_window_us) as srv:
        warm_workers(srv, args.serve_inflight)
        log("serving: workers warm; starting the Poisson stream")
        t_start = time.perf_counter()
        t_next = t_start
        for (which, snps, cfg, annot, _), gap in zip(reqs, gaps):
            t_next += gap
            delay = t_next - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            tickets.append(srv.submit(panels[which], snps, cfg, annot))
        for t in tickets:
            lats.append(srv.wait(t))
        t_end = time.perf_counter()
        poisson_stats = srv.stats()
        # the same stream at higher offered rates (same requests, rescaled gaps)
        sweep = []
        for q in [float(x) for x in args.qps_sweep.split(",") if x]:
            sl, st = [], []
            t0 = time.perf_counter()
            t_next = t0
            for (which, snps, cfg, annot, _), gap in zip(reqs, gaps * (args.qps / q)):
                t_next += gap
                delay = t_next - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                st.append(srv.submit(panels[which], snps, cfg, annot))
            sl = np.array([srv.wait(t) for t in st])
            t1 = time.perf_counter()
            sweep.append({"offered_qps": q, "achieved_qps": len(st) / (t1 - t0),
                          "p50_ms": float(np.percentile(sl, 50)) * 1e3,
                          "p99_ms": float(np.percentile(sl, 99)) * 1e3,
                          "snp_pairs_per_s": float(sum(r[4] for r in reqs)) / (t1 - t0)})
            log(f"serving at {q} QPS: {sweep[-1]}")
    # capacity: the same requests submitted at once (a closed burst), on more
    # workers, with band batching
    def run_burst(window_us):
        with Server(hbm_budget_bytes=96 << 30, max_inflight=args.serve_burst_inflight,
                    queue_capacity=4096, batch_window_us=window_us) as srv:
            warm_workers(srv, args.serve_burst_inflight)
            s0 = srv.stats()
            tb = time.perf_counter()
            bt = [srv.submit(panels[w], snps, cfg, annot) for (w, snps, cfg, annot, _) in reqs]
            blats = np.array([srv.wait(t) for t in bt])
            tb_end = time.perf_counter()
            s1 = srv.stats()
        out = {"requests": len(reqs), "wall_s": tb_end - tb, "qps": len(reqs) / (tb_end - tb),
               "snp_pairs_per_s": float(sum(r[4] for r in reqs)) / (tb_end - tb),
               "p50_ms": float(np.percentile(blats, 50)) * 1e3,
               "p99_ms": float(np.percentile(blats, 99)) * 1e3,
               "max_inflight": args.serve_burst_inflight, "batch_window_us": window_us,
               "band_launches": s1["band_launches"] - s0["band_launches"],
               "batched_requests": s1["batched_requests"] - s0["batched_requests"],
               "max_batch": s1["max_batch"]}
        log(f"serving burst (window {window_us} us): {out}")
        return out

    burst = burst_batched = None
    if args.serve_burst_inflight > 0:
        run_burst(args.serve_window_us)  # warm-up: grows the pooled workspaces
        burst = run_burst(args.serve_window_us)
        burst_batched = run_burst(args.serve_batched_window_us)
    for pn in panels:
        pn.close()
    lats = np.array(lats)
    total_pairs = float(sum(r[4] for r in reqs))
    res = {
        "requests": len(reqs), "offered_qps": args.qps,
        "achieved_qps": len(reqs) / (t_end - t_start),
        "p50_ms": float(np.percentile(lats, 50)) * 1e3,
        "p99_ms": float(np.percentile(lats, 99)) * 1e3,
        "mean_ms": float(lats.mean()) * 1e3,
        "snp_pairs_per_s": total_pairs / (t_end - t_start),
        "max_inflight": args.serve_inflight,
        "batch_window_us": args.serve_window_us,
        "band_launches": poisson_stats["band_launches"],
        "batched_requests": poisson_stats["batched_requests"],
        "panels": "N=50000 x M=%d and N=10000 x M=%d resident" % (args.serve_m_big,
                                                                  args.serve_m_small),
        "qps_sweep": sweep,
        "burst": burst,
        "burst_shared_band_launches": burst_batched,
    }
    log(f"serving: {res}")
    return res


# ---------------------------------------------------------------------------
# CountSketch
# ---------------------------------------------------------------------------
SKETCH_PROFILE = "profiles/sketch_100k_x_200k.md"
PART_PROFILE = "profiles/partitioned_20k_x_500k.md"
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense bf16 MFMA, TFLOP/s
# the d-length contraction of the sketch path
SKETCH_CONTRACTION = {"kernel": "k_band_sk8_2w", "peak": INT8_PEAK_TOPS,
                      "note": "algorithmic 2*d*P ops per launch (one d-long dot product per "
                              "included pair); the kernel runs four exact int8 products (W.W, "
                              "W.E, E.W, E.E) per pair, so 25 % of the dense int8 peak is its "
                              "ceiling in these units"}


def leg_sketch(args, L, synth):
    """C3: CountSketch LD scores at N=100,000 x M=200,000 for d = 200, 1000,
    5000.  The preparation (decode, bucket-order gather and the exact integer
    bucket sums) and the d-length contraction (MFMA) are timed with HIP
    events inside the call (L2Perf); algorithmic bytes of the preparation:
    M*ceil(N/4) .bed bytes read + 8N bucket/sign bytes + 2*M*d bytes of the
    int8 bucket values (W and E) written; of the contraction, 2*d*P ops."""
    n, m = 100_000, args.sketch_m
    p = synth.make_panel(n, m, 20261206, miss_rate=0.01)
    snps = L.make_snp_array(p.chr, p.cm, p.bp)
    out = {"workload": f"CountSketch N={n} x M={m}, 1% missing, cM 1, chunk 200, masked",
           "d": {}}
    with L.Panel(p.bed_bytes(), n, m) as panel:
        ex_cfg = L.L2Config(mode=L.WindowMode.Cm(1.0), snp_level_masking=True)
        w, pf, (exact, _) = timed_steps(L, panel, snps, ex_cfg, 1, 1, "sketch exact")
        out["exact_ms_per_step"] = w[0] * 1e3
        P = pf[0].pairs_included
        for d in (200, 1000, 5000):
            cfg = L.L2Config(mode=L.WindowMode.Cm(1.0), snp_level_masking=True, sketch=d)
            w, pf, (sk, _) = timed_steps(L, panel, snps, cfg, args.steps, 1, f"sketch d={d}")
            t = float(np.mean(w))
            prep_s = float(np.mean([q.norm_secs + (q.sketch_secs or 0.0) for q in pf]))
            band_s = float(np.mean([q.bb_dot_secs for q in pf]))
            bytes_sk = m * ((n + 3) // 4) + 8 * n + 2 * m * d
            out["d"][str(d)] = {
                "snp_pairs_per_s": P / t, "ms_per_step": t * 1e3,
                "prep_ms": prep_s * 1e3, "contraction_ms": band_s * 1e3,
                "prep_roofline": {"bound": "hbm", "achieved": bytes_sk / prep_s / 1e9,
                                  "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": bytes_sk / prep_s / 1e9 / HBM_PEAK_GBS,
                                  "bytes": bytes_sk,
                                  "traffic": profiled_traffic(f"prep_d{d}", SKETCH_PROFILE)},
                "contraction_roofline": {
                    "bound": "mfma", "kernel": SKETCH_CONTRACTION["kernel"],
                    "achieved": 2.0 * d * P / band_s / 1e12,
                    "peak": SKETCH_CONTRACTION["peak"], "unit": "TFLOP/s",
                    "frac": 2.0 * d * P / band_s / 1e12 / SKETCH_CONTRACTION["peak"],
                    "profile_trace_avg_ms": profiled("trace_avg_ms",
                                                     f"{SKETCH_CONTRACTION['kernel']}_d{d}",
                                                     SKETCH_PROFILE),
                    "note": SKETCH_CONTRACTION["note"]},
                "speedup_vs_exact": out["exact_ms_per_step"] / (t * 1e3),
                "pearson_r_vs_exact": float(np.corrcoef(sk[:, 0], exact[:, 0])[0, 1]),
                "max_abs_diff_vs_exact": float(np.max(np.abs(sk - exact))),
            }
            log(f"sketch d={d}: {out['d'][str(d)]}")
        out["pairs"] = P
        if not args.skip_cpu:
            from oracle import oracle as O

            ms, d = min(args.cpu_sample_snps, m), 1000
            sl = slice(m // 2, m // 2 + ms)
            idx = np.arange(sl.start, sl.stop, dtype=np.uint64)
            cfg = L.L2Config(mode=L.WindowMode.Cm(1.0), snp_level_masking=True, sketch=d)
            got, _, perf_s = L.compute_ldscore(panel, snps[sl], cfg)
            wall, exp = time_cpu(lambda: O.l2(p.bed_bytes(), n, p.chr[sl], p.cm[sl], p.bp[sl],
                                              idx, mode=O.WINDOW_CM, window=1.0, chunk=200,
                                              masking=True, use_f32=True, sketch=d,
                                              threads=args.cpu_threads)[0])
            Ps = perf_s.pairs_included
            out["cpu_baseline"] = {
                "value": Ps / wall, "unit": "SNP-pairs/s", "cores": args.cpu_threads,
                "kind": "port",
                "sample": f"{ms} SNPs from the middle of the same panel (N={n}, cM 1, masked), "
                          f"{Ps} pairs, d={d}: C restatement of the reference's f32 sketch path "
                          f"(sequential f32 scatter, f32 GEMM), median of 3 runs after 1 "
                          f"warmup: {wall:.2f}s",
                "host": {"cpu_model": cpu_model(), "nproc": os.cpu_count()},
                "device_vs_restatement_max_rel": float(np.max(np.abs(got - exp) /
                                                              np.maximum(1.0, np.abs(exp)))),
            }
            log(f"sketch cpu baseline: {out['cpu_baseline']}")
    return out


# ---------------------------------------------------------------------------
# Partitioned LD scores
# ---------------------------------------------------------------------------
def leg_partitioned(args, L, synth):
    """C4: N=20,000 x M=500,000, K=50 annotations (40 binary in runs, 10
    LogNormal), --ld-wind-kb 1000, chunk 200.  The 200 MB annotation goes
    up and the 200 MB result comes back within each step (PCIe-inclusive);
    the band's roofline counts 2*N*P per launch over its HIP-event time."""
    n, m = 20_000, args.partitioned_m
    p = synth.make_panel(n, m, 20261207 + 100, miss_rate=0.01)
    snps = L.make_snp_array(p.chr, p.cm, p.bp)
    annot = synth.make_annot(m, 20261210, n_bin=40, n_cont=10)
    out = {"workload": f"partitioned N={n} x M={m}, K=50, kb 1000, chunk 200"}
    res = {}
    with L.Panel(p.bed_bytes(), n, m) as panel:
        for tag, f32, kern, steps in (("exact", False, "auto", args.steps),
                                      ("exact_int8", False, "i8", args.i8_steps),
                                      ("fp32", True, "auto", args.fp32_steps)):
            if steps <= 0:
                continue
            cfg = L.L2Config(mode=L.WindowMode.Kb(1000.0), chunk_size=200, use_f32=f32,
                             exact_kernel=kern)
            # the 200 MB result lands in one reused buffer, as in a loop over
            # annotation sets; a fresh array per call adds its page faults
            dest = np.empty((annot.shape[1], m)).T
            w, pf, (l2, _) = timed_steps(L, panel, snps, cfg, steps, 1, f"partitioned {tag}",
                                         annot=annot, dest=dest)
            P = pf[0].pairs_included
            t = float(np.mean(w))
            band = float(np.mean([q.bb_dot_secs for q in pf]))
            peak = FP32_PEAK_TFLOPS if f32 else INT8_PEAK_TOPS if kern == "i8" else FP4_PEAK_TFLOPS
            kname = "k_band_f32_2w" if f32 else "k_band_i8" if kern == "i8" else "k_band_fp4"
            out[tag] = {"snp_pairs_per_s": P / t, "ms_per_step": t * 1e3,
                        "band_ms": band * 1e3, "pairs": P,
                        "roofline": {"bound": "mfma", "kernel": kname,
                                     "achieved": 2.0 * n * P / band / 1e12,
                                     "peak": peak, "unit": "TFLOP/s",
                                     "frac": 2.0 * n * P / band / 1e12 / peak,
                                     "traffic_per_call": profiled_traffic(
                                         f"{kname}_per_call", PART_PROFILE),
                                     "profile_trace_ms_per_call": profiled(
                                         "trace_ms_per_call", kname, PART_PROFILE),
                                     "note": "2*N*P per call over the band's HIP-event time "
                                             "(its row-segment launches together); the FP64 "
                                             "annotation contraction, 4*P*K flops, runs in the "
                                             "same kernels"}}
            res[tag] = l2.copy()
        if "fp32" in res:
            out["fp32_vs_exact_max_rel"] = float(np.max(np.abs(res["fp32"] - res["exact"]) /
                                                        np.maximum(1.0, np.abs(res["exact"]))))
        if "exact_int8" in res:
            out["exact_int8_max_rel"] = float(np.max(np.abs(res["exact_int8"] - res["exact"]) /
                                                     np.maximum(1.0, np.abs(res["exact"]))))
        if not args.skip_cpu:
            from oracle import oracle as O

            ms = min(args.cpu_sample_snps, m)
            sl = slice(m // 2, m // 2 + ms)
            idx = np.arange(sl.start, sl.stop, dtype=np.uint64)
            cfg = L.L2Config(mode=L.WindowMode.Kb(1000.0), chunk_size=200)
            got, _, perf_s = L.compute_ldscore(panel, snps[sl], cfg, annot=annot[sl])
            wall, exp = time_cpu(lambda: O.l2(p.bed_bytes(), n, p.chr[sl], p.cm[sl], p.bp[sl],
                                              idx, mode=O.WINDOW_KB, window=1000.0, chunk=200,
                                              annot=annot[sl], threads=args.cpu_threads)[0])
            Ps = perf_s.pairs_included
            out["cpu_baseline"] = {
                "value": Ps / wall, "unit": "SNP-pairs/s", "cores": args.cpu_threads,
                "kind": "port",
                "sample": f"{ms} SNPs from the middle of the same panel (N={n}, kb 1000, K=50), "
                          f"{Ps} pairs: C restatement of the reference's f64 path, median of 3 "
                          f"runs after 1 warmup: {wall:.2f}s",
                "host": {"cpu_model": cpu_model(), "nproc": os.cpu_count()},
                "device_vs_restatement_max_rel": float(np.max(np.abs(got - exp) /
                                                              np.maximum(1.0, np.abs(exp)))),
            }
            log(f"partitioned cpu baseline: {out['cpu_baseline']}")
    return out


# ---------------------------------------------------------------------------
# The shapes of the reference's published CPU timings
# ---------------------------------------------------------------------------
def leg_reference_shapes(args, L, synth):
    """Synthetic panels of the shapes the reference publishes l2 wall times
    for: 1000G (N=2,490, M=1.66 M, --ld-wind-kb 1000; Rust exact f64 41.1 s,
    --fast-f32 ~32 s on 16 vCPU of an EPYC 7R13) and the N=50,000 biobank
    panel (M=1.66 M, kb 1000, chunk 200; exact f64 727 s, f32 424 s).  The
    data differ (founder-haplotype generator, one chromosome), so these are
    same-shape wall times, not a parity comparison."""
    out = {}
    for name, n, m, seed in (("1000G_shape", 2490, 1_660_000, 20261211),
                             ("biobank_shape", 50_000, 1_660_000, 20261212)):
        t0 = time.perf_counter()
        p = synth.make_panel(n, m, seed, miss_rate=0.01)
        snps = L.make_snp_array(p.chr, p.cm, p.bp)
        log(f"{name}: panel N={n} M={m} generated ({time.perf_counter() - t0:.1f}s)")
        res = {"n_indiv": n, "m_snps": m}
        with L.Panel(p.bed_bytes(), n, m) as panel:
            del p
            for tag, f32 in (("exact", False), ("fp32", True)):
                cfg = L.L2Config(mode=L.WindowMode.Kb(1000.0), chunk_size=200, use_f32=f32)
                w, pf, _ = timed_steps(L, panel, snps, cfg, args.steps, 1, f"{name} {tag}")
                res[tag] = {"wall_s": float(np.mean(w)), "pairs": pf[0].pairs_included,
                            "snp_pairs_per_s": pf[0].pairs_included / float(np.mean(w)),
                            "band_ms": float(np.mean([q.bb_dot_secs for q in pf])) * 1e3,
                            "tile_ms": float(np.mean([q.norm_secs for q in pf])) * 1e3}
        out[name] = res
        log(f"{name}: {res}")
    out["reference_published_s"] = {"1000G_shape": {"exact_f64": 41.1, "fast_f32": 32.0},
                                     "biobank_shape": {"exact_f64": 727.0, "f32": 424.0}}
    return out


def leg_cli(args, L, synth):
    """The `ldsc l2` command end to end (ldsc_amd.cli): BIM/FAM/BED read, the
    panel upload, the LD scores, the .l2.ldscore.gz / .l2.M / .l2.M_5_50 files
    of the genome and of each chromosome, and the summary.  The panel has the
    1000G shape (N=2,490, M=1.66 M, here over 22 chromosomes of equal SNP
    count) with --ld-wind-kb 1000, the default chunk 200 and the default
    per-chromosome pass.  The reference publishes 41.1 s (exact f64) and
    about 32 s (--fast-f32) for its `l2` run on 1000G on 16 vCPU of an EPYC
    7R13; the data differ, so this is a same-shape wall time."""
    import contextlib
    import io
    import shutil
    import tempfile

    from ldsc_amd import cli

    n, m = 2490, 1_660_000
    tmp = tempfile.mkdtemp(prefix="ldsc_cli_")
    try:
        t0 = time.perf_counter()
        p = synth.make_panel(n, m, 20261213, miss_rate=0.01)
        p.chr[:] = 1 + np.arange(m) * 22 // m
        prefix = os.path.join(tmp, "g")
        synth.write_plink(prefix, p)
        del p
        log(f"cli: PLINK files N={n} M={m} written ({time.perf_counter() - t0:.1f}s)")
        res = {"n_indiv": n, "m_snps": m, "chromosomes": 22, "window": "--ld-wind-kb 1000"}
        for tag, extra in (("exact", []), ("fast_f32", ["--fast-f32"])):
            walls, stages = [], []
            for i in range(1 + args.steps):
                err = io.StringIO()
                argv = ["l2", "--bfile", prefix, "--out", os.path.join(tmp, "o"),
                        "--ld-wind-kb", "1000", "--verbose-timing", *extra]
                t1 = time.perf_counter()
                with contextlib.redirect_stdout(io.StringIO()), contextlib.redirect_stderr(err):
                    rc = cli.main(argv)
                w = time.perf_counter() - t1
                if rc != 0:
                    raise RuntimeError(f"ldsc l2 {tag} failed: {err.getvalue()[-2000:]}")
                st = {k: float(v) for k, v in
                      re.findall(r"\[perf\] (\w+)=([0-9.]+)s", err.getvalue())}
                log(f"cli {tag} run {i}: {w:.3f}s {st}")
                if i:
                    walls.append(w)
                    stages.append(st)
            keys = sorted(set().union(*stages))
            res[tag] = {"wall_s": float(np.mean(walls)), "runs": len(walls),
                        "stages_s": {k: float(np.mean([s_.get(k, 0.0) for s_ in stages]))
                                     for k in keys}}
        res["reference_published_s"] = {"exact_f64": 41.1, "fast_f32": 32.0}
        return res
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="panel",
                    choices=["panel", "sketch", "partitioned", "reference_shapes", "cli"])
    ap.add_argument("--n", type=int, default=50_000)
    ap.add_argument("--m", type=int, default=100_000)
    ap.add_argument("--fp32-steps", type=int, default=2)
    ap.add_argument("--i8-steps", type=int, default=3)
    ap.add_argument("--cpu-sample-snps", type=int, default=8000)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-serving", action="store_true")
    ap.add_argument("--skip-sketch", action="store_true")
    ap.add_argument("--skip-partitioned", action="store_true")
    ap.add_argument("--skip-peaks", action="store_true",
                    help="do not measure the sustained MFMA and copy peaks")
    ap.add_argument("--qps", type=float, default=200.0)
    ap.add_argument("--qps-sweep", default="500,1000,2000",
                    help="further offered rates for the same request stream (comma list)")
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--serve-inflight", type=int, default=4)
    ap.add_argument("--serve-window-us", type=int, default=0,
                    help="band batching window of the request server (0: one launch each)")
    ap.add_argument("--serve-batched-window-us", type=int, default=200,
                    help="batching window of the comparison burst with shared band launches")
    ap.add_argument("--serve-burst-inflight", type=int, default=8,
                    help="workers of the closed-burst capacity run (0: skip it)")
    ap.add_argument("--serve-m-big", type=int, default=1_000_000)
    ap.add_argument("--serve-m-small", type=int, default=500_000)
    ap.add_argument("--sketch-m", type=int, default=200_000)
    ap.add_argument("--partitioned-m", type=int, default=500_000)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB (torch) to "
                         "calibrate HBM byte counters, then exit")
    args = ap.parse_args()
    if args.calib_copy_mib:
        import torch

        n = args.calib_copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calib_copy_bytes_each_way": n * 4}), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the LD-score "
              "engine runs on one GPU and has no multi-GPU path.", file=sys.stderr)
        return 2

    import ldsc_amd as L
    from ldsc_amd import synth

    args.peaks = None if args.skip_peaks else measured_peaks()
    if args.workload == "sketch":
        print(json.dumps({"metric": "CountSketch LD scores", **leg_sketch(args, L, synth)}),
              flush=True)
        return 0
    if args.workload == "reference_shapes":
        print(json.dumps({"metric": "l2 wall time at the reference's published shapes",
                          **leg_reference_shapes(args, L, synth)}), flush=True)
        return 0
    if args.workload == "cli":
        print(json.dumps({"metric": "ldsc l2 command wall time, PLINK files to LD-score files",
                          **leg_cli(args, L, synth)}), flush=True)
        return 0
    if args.workload == "partitioned":
        print(json.dumps({"metric": "partitioned LD scores (K=50)", **leg_partitioned(args, L, synth)}),
              flush=True)
        return 0

    pan = leg_panel(args, L, synth)
    serving = None if args.skip_serving else leg_serving(args, L, synth)
    sketch = None if args.skip_sketch else leg_sketch(args, L, synth)
    part = None if args.skip_partitioned else leg_partitioned(args, L, synth)
    out = {
        "metric": METRIC,
        "value": pan["value"],
        "unit": "SNP-pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": pan["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "exact integer products on e2m1 MFMA (f32 acc) + f64 finish",
        "data": "synthetic (seeded founder-haplotype generator)",
        "config": {"workload": "exact LD scores, N=50000 x M=100000, 1% missing, "
                               "--ld-wind-cm 1, chunk 200, --snp-level-masking",
                   "n_indiv": args.n, "m_snps": args.m, "pairs": pan["pairs"],
                   "pairs_evaluated": pan["pairs_evaluated"], "tiles": pan["tiles"]},
        "fp32_plain": pan["fp32"],
        "exact_int8": pan["exact_int8"],
        "exact_over_fp32": pan["value"] / pan["fp32"]["value"] if pan["fp32"] else None,
        "ld_scored_snps_per_s": pan["ld_scored_snps_per_s"],
        "phases_ms": {"tile_stats": pan["tile_ms"], "band": pan["roofline"]["band_kernel_ms"],
                      "reduce": pan["reduce_ms"], "host_prep": pan["host_prep_ms"],
                      "library_call": pan["library_ms"]},
        "roofline": pan["roofline"],
        "decode_roofline": pan["decode_roofline"],
        "measured_peaks": args.peaks,
        "cpu_baseline": pan.get("cpu_baseline"),
        "sample_accuracy": pan.get("sample_accuracy"),
        "pair_accuracy": pan.get("pair_accuracy"),
        "serving": serving,
        "sketch": sketch,
        "partitioned": part,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
