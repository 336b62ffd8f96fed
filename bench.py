"""LD-score throughput on one MI355X.

Workload: exact LD scores at panel scale -- synthetic founder-haplotype PLINK
panel with N=50,000 individuals, M=100,000 SNPs on one chromosome, 1 %
missing, --ld-wind-cm 1, --chunk-size 200, --snp-level-masking.  A step is
one full ldsc_l2_compute over the resident panel.  The headline `value` is
SNP-pairs/s of the int8-MFMA + FP64-finish form; the plain FP32 form runs in
the same process and is reported beside it.

Prints one JSON line.  `--gpus` above 1 is refused: this build targets one
GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

INT8_PEAK_TOPS = 5000.0    # MI355X dense int8 MFMA, TOP/s
FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 matrix, TFLOP/s


def profiled_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 summary
    (FETCH_SIZE/WRITE_SIZE scaled by a copy-kernel calibration), or None."""
    path = os.path.join(ROOT, "profiles", "k_band_i8_c2.md")
    try:
        for line in open(path):
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
                return int(parts[2])
    except OSError:
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=50_000)
    ap.add_argument("--m", type=int, default=100_000)
    ap.add_argument("--fp32-steps", type=int, default=2)
    ap.add_argument("--cpu-sample-snps", type=int, default=6000)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--skip-cpu", action="store_true")
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
              "engine has no multi-GPU path in this build.", file=sys.stderr)
        return 2

    import ldsc_amd as L
    from ldsc_amd import synth

    seed = 20261205
    t0 = time.perf_counter()
    p = synth.make_panel(args.n, args.m, seed, miss_rate=0.01)
    log(f"generated panel N={p.n} M={p.m} in {time.perf_counter() - t0:.1f}s")
    snps = L.make_snp_array(p.chr, p.cm, p.bp)
    cfg8 = L.L2Config(mode=L.WindowMode.Cm(1.0), chunk_size=200, snp_level_masking=True)
    cfg32 = L.L2Config(mode=L.WindowMode.Cm(1.0), chunk_size=200, snp_level_masking=True,
                       use_f32=True)
    panel = L.Panel(p.bed_bytes(), p.n, p.m)
    log("panel resident on device")

    def run(cfg, steps, warmup, tag):
        for w in range(warmup):
            L.compute_ldscore(panel, snps, cfg)
            log(f"{tag} warmup {w + 1}/{warmup} done")
        walls, kern, outs = [], [], None
        for s in range(steps):
            t = time.perf_counter()
            l2, maf, perf = L.compute_ldscore(panel, snps, cfg)
            walls.append(time.perf_counter() - t)
            kern.append(perf.bb_dot_secs)
            outs = (l2, maf, perf)
            log(f"{tag} step {s + 1}/{steps}: {walls[-1] * 1e3:.2f} ms "
                f"(band kernel {perf.bb_dot_secs * 1e3:.2f} ms)")
        return walls, kern, outs

    w8, k8, (l2_8, maf8, perf8) = run(cfg8, args.steps, args.warmup, "int8")
    w32, k32, (l2_32, _, perf32) = run(cfg32, args.fp32_steps, 1, "fp32")

    P = perf8.pairs_included
    t_step = float(np.mean(w8))
    value = P / t_step
    fp32_value = P / float(np.mean(w32))
    algo_ops = 2.0 * p.n * P
    k8m, k32m = float(np.mean(k8)), float(np.mean(k32))
    achieved_tops = algo_ops / k8m / 1e12
    rel32 = float(np.max(np.abs(l2_32 - l2_8) / np.maximum(1.0, np.abs(l2_8))))

    # CPU baseline: the C restatement (f64, the reference default path) on the
    # first `cpu-sample-snps` SNPs of the same panel, all pairs of that slice.
    cpu = None
    sample_err = None
    if not args.skip_cpu:
        from oracle import oracle as O

        ms = min(args.cpu_sample_snps, p.m)
        sub = p.body[:ms]
        bed_s = np.concatenate([np.array([0x6C, 0x1B, 0x01], np.uint8), sub.reshape(-1)])
        t = time.perf_counter()
        exp, _ = O.l2(bed_s, p.n, p.chr[:ms], p.cm[:ms], p.bp[:ms], mode=O.WINDOW_CM,
                      window=1.0, chunk=200, masking=True, threads=args.cpu_threads)
        cpu_wall = time.perf_counter() - t
        with L.Panel(bed_s, p.n, ms) as ps:
            got, _, perf_s = L.compute_ldscore(ps, snps[:ms], cfg8)
            got32, _, _ = L.compute_ldscore(ps, snps[:ms], cfg32)
        sample_err = {
            "int8_fp64_max_rel": float(np.max(np.abs(got - exp) / np.abs(exp))),
            "fp32_max_abs": float(np.max(np.abs(got32 - exp))),
            "fp32_max_rel": float(np.max(np.abs(got32 - exp) / np.abs(exp))),
        }
        cpu = {
            "value": perf_s.pairs_included / cpu_wall,
            "unit": "SNP-pairs/s",
            "cores": args.cpu_threads,
            "kind": "port",
            "sample": f"first {ms} SNPs of the same panel (N={p.n}, cM 1, masked), "
                      f"{perf_s.pairs_included} pairs, FP64 restatement, {cpu_wall:.1f}s",
        }
        log(f"cpu baseline: {cpu['value']:.3e} pairs/s over {cpu_wall:.1f}s; sample errors "
            f"{sample_err}")

    out = {
        "metric": "SNP-pairs/s int8-MFMA vs FP32 plain (≤1e-6 rel); p50/p99 LD-request latency",
        "value": value,
        "unit": "SNP-pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8 (int32 acc) + f64 finish",
        "data": "synthetic (seeded founder-haplotype generator)",
        "config": {"workload": "exact LD scores, N=50000 x M=100000, 1% missing, "
                               "--ld-wind-cm 1, chunk 200, --snp-level-masking",
                   "n_indiv": p.n, "m_snps": p.m, "pairs": P,
                   "pairs_evaluated": perf8.pairs_evaluated, "tiles": perf8.tiles},
        "fp32_plain": {"value": fp32_value, "unit": "SNP-pairs/s",
                       "ms_per_step": float(np.mean(w32)) * 1e3,
                       "band_kernel_ms": k32m * 1e3,
                       "tflops_algorithmic": algo_ops / k32m / 1e12,
                       "frac_of_fp32_peak": algo_ops / k32m / 1e12 / FP32_PEAK_TFLOPS,
                       "max_rel_vs_int8": rel32},
        "int8_over_fp32": value / fp32_value,
        "ld_scored_snps_per_s": p.m / t_step,
        "roofline": {"bound": "mfma", "kernel": "k_band_i8", "achieved": achieved_tops,
                     "peak": INT8_PEAK_TOPS, "unit": "TFLOP/s", "frac": achieved_tops / INT8_PEAK_TOPS,
                     "traffic": profiled_traffic("k_band_i8"), "band_kernel_ms": k8m * 1e3,
                     "ops_per_launch": algo_ops},
        "cpu_baseline": cpu,
        "sample_accuracy": sample_err,
    }
    print(json.dumps(out), flush=True)
    panel.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
