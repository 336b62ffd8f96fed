"""`ldsc l2` / `ldsc --l2` command line over the MI355X engine.

Mirrors the reference's `l2` subcommand: flags of L2Args (src/cli.rs:169-355),
the Python-style `--l2` injection (src/main.rs:39-78) and the orchestration
and outputs of `run` (src/l2/mod.rs:58-795): BIM/FAM/annot/extract/keep/
print-snps inputs, the MAF prefilter, `{out}.l2.ldscore.gz` (values `{:.3}`),
`{out}.l2.M` / `.l2.M_5_50` (shortest round-trip decimals), per-chromosome
files and the "Summary of LD Scores" block.  All LD-score arithmetic runs in
libldsc_l2.so; the BIM reader and the gzip writer run in libldsc_io.so.

Usage:  python -m ldsc_amd.cli l2 --bfile PREFIX --out OUT [--ld-wind-cm 1] ...
        python -m ldsc_amd.cli --bfile PREFIX --l2 --out OUT ...
"""
from __future__ import annotations

import argparse
import bz2
import gzip
import math
import os
import sys
import time
from typing import List, Optional

import numpy as np

from . import _native as N
from .formats import BimTable, parse_annot_text, read_text_bytes, write_ldscore
from .l2 import L2Config, Panel, WindowMode, compute_ldscore, decode, make_snp_array


# ----------------------------------------------------------------------------
# number formatting
# ----------------------------------------------------------------------------
def rust_display(v: float) -> str:
    """Rust `{}` for f64: shortest round-trip digits, never an exponent, no
    trailing `.0` (src/l2/io.rs:161-168)."""
    if math.isnan(v):
        return "NaN"
    if math.isinf(v):
        return "inf" if v > 0 else "-inf"
    return np.format_float_positional(v, unique=True, trim="-")


def format_m_vals(vals) -> str:
    return "\t".join(rust_display(float(v)) for v in vals) + "\n"


# ----------------------------------------------------------------------------
# inputs
# ----------------------------------------------------------------------------
def open_text(path: str):
    if path.endswith(".gz"):
        return gzip.open(path, "rt")
    if path.endswith(".bz2"):
        return bz2.open(path, "rt")
    return open(path, "r")


def count_fam(path: str) -> int:
    with open(path) as f:
        return sum(1 for _ in f)


def parse_fam_ids(path: str):
    ids = []
    with open(path) as f:
        for i, line in enumerate(f):
            cols = line.split()
            if len(cols) < 2:
                raise ValueError(f"FAM line {i + 1}: expected ≥ 2 columns (FID IID ...), "
                                 f"got {len(cols)}")
            ids.append((cols[0], cols[1]))
    return ids


def load_snp_set(path: str) -> set:
    """load_snp_set (src/l2/io.rs:171-187)."""
    out = set()
    with open(path) as f:
        for line in f:
            if line.lstrip().startswith("#"):
                continue
            cols = line.split()
            if cols:
                out.add(cols[0])
    print(f"  Loaded {len(out)} SNP IDs from '{path}'")
    return out


def load_print_snps(path: str) -> set:
    """load_print_snps (src/l2/io.rs:190-209): exactly one column."""
    out = set()
    with open_text(path) as f:
        for line in f:
            if line.lstrip().startswith("#"):
                continue
            cols = line.split()
            if not cols:
                continue
            if len(cols) != 1:
                raise ValueError("--print-snps must refer to a file with a one column of SNP IDs.")
            out.add(cols[0])
    print(f"  Loaded {len(out)} SNP IDs from '{path}'")
    return out


def load_individual_indices(keep_path: str, fam_ids) -> np.ndarray:
    """load_individual_indices (src/l2/io.rs:53-91): FAM rows in FAM order."""
    keep = set()
    with open(keep_path) as f:
        for line in f:
            if line.lstrip().startswith("#"):
                continue
            cols = line.split()
            if len(cols) >= 2:
                keep.add((cols[0], cols[1]))
    idx = [i for i, fi in enumerate(fam_ids) if fi in keep]
    if not idx:
        raise ValueError(f"--keep '{keep_path}': none of {len(keep)} individuals matched the "
                         f"FAM ({len(fam_ids)} individuals)")
    print(f"  --keep: retaining {len(idx)}/{len(fam_ids)} individuals")
    return np.array(idx, np.int64)


def resolve_annot_path(prefix: str) -> str:
    """resolve_annot_path (src/parse.rs:166-187)."""
    if prefix.endswith((".annot", ".annot.gz", ".annot.bz2")):
        if os.path.exists(prefix):
            return prefix
        raise FileNotFoundError(f"Annotation file not found: '{prefix}'")
    for suf in (".annot.gz", ".annot.bz2", ".annot"):
        if os.path.exists(prefix + suf):
            return prefix + suf
    raise FileNotFoundError(f"Annotation file not found: '{prefix}.annot[.gz|.bz2]'")


def read_annot_path(path: str, thin: bool):
    """read_annot_path (src/parse.rs:191-229) through libldsc_io.so: header
    row; tab-separated if the header has a tab, else whitespace; the full
    format skips CHR SNP BP CM; NA / NaN / non-numbers -> 0."""
    return parse_annot_text(read_text_bytes(path), thin, path)


def read_annotations(prefix: str, thin: bool, chr_: np.ndarray) -> tuple:
    """The three --annot forms of run (src/l2/mod.rs:147-252); ``chr_`` is the
    chromosome of each SNP of the LD-score list."""
    explicit = prefix.endswith((".annot", ".annot.gz", ".annot.bz2"))
    try_single = explicit
    if not explicit:
        try:
            resolve_annot_path(prefix)
            try_single = True
        except FileNotFoundError:
            try_single = False
    if try_single:
        path = resolve_annot_path(prefix)
        mat, names = read_annot_path(path, thin)
        src = path
    else:
        chrs = [int(c) for c in np.unique(chr_)]
        mats, names = [], []
        for c in chrs:
            m, nm = read_annot_path(resolve_annot_path(f"{prefix}{c}"), thin)
            names = names or nm
            mats.append(m)
        mat = np.concatenate(mats, axis=0) if mats else np.zeros((0, 0))
        src = prefix + "*"
    if mat.shape[0] != len(chr_):
        raise ValueError(f"Annotation file has {mat.shape[0]} rows but BIM has {len(chr_)} "
                         "SNPs — they must match exactly")
    print(f"Read {len(names)} annotations for {mat.shape[0]} SNPs from '{src}'")
    return mat, names


# ----------------------------------------------------------------------------
# outputs (the .l2.ldscore.gz writer is formats.write_ldscore)
# ----------------------------------------------------------------------------
def _percentile(v: np.ndarray, q: float) -> float:
    """Linear-interpolation percentile of the ascending-sorted ``v``."""
    if v.size == 1:
        return float(v[0])
    pos = q * (v.size - 1)
    lo = int(math.floor(pos))
    hi = lo + 1
    frac = pos - lo
    return float(v[lo]) if hi >= v.size else float(v[lo] * (1 - frac) + v[hi] * frac)


def _pearson(x, y) -> float:
    mx, my = x.mean(), y.mean()
    dx, dy = x - mx, y - my
    vx, vy = float((dx * dx).sum()), float((dy * dy).sum())
    if vx == 0.0 or vy == 0.0:
        return float("nan")
    return float((dx * dy).sum()) / math.sqrt(vx * vy)


def print_ld_score_summary(out_path, l2_out: np.ndarray, col_names, maf_out: np.ndarray):
    """print_ld_score_summary (src/l2/mod.rs:799-865)."""
    n = l2_out.shape[0]
    if n == 0:
        return
    cols = [("MAF", maf_out)] + [(nm, l2_out[:, k]) for k, nm in enumerate(col_names)]
    print(f"\nSummary of LD Scores in {out_path}")
    print(" " * 8 + "".join(f"{nm:>8}" for nm, _ in cols))
    srt = [np.sort(v) for _, v in cols]
    for lab in ("mean", "std", "min", "25%", "50%", "75%", "max"):
        row = f"{lab:>8}"
        for ci, (_, v) in enumerate(cols):
            if lab == "mean":
                st = v.sum() / n
            elif lab == "std":
                mu = v.sum() / n
                st = math.sqrt(((v - mu) ** 2).sum() / max(n - 1, 1))
            elif lab == "min":
                st = v.min()
            elif lab == "max":
                st = v.max()
            else:
                st = _percentile(srt[ci], {"25%": 0.25, "50%": 0.5, "75%": 0.75}[lab])
            row += f"{st:>8.4f}"
        print(row)
    print()
    print("MAF/LD Score Correlation Matrix")
    print(" " * 8 + "".join(f"{nm:>8}" for nm, _ in cols))
    for i, (ni, vi) in enumerate(cols):
        row = f"{ni:>8}"
        for j, (_, vj) in enumerate(cols):
            row += f"{(1.0 if i == j else _pearson(vi, vj)):>8.4f}"
        print(row)


# ----------------------------------------------------------------------------
# run
# ----------------------------------------------------------------------------
def build_parser() -> argparse.ArgumentParser:
    ap = argparse.ArgumentParser(prog="ldsc l2", description="Compute LD scores (MI355X)")
    a = ap.add_argument
    a("--bfile", required=True)
    a("--out", required=True)
    a("--ld-wind-cm", type=float)
    a("--ld-wind-kb", type=float)
    a("--ld-wind-snps", type=int)
    a("--annot")
    a("--cts-bin")
    a("--cts-breaks")
    a("--cts-names")
    a("--thin-annot", action="store_true")
    a("--extract")
    a("--maf", type=float)
    a("--maf-pre", nargs="?", const="true", default="true")
    a("--mmap", action="store_true")
    a("--global-pass", action="store_true")
    a("--snp-level-masking", action="store_true")
    a("--print-snps")
    a("--keep")
    a("--per-allele", action="store_true")
    a("--pq-exp", type=float)
    a("--no-print-annot", action="store_true")
    a("--chunk-size", type=int, default=200)
    a("--python-compat", action="store_true")
    a("--yes-really", action="store_true")
    a("--fast-f32", action="store_true")
    a("--sketch", type=int)
    a("--sketch-maf-aware", action="store_true")
    a("--verbose-timing", action="store_true")
    # reference GPU knobs: the whole path runs on the device here; accepted
    a("--gpu", action="store_true")
    a("--gpu-tile-cols", type=int)
    a("--gpu-flex32", action="store_true")
    a("--gpu-f64", action="store_true")
    return ap


def run(args) -> int:
    t_start = time.perf_counter()
    if args.per_allele and args.pq_exp is not None:
        raise SystemExit("Cannot set both --per-allele and --pq-exp (--per-allele is "
                         "equivalent to --pq-exp 1).")
    if args.cts_bin or args.cts_breaks or args.cts_names:
        raise SystemExit("--cts-bin is not supported by this engine (input preparation is "
                         "outside the LD-score path); build the annotation with the "
                         "reference and pass it with --annot.")
    if args.python_compat and args.snp_level_masking:
        raise SystemExit("--python-compat conflicts with --snp-level-masking")
    if args.sketch_maf_aware and args.sketch is None:
        raise SystemExit("--sketch-maf-aware requires --sketch")
    if args.chunk_size < 1:
        raise SystemExit("compute_ldscore_global: chunk_size must be >= 1, got 0")
    maf_pre = str(args.maf_pre).lower() not in ("false", "0", "no")
    pq_exp = 1.0 if args.per_allele else args.pq_exp
    nwind = sum(x is not None for x in (args.ld_wind_cm, args.ld_wind_kb, args.ld_wind_snps))
    if nwind > 1:
        raise SystemExit("Must specify exactly one --ld-wind option")
    if args.ld_wind_kb is not None:
        mode = WindowMode.Kb(args.ld_wind_kb)
    elif args.ld_wind_snps is not None:
        mode = WindowMode.Snp(args.ld_wind_snps)
    else:
        mode = WindowMode.Cm(1.0 if args.ld_wind_cm is None else args.ld_wind_cm)

    bim_path, fam_path, bed_path = (args.bfile + ext for ext in (".bim", ".fam", ".bed"))
    bim = BimTable.from_file(bim_path)
    t_bim = time.perf_counter()
    n_fam = count_fam(fam_path)
    print(f"Loaded {bim.n} SNPs, {n_fam} individuals from '{args.bfile}'")
    sel = bim.bed_idx  # BIM row of each SNP of the LD-score list, in BIM order
    if args.extract:
        ext = load_snp_set(args.extract)
        sel = sel[bim.name_in(ext)]
        print(f"  After --extract: {sel.size} SNPs")
    print_set = load_print_snps(args.print_snps) if args.print_snps else None
    if args.annot and args.extract:
        print("WARNING: --annot with --extract is not supported. Ensure your annot files "
              "match the extracted SNP set.")
    annot, annot_names = (None, None)
    if args.annot:
        annot, annot_names = read_annotations(args.annot, args.thin_annot, bim.chr[sel])
    keep = load_individual_indices(args.keep, parse_fam_ids(fam_path)) if args.keep else None

    t_in = time.perf_counter()
    # mapped, not read: the upload copies straight from the page cache (an
    # empty file cannot be mapped; the panel then reports it too short)
    bed = (np.memmap(bed_path, dtype=np.uint8, mode="r") if os.path.getsize(bed_path) > 0
           else np.zeros(0, np.uint8))
    panel = Panel(bed, n_fam, bim.n, keep=keep)
    del bed
    n_indiv = panel.n_indiv
    t_panel = time.perf_counter()

    # MAF prefilter (src/l2/mod.rs:269-336; snp_stats.rs:150-195)
    pq_scaled = False
    if maf_pre and (args.maf is not None or pq_exp is not None):
        idx = sel.astype(np.uint64)
        _, _, st, _ = decode(panel, idx, dosage=False, miss=False, stats=True, std_f32=False)
        cnt = st[:, 1].astype(np.float64)
        freq = np.where(cnt > 0, st[:, 0] / np.maximum(2.0 * cnt, 1.0), 0.0)
        maf_all = np.minimum(freq, 1.0 - freq)
        het_miss_ok = (st[:, 3].astype(np.int64) + (n_indiv - st[:, 1].astype(np.int64))) < n_indiv
        thr = args.maf if args.maf is not None else 0.0
        keep_mask = het_miss_ok & (maf_all > thr)
        if not keep_mask.any():
            raise SystemExit("SNP prefilter removed all SNPs")
        if not keep_mask.all():
            sel = sel[keep_mask]
            if annot is not None:
                annot = annot[keep_mask]
        if args.maf is not None:
            print(f"--maf-pre: kept {sel.size} / {keep_mask.size} SNPs (MAF > {thr})")
        if pq_exp is not None and annot is not None:
            annot = annot * ((maf_all[keep_mask] * (1.0 - maf_all[keep_mask])) ** pq_exp)[:, None]
            pq_scaled = True

    chunk = 50 if args.python_compat else args.chunk_size
    if args.python_compat and args.chunk_size != 200:
        print(f"--python-compat overrides --chunk-size {args.chunk_size} with c=50 for "
              "Python LDSC parity", file=sys.stderr)
    global_pass = args.global_pass or args.python_compat
    use_f32 = args.fast_f32 or args.sketch is not None
    cfg = L2Config(mode=mode, chunk_size=chunk, use_f32=use_f32, sketch=args.sketch,
                   sketch_maf_aware=args.sketch_maf_aware,
                   snp_level_masking=args.snp_level_masking, yes_really=args.yes_really,
                   pq_exp=None if pq_scaled else pq_exp, verbose_timing=args.verbose_timing,
                   per_chromosome=not global_pass)
    t_c = time.perf_counter()
    snps = make_snp_array(bim.chr[sel], bim.cm[sel], bim.bp[sel], bed_idx=sel)
    l2, maf, perf = compute_ldscore(panel, snps, cfg, annot=annot)
    panel.close()
    t_done = time.perf_counter()
    if args.verbose_timing:
        print(f"[perf] bim_parse={t_bim - t_start:.3f}s", file=sys.stderr)
        print(f"[perf] other_inputs={t_in - t_bim:.3f}s", file=sys.stderr)
        print(f"[perf] bed_read_and_panel={t_panel - t_in:.3f}s", file=sys.stderr)
        print(f"[perf] maf_prefilter={t_c - t_panel:.3f}s", file=sys.stderr)
        print(f"[perf] compute_ldscore_total={t_done - t_c:.3f}s", file=sys.stderr)

    # outputs (src/l2/mod.rs:584-795)
    suffix = "" if pq_exp is None else f"_S{rust_display(pq_exp)}"
    col_names = ([f"L2{suffix}"] if annot is None else [f"{nm}L2{suffix}" for nm in annot_names])
    pq = None if pq_exp is None else (maf * (1.0 - maf)) ** pq_exp
    pq_for_m = None if pq_scaled else pq

    ok = np.ones(sel.size, dtype=bool)
    if args.maf is not None:
        ok &= maf > args.maf
    if print_set is not None:
        ok &= bim.name_in(print_set, sel)
    out_pos = np.flatnonzero(ok)
    if print_set is not None and not out_pos.size:
        raise SystemExit("After merging with --print-snps, no SNPs remain.")

    def seq_sum(x):  # sequential f64 sums, as the reference accumulates M
        x = np.asarray(x, dtype=np.float64)
        return np.cumsum(x, axis=0)[-1] if x.shape[0] else np.zeros(x.shape[1:])

    def m_vals(pos):
        pos = np.asarray(pos, dtype=np.int64)
        big = maf[pos] > 0.05
        if annot is not None:
            base = annot[pos] * (pq_for_m[pos][:, None] if pq_for_m is not None else 1.0)
            return seq_sum(base), seq_sum(base[big])
        if pq is not None:
            return [float(seq_sum(pq[pos]))], [float(seq_sum(pq[pos][big]))]
        return [float(pos.size)], [float(big.sum())]

    out_path = f"{args.out}.l2.ldscore.gz"
    write_ldscore(out_path, bim, sel, l2, out_pos, col_names)
    ma, m5 = m_vals(np.arange(sel.size))
    with open(f"{args.out}.l2.M", "w") as f:
        f.write(format_m_vals(ma))
    with open(f"{args.out}.l2.M_5_50", "w") as f:
        f.write(format_m_vals(m5))
    chr_sel = bim.chr[sel]
    for c in (int(x) for x in np.unique(chr_sel)):
        pos_all = np.flatnonzero(chr_sel == c)
        pos_out = pos_all[ok[pos_all]]
        path = f"{args.out}{c}.l2.ldscore.gz"
        write_ldscore(path, bim, sel, l2, pos_out, col_names)
        ma_c, m5_c = m_vals(pos_all)
        with open(f"{args.out}{c}.l2.M", "w") as f:
            f.write(format_m_vals(ma_c))
        with open(f"{args.out}{c}.l2.M_5_50", "w") as f:
            f.write(format_m_vals(m5_c))
        print(f"chr {c}: {len(pos_out)} SNPs → {path} (M={','.join(rust_display(v) for v in ma_c)}, "
              f"M_5_50={','.join(rust_display(v) for v in m5_c)})")
    print_ld_score_summary(out_path, l2[out_pos], col_names, maf[out_pos])
    if args.verbose_timing:
        t_end = time.perf_counter()
        print(f"[perf] outputs={t_end - t_done:.3f}s", file=sys.stderr)
        print(f"[perf] total={t_end - t_start:.3f}s", file=sys.stderr)
    return 0


def inject_implicit_subcommand(argv: List[str]) -> List[str]:
    """`ldsc --l2 ...` -> `ldsc l2 ...` (src/main.rs:39-78)."""
    if argv and argv[0] == "l2":
        return argv[1:]
    if "--l2" in argv:
        return [a for a in argv if a != "--l2"]
    return argv


def main(argv: Optional[List[str]] = None) -> int:
    argv = list(sys.argv[1:] if argv is None else argv)
    argv = inject_implicit_subcommand(argv)
    try:
        return run(build_parser().parse_args(argv))
    except N.LdscError as e:
        print(f"Error: {e}", file=sys.stderr)
        return 1


if __name__ == "__main__":
    sys.exit(main())
