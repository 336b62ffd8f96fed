"""ldsc_amd -- MI355X-native LD-score (`l2`) engine of sharifhsn/ldsc.

The compute path is libldsc_l2.so (HIP for gfx950) behind include/ldsc_l2.h;
this package is the host mirror of the reference's l2 library API.
"""
from .l2 import (BimRecord, L2Config, L2Output, L2Perf, L2Progress, Panel, WindowMode,
                 block_left, compute_l2_from_bed, compute_l2_from_bed_with_progress,
                 compute_l2_from_bytes, compute_l2_from_bytes_with_progress, compute_ldscore,
                 count_fam_str, countsketch_hash, decode, make_snp_array, pair_bounds,
                 parse_bim_str, sketch_buckets, snps_to_array, unbiased_r2_constants)
from ._native import LdscError

__all__ = [
    "BimRecord", "L2Config", "L2Output", "L2Perf", "L2Progress", "Panel", "WindowMode",
    "block_left", "compute_l2_from_bed", "compute_l2_from_bed_with_progress",
    "compute_l2_from_bytes", "compute_l2_from_bytes_with_progress", "compute_ldscore",
    "count_fam_str", "countsketch_hash", "decode", "make_snp_array", "pair_bounds",
    "parse_bim_str", "sketch_buckets", "snps_to_array", "unbiased_r2_constants", "LdscError",
]
