/*
 * synth.c -- seeded synthetic PLINK panels for the LD-score workloads.
 *
 * Genotypes come from two haplotypes per individual.  Each haplotype is a
 * Markov chain over SNPs that copies one of `n_founders` founder haplotypes
 * and switches to a uniformly drawn founder with probability `switch_p` per
 * SNP.  Founder allele bits are Bernoulli(MAF_j) with MAF_j ~ U(0.01, 0.5).
 * Dosage (count of A1) = sum of the two haplotype bits, encoded the PLINK
 * way (2 -> 00, 1 -> 10, 0 -> 11); missing genotypes (01) are iid with rate
 * `miss_rate`.  bp gaps are Exp(mean 2000) rounded to integers (bp starts
 * at 1) and cM = bp * 1e-6.
 *
 * Deterministic for a given seed regardless of thread count: every block of
 * 64 individuals draws from its own splitmix64 stream.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static inline uint64_t splitmix64(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static inline double unif(uint64_t* s) { /* (0, 1] */
    return ((splitmix64(s) >> 11) + 1) * (1.0 / 9007199254740992.0);
}
/* number of trials until the next success of a Bernoulli(p), >= 1 */
static inline uint64_t geom(uint64_t* s, double log1mp) {
    if (log1mp == 0.0) return UINT64_MAX / 4;
    double u = unif(s);
    double k = floor(log(u) / log1mp);
    if (k > 1e15) return UINT64_MAX / 4;
    return (uint64_t)k + 1;
}

/* SNP metadata + founder matrix: maf[m], bp[m], cm[m], founders[m] (bit f =
 * allele of founder f). */
void synth_snps(uint64_t seed, uint64_t m, uint32_t n_founders, double* maf, uint32_t* bp,
                double* cm, uint32_t* founders) {
    uint64_t s = seed ^ 0x5EEDF00DULL;
    uint64_t pos = 1;
    for (uint64_t j = 0; j < m; ++j) {
        double mf = 0.01 + 0.49 * unif(&s);
        maf[j] = mf;
        uint32_t bits = 0;
        for (uint32_t f = 0; f < n_founders; ++f)
            if (unif(&s) <= mf) bits |= 1u << f;
        founders[j] = bits;
        if (j > 0) {
            double gap = -2000.0 * log(unif(&s));
            pos += (uint64_t)llround(gap);
        }
        bp[j] = (uint32_t)pos;
        cm[j] = (double)pos * 1e-6;
    }
}

/* Packed SNP-major .bed body (no header): m rows of ceil(n/4) bytes. */
void synth_bed(uint64_t seed, uint64_t n, uint64_t m, uint32_t n_founders, double switch_p,
               double miss_rate, const uint32_t* founders, uint8_t* out) {
    const uint64_t bps = (n + 3) / 4;
    const uint64_t nblk = (n + 63) / 64;
    const double lsw = switch_p > 0 ? log1p(-switch_p) : 0.0;
    const double lms = miss_rate > 0 ? log1p(-miss_rate) : 0.0;
#pragma omp parallel for schedule(dynamic, 1)
    for (uint64_t blk = 0; blk < nblk; ++blk) {
        uint64_t s = seed * 0x2545F4914F6CDD1DULL + blk * 0x9E3779B97F4A7C15ULL + 17;
        const uint64_t i0 = blk * 64;
        const uint64_t cnt = (n - i0) < 64 ? (n - i0) : 64;
        uint32_t fnd[128];
        uint64_t next_sw[128];
        uint64_t next_miss[64];
        for (uint64_t h = 0; h < 2 * cnt; ++h) {
            fnd[h] = (uint32_t)(splitmix64(&s) % n_founders);
            next_sw[h] = geom(&s, lsw);
        }
        for (uint64_t q = 0; q < cnt; ++q) next_miss[q] = geom(&s, lms) - 1;
        uint8_t row[16];
        for (uint64_t j = 0; j < m; ++j) {
            memset(row, 0, sizeof(row));
            const uint32_t fb = founders[j];
            for (uint64_t q = 0; q < cnt; ++q) {
                for (int hh = 0; hh < 2; ++hh) {
                    uint64_t h = 2 * q + hh;
                    if (j > 0 && next_sw[h] == j) {
                        fnd[h] = (uint32_t)(splitmix64(&s) % n_founders);
                        next_sw[h] = j + geom(&s, lsw);
                    }
                }
                uint32_t dos = ((fb >> fnd[2 * q]) & 1u) + ((fb >> fnd[2 * q + 1]) & 1u);
                uint32_t code = dos == 2 ? 0u : (dos == 1 ? 2u : 3u);
                if (next_miss[q] == j) {
                    code = 1u;
                    next_miss[q] = j + geom(&s, lms);
                }
                row[q >> 2] |= (uint8_t)(code << (2 * (q & 3)));
            }
            const uint64_t nb = (cnt + 3) / 4;
            memcpy(out + j * bps + i0 / 4, row, nb);
        }
    }
}

/* Annotation matrix (m x k column-major): n_bin binary columns in contiguous
 * runs (run length ~ Geom(mean run_mean), column density ~ U(0.005, 0.2))
 * followed by n_cont LogNormal(0, 1) columns. */
void synth_annot(uint64_t seed, uint64_t m, uint32_t n_bin, uint32_t n_cont, double run_mean,
                 double* out) {
    for (uint32_t k = 0; k < n_bin + n_cont; ++k) {
        uint64_t s = seed * 0x9E3779B97F4A7C15ULL + (k + 1) * 0xD1B54A32D192ED03ULL;
        double* col = out + (uint64_t)k * m;
        if (k < n_bin) {
            double dens = 0.005 + 0.195 * unif(&s);
            /* alternate off/on runs: on-runs mean run_mean, off-runs sized so the
             * long-run fraction of ones equals dens */
            double off_mean = run_mean * (1.0 - dens) / dens;
            double l_on = log1p(-1.0 / run_mean);
            double l_off = log1p(-1.0 / (off_mean > 1.0 ? off_mean : 1.0000001));
            uint64_t j = 0;
            int on = unif(&s) < dens;
            while (j < m) {
                uint64_t len = geom(&s, on ? l_on : l_off);
                for (uint64_t t = 0; t < len && j < m; ++t, ++j) col[j] = on ? 1.0 : 0.0;
                on = !on;
            }
        } else {
            for (uint64_t j = 0; j < m; ++j) {
                double u1 = unif(&s), u2 = unif(&s);
                double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
                col[j] = exp(z);
            }
        }
    }
}
