/*
 * ldsc_io.h — C ABI of the host-side file formats around the LD-score path.
 *
 * libldsc_io.so is plain C++ (no device code).  It reads the PLINK `.bim`
 * text and the `.annot` tables, and writes the `.l2.ldscore.gz` tables of the `l2` subcommand on
 * all host cores: a 1.66 M-SNP BIM parses in about 0.1 s on 8 cores, and
 * the gzip writer formats and deflates in parallel pieces that form one
 * standard gzip member.
 *
 * Error convention as in ldsc_l2.h: 0 on success, LDSC_ERR_INVALID_ARG (1)
 * otherwise, with the message written into `err`.
 */
#ifndef LDSC_IO_H
#define LDSC_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/*
 * Number of lines `BufRead::lines` yields for `text` (split on '\n'; a
 * final line without '\n' counts; no trailing empty record).
 * Replaces the line loop of parse_bim_reader_filtered
 * (src/parse.rs:550-592) and count_fam_str (src/l2/io.rs:16-22).
 */
uint64_t ldsc_io_count_lines(const char* text, uint64_t len);

/*
 * parse_bim_reader_filtered (src/parse.rs:550-592) with every chromosome
 * kept, over an in-memory BIM of `n` lines (ldsc_io_count_lines):
 *   chr = column 1 as u8, else 0;  cm = column 3 as f64, else 0;
 *   bp  = column 4 as u32, else 0; the SNP ID (column 2) is returned as a
 *   byte span of `text` (name_off[i], name_len[i]); bed_idx is the line
 *   index.  A line without columns 2, 5 or 6, or that is not UTF-8, is an
 *   error naming the first such line, with the reference's messages.
 * Field grammars are Rust's `str::parse` for u8 / u32 / f64.
 * `threads` = 0 uses the OpenMP default.
 */
int ldsc_io_parse_bim(const char* text, uint64_t len, uint64_t n,
                      uint8_t* chr, double* cm, uint32_t* bp,
                      uint64_t* name_off, uint32_t* name_len,
                      int threads, char* err, uint64_t err_len);

/*
 * write_ldscore_refs (src/l2/io.rs:94-125): a gzip TSV (deflate level 1,
 * as `Compression::fast()`), header "CHR\tSNP\tBP" + `header_tail`
 * (e.g. "\tL2"), then for each r in rows[0..n_rows):
 *   chr[r] \t names[name_off[r]..+name_len[r]] \t bp[r] \t l2[r*K+k] (k<K)
 * with values written as Rust's `{:.3}` (exact decimal value, rounded
 * half to even; "NaN", "inf", "-inf"; a negative value that rounds to zero
 * keeps its sign).  chr, bp, name_off, name_len and the rows of `l2`
 * (row-major, K columns) describe n_snps SNPs; `names` holds names_len
 * bytes.  A row index or name span outside them is an error, and no file
 * is created then.
 */
int ldsc_io_write_ldscore(const char* path, const char* header_tail, uint64_t n_snps,
                          const uint8_t* chr, const uint32_t* bp,
                          const char* names, uint64_t names_len, const uint64_t* name_off,
                          const uint32_t* name_len, const double* l2, uint64_t K,
                          const int64_t* rows, uint64_t n_rows, int threads,
                          char* err, uint64_t err_len);

/*
 * Lines of `text` that are not empty once every trailing '\r' / '\n' is
 * removed: the header plus body rows of read_tsv (src/frame.rs:319-344).
 */
uint64_t ldsc_io_count_nonempty_lines(const char* text, uint64_t len);

/*
 * The body pass of read_tsv (src/frame.rs:319-413) and read_annot_path
 * (src/parse.rs:191-229) over a decompressed `.annot` text: the first
 * non-empty line is the header (its `ncol` columns and the separator --
 * `use_tab` when the header holds a tab, else runs of whitespace -- are
 * found by the caller); every later non-empty line is a row, padded with
 * empty cells to ncol and cut at ncol.  Cell (i, j) for j >= skip goes to
 * out[i * (ncol - skip) + (j - skip)]: trimmed, NA / NaN / "." / empty or
 * not a Rust f64 -> 0, else the value.  n_rows = count_nonempty - 1.
 * Invalid UTF-8 is an error naming the line.
 */
int ldsc_io_parse_annot(const char* text, uint64_t len, int use_tab, uint64_t ncol,
                        uint64_t skip, uint64_t n_rows, double* out, int threads,
                        char* err, uint64_t err_len);

/*
 * Rust `{:.3}` of one f64 into `out` (at least 32 bytes for |v| < 1e25,
 * 330 bytes for any v); returns the length.  Exposed for the tests.
 */
uint64_t ldsc_io_format_fixed3(double v, char* out, uint64_t cap);

#ifdef __cplusplus
}
#endif

#endif /* LDSC_IO_H */
