// io.cpp -- host-side BIM reader and .l2.ldscore.gz writer (include/ldsc_io.h).
//
// BIM: the text is cut into one segment per thread at line starts, each
// segment counts its lines, and a prefix sum gives every segment the line
// index it starts at; the segments are then parsed independently.  The
// first failing line over all segments is the error reported, as a
// sequential reader would report it.
//
// ldscore: the rows of a batch are formatted in blocks of 8192 rows, each
// into its own text buffer, and each block is deflated as one piece of the
// stream, on a thread of its own.  A piece is a raw
// deflate stream primed with the 32 KiB of text before it (the end of the
// previous block); all but the last end on a byte-aligned sync flush, so
// their concatenation is one valid deflate stream and the file is a single
// gzip member (readable by any gzip decoder, including single-member ones).
#include "../../include/ldsc_io.h"

#include <locale.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

namespace {

constexpr int kErrInvalid = 1;

void set_err(char* err, uint64_t err_len, const std::string& msg) {
    if (err && err_len) {
        size_t k = std::min<size_t>(msg.size(), err_len - 1);
        memcpy(err, msg.data(), k);
        err[k] = 0;
    }
}

int threads_or_default(int t) { return t > 0 ? t : omp_get_max_threads(); }

// ---------------------------------------------------------------------------
// Rust `str::parse` grammars
// ---------------------------------------------------------------------------
// u8 / u32: optional '+', then one or more ASCII digits, value in range.
template <typename T>
bool parse_unsigned(const char* s, size_t n, T* out) {
    if (n && s[0] == '+') { ++s; --n; }
    if (!n) return false;
    uint64_t v = 0;
    const uint64_t lim = (uint64_t)(T)~(T)0;
    for (size_t i = 0; i < n; ++i) {
        unsigned d = (unsigned char)s[i] - '0';
        if (d > 9) return false;
        v = v * 10 + d;
        if (v > lim) return false;
    }
    *out = (T)v;
    return true;
}

bool ieq(const char* s, size_t n, const char* lit) {
    size_t k = strlen(lit);
    if (n != k) return false;
    for (size_t i = 0; i < n; ++i)
        if ((s[i] | 0x20) != lit[i]) return false;
    return true;
}

locale_t c_locale() {
    static locale_t loc = newlocale(LC_ALL_MASK, "C", (locale_t)0);
    return loc;
}

// f64: [+-]? ( inf | infinity | nan | digits [. digits?] | . digits ) ([eE][+-]?digits)?
// (inf/infinity/nan without exponent, any case).  The decimal value is
// rounded correctly by strtod_l in the C locale.
bool parse_f64(const char* s, size_t n, double* out) {
    size_t i = 0;
    bool neg = false;
    if (i < n && (s[i] == '+' || s[i] == '-')) { neg = s[i] == '-'; ++i; }
    const char* b = s + i;
    size_t m = n - i;
    if (ieq(b, m, "inf") || ieq(b, m, "infinity")) { *out = neg ? -HUGE_VAL : HUGE_VAL; return true; }
    if (ieq(b, m, "nan")) { *out = neg ? -NAN : NAN; return true; }
    size_t int_d = 0, frac_d = 0;
    while (i < n && (unsigned)(s[i] - '0') <= 9) { ++i; ++int_d; }
    if (i < n && s[i] == '.') {
        ++i;
        while (i < n && (unsigned)(s[i] - '0') <= 9) { ++i; ++frac_d; }
    }
    if (int_d + frac_d == 0) return false;
    if (i < n && (s[i] == 'e' || s[i] == 'E')) {
        ++i;
        if (i < n && (s[i] == '+' || s[i] == '-')) ++i;
        size_t ed = 0;
        while (i < n && (unsigned)(s[i] - '0') <= 9) { ++i; ++ed; }
        if (!ed) return false;
    }
    if (i != n) return false;
    char small[64];
    std::string big;
    const char* z;
    if (n < sizeof(small)) {
        memcpy(small, s, n);
        small[n] = 0;
        z = small;
    } else {
        big.assign(s, n);
        z = big.c_str();
    }
    *out = strtod_l(z, nullptr, c_locale());
    return true;
}

// Strict UTF-8 (no overlongs, no surrogates, at most U+10FFFF).
bool valid_utf8(const unsigned char* s, size_t n) {
    size_t i = 0;
    while (i < n) {
        unsigned c = s[i];
        if (c < 0x80) { ++i; continue; }
        size_t k;
        unsigned lo = 0x80, hi = 0xBF;
        if (c >= 0xC2 && c <= 0xDF) k = 1;
        else if (c == 0xE0) { k = 2; lo = 0xA0; }
        else if (c >= 0xE1 && c <= 0xEC) k = 2;
        else if (c == 0xED) { k = 2; hi = 0x9F; }
        else if (c >= 0xEE && c <= 0xEF) k = 2;
        else if (c == 0xF0) { k = 3; lo = 0x90; }
        else if (c >= 0xF1 && c <= 0xF3) k = 3;
        else if (c == 0xF4) { k = 3; hi = 0x8F; }
        else return false;
        if (i + k >= n) return false;
        unsigned c1 = s[i + 1];
        if (c1 < lo || c1 > hi) return false;
        for (size_t j = 2; j <= k; ++j)
            if ((s[i + j] & 0xC0) != 0x80) return false;
        i += k + 1;
    }
    return true;
}

// One BIM line (without its terminator).  Returns 0, or the error kind:
// 1 not UTF-8, 2 no SNP column, 3 no A1 column, 4 no A2 column.
int parse_bim_line(const char* s, size_t n, uint64_t base_off, uint8_t* chr, double* cm,
                   uint32_t* bp, uint64_t* name_off, uint32_t* name_len) {
    bool ascii = true;
    for (size_t i = 0; i < n; ++i) ascii &= (unsigned char)s[i] < 0x80;
    if (!ascii && !valid_utf8((const unsigned char*)s, n)) return 1;
    const char* f[6];
    size_t fl[6];
    int nf = 0;
    size_t st = 0;
    for (size_t i = 0; i <= n && nf < 6; ++i) {
        if (i == n || s[i] == '\t') {
            f[nf] = s + st;
            fl[nf] = i - st;
            ++nf;
            st = i + 1;
        }
    }
    uint8_t c8 = 0;
    if (!parse_unsigned<uint8_t>(f[0], fl[0], &c8)) c8 = 0;
    if (nf < 2) return 2;
    double c = 0.0;
    if (nf < 3 || !parse_f64(f[2], fl[2], &c)) c = 0.0;
    uint32_t b = 0;
    if (nf < 4 || !parse_unsigned<uint32_t>(f[3], fl[3], &b)) b = 0;
    if (nf < 5) return 3;
    if (nf < 6) return 4;
    *chr = c8;
    *cm = c;
    *bp = b;
    *name_off = base_off + (uint64_t)(f[1] - s);
    *name_len = (uint32_t)fl[1];
    return 0;
}

// ---------------------------------------------------------------------------
// Rust `{:.3}` for f64
// ---------------------------------------------------------------------------
char* put_u64(char* p, uint64_t v) {
    char tmp[24];
    int k = 0;
    do { tmp[k++] = (char)('0' + v % 10); v /= 10; } while (v);
    while (k) *p++ = tmp[--k];
    return p;
}

// av = m * 2^sh exactly, m < 2^53 an integer, so av * 1000 = (m * 1000) * 2^sh
// with m * 1000 < 2^63, and q = round_half_even(av * 1000) follows from
// 128-bit integer shifts.  Returns false when q needs more than 63 bits
// (av above about 9.2e15); the caller then uses snprintf, which glibc also
// rounds exactly.
bool scaled_round(double av, uint64_t* q) {
    int e;
    double fr = frexp(av, &e);             // av = fr * 2^e, fr in [0.5, 1)
    if (av == 0.0) { *q = 0; return true; }
    uint64_t m = (uint64_t)ldexp(fr, 53);  // exact: av = m * 2^(e-53)
    int sh = e - 53;
    unsigned __int128 p = (unsigned __int128)m * 1000u;
    if (sh >= 0) {
        if (sh > 10) return false;
        unsigned __int128 r = p << sh;
        if (r >> 63) return false;
        *q = (uint64_t)r;
        return true;
    }
    int s = -sh;
    if (s >= 127) { *q = 0; return true; }  // p < 2^63 < 2^(s-1): below one half
    unsigned __int128 qq = p >> s;
    unsigned __int128 rem = p - (qq << s);
    unsigned __int128 half = (unsigned __int128)1 << (s - 1);
    if (rem > half || (rem == half && (qq & 1))) ++qq;
    if (qq >> 63) return false;
    *q = (uint64_t)qq;
    return true;
}

size_t fmt3(double v, char* out) {
    if (std::isnan(v)) { memcpy(out, "NaN", 3); return 3; }
    if (std::isinf(v)) {
        if (v < 0) { memcpy(out, "-inf", 4); return 4; }
        memcpy(out, "inf", 3);
        return 3;
    }
    uint64_t q;
    if (!scaled_round(std::fabs(v), &q)) return (size_t)snprintf(out, 330, "%.3f", v);
    char* p = out;
    if (std::signbit(v)) *p++ = '-';
    p = put_u64(p, q / 1000);
    unsigned f = (unsigned)(q % 1000);
    *p++ = '.';
    *p++ = (char)('0' + f / 100);
    *p++ = (char)('0' + f / 10 % 10);
    *p++ = (char)('0' + f % 10);
    return (size_t)(p - out);
}

// ---------------------------------------------------------------------------
// gzip writer: parallel pieces that form one deflate stream
// ---------------------------------------------------------------------------
struct GzOut {
    FILE* f = nullptr;
    uLong crc = 0;
    uint64_t total = 0;
    std::string window;  // the last <= 32 KiB of text written so far
};

constexpr size_t kDict = 32768;

bool write_gz_header(GzOut& g) {
    const unsigned char h[10] = {0x1f, 0x8b, 8, 0, 0, 0, 0, 0, 4, 255};
    return fwrite(h, 1, 10, g.f) == 10;
}

// The kDict bytes of text that come before piece i of `pieces`: the end of
// the pieces before it, then of the text already written (g.window).
static std::string dict_before(const GzOut& g, const std::vector<std::string>& pieces, size_t i) {
    std::string d(kDict, '\0');
    size_t have = 0;  // bytes filled at the end of d
    for (size_t k = i; k-- > 0 && have < kDict;) {
        const std::string& p = pieces[k];
        const size_t take = std::min(p.size(), kDict - have);
        memcpy(&d[kDict - have - take], p.data() + p.size() - take, take);
        have += take;
    }
    if (have < kDict) {
        const size_t take = std::min(g.window.size(), kDict - have);
        memcpy(&d[kDict - have - take], g.window.data() + g.window.size() - take, take);
        have += take;
    }
    return d.substr(kDict - have);
}

// Deflate `pieces` (consecutive text, continuing the stream in g), each on
// its own thread, primed with the 32 KiB of text before it and ended with a
// sync flush, so that the outputs concatenate into one deflate stream;
// `last` closes the stream after the final piece.
bool gz_append(GzOut& g, const std::vector<std::string>& pieces, bool last, int threads,
               std::string& why) {
    const size_t np = pieces.size();
    std::vector<std::string> outs(np);
    std::vector<uLong> crcs(np);
    std::vector<int> ok(np, 1);
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
    for (size_t i = 0; i < np; ++i) {
        const std::string& in = pieces[i];
        if (in.size() > 0xFFFF0000u) { ok[i] = 0; continue; }  // one deflate call's uInt
        z_stream zs;
        memset(&zs, 0, sizeof(zs));
        if (deflateInit2(&zs, 1, Z_DEFLATED, -15, 8, Z_DEFAULT_STRATEGY) != Z_OK) { ok[i] = 0; continue; }
        const std::string dict = dict_before(g, pieces, i);
        if (!dict.empty() &&
            deflateSetDictionary(&zs, (const Bytef*)dict.data(), (uInt)dict.size()) != Z_OK) {
            ok[i] = 0;
            deflateEnd(&zs);
            continue;
        }
        std::string& o = outs[i];
        o.resize(deflateBound(&zs, in.size()) + 64);
        zs.next_in = (Bytef*)in.data();
        zs.avail_in = (uInt)in.size();
        zs.next_out = (Bytef*)&o[0];
        zs.avail_out = (uInt)o.size();
        int flush = (last && i + 1 == np) ? Z_FINISH : Z_SYNC_FLUSH;
        int rc = deflate(&zs, flush);
        if ((flush == Z_FINISH && rc != Z_STREAM_END) || (flush != Z_FINISH && rc != Z_OK) ||
            zs.avail_in != 0)
            ok[i] = 0;
        o.resize(o.size() - zs.avail_out);
        deflateEnd(&zs);
        crcs[i] = crc32(0L, (const Bytef*)in.data(), (uInt)in.size());
    }
    for (size_t i = 0; i < np; ++i) {
        if (!ok[i]) { why = "deflate failed"; return false; }
        g.crc = crc32_combine(g.crc, crcs[i], (z_off_t)pieces[i].size());
        g.total += pieces[i].size();
        if (!outs[i].empty() && fwrite(outs[i].data(), 1, outs[i].size(), g.f) != outs[i].size()) {
            why = "write failed";
            return false;
        }
    }
    if (np) g.window = dict_before(g, pieces, np);
    return true;
}

bool write_gz_trailer(GzOut& g) {
    unsigned char t[8];
    for (int i = 0; i < 4; ++i) t[i] = (unsigned char)(g.crc >> (8 * i));
    for (int i = 0; i < 4; ++i) t[4 + i] = (unsigned char)(g.total >> (8 * i));
    return fwrite(t, 1, 8, g.f) == 8;
}


// ---------------------------------------------------------------------------
// annotation tables (read_tsv + read_annot_path)
// ---------------------------------------------------------------------------
// Length of the `char::is_whitespace` character starting at p (0 if none):
// U+0009-000D, U+0020, U+0085, U+00A0, U+1680, U+2000-200A, U+2028, U+2029,
// U+202F, U+205F, U+3000.
size_t ws_fwd(const unsigned char* p, const unsigned char* end) {
    const size_t n = (size_t)(end - p);
    if (!n) return 0;
    const unsigned c = p[0];
    if (c == ' ' || (c >= 0x09 && c <= 0x0D)) return 1;
    if (c < 0x80) return 0;
    if (n >= 2 && c == 0xC2 && (p[1] == 0x85 || p[1] == 0xA0)) return 2;
    if (n < 3) return 0;
    if (c == 0xE1 && p[1] == 0x9A && p[2] == 0x80) return 3;
    if (c == 0xE2 && p[1] == 0x80 &&
        ((p[2] >= 0x80 && p[2] <= 0x8A) || p[2] == 0xA8 || p[2] == 0xA9 || p[2] == 0xAF))
        return 3;
    if (c == 0xE2 && p[1] == 0x81 && p[2] == 0x9F) return 3;
    if (c == 0xE3 && p[1] == 0x80 && p[2] == 0x80) return 3;
    return 0;
}

// Length of the whitespace character that ends just before `end` (0 if none).
size_t ws_back(const unsigned char* begin, const unsigned char* end) {
    const size_t n = (size_t)(end - begin);
    if (n >= 1 && ws_fwd(end - 1, end) == 1) return 1;
    if (n >= 2 && ws_fwd(end - 2, end) == 2) return 2;
    if (n >= 3 && ws_fwd(end - 3, end) == 3) return 3;
    return 0;
}

// parse_f64_or_nan (src/frame.rs:68-74) after `trim()`, then NaN -> 0 as
// read_annot_path stores it (src/parse.rs:219-223).
double annot_cell(const char* s, size_t n) {
    const unsigned char* b = (const unsigned char*)s;
    const unsigned char* e = b + n;
    for (size_t k; b < e && (k = ws_fwd(b, e)) != 0;) b += k;
    for (size_t k; e > b && (k = ws_back(b, e)) != 0;) e -= k;
    const char* t = (const char*)b;
    const size_t m = (size_t)(e - b);
    if (m == 0 || ieq(t, m, "na") || ieq(t, m, "nan") || (m == 1 && t[0] == '.')) return 0.0;
    double v;
    if (!parse_f64(t, m, &v) || std::isnan(v)) return 0.0;
    return v;
}

// One body row: the first ncol cells (missing ones are empty), columns
// skip.. stored.  Tab mode splits on every '\t'; otherwise on runs of
// whitespace with none leading or trailing (`split_whitespace`).
void annot_row(const char* s, size_t n, bool use_tab, uint64_t ncol, uint64_t skip, double* out) {
    const unsigned char* p = (const unsigned char*)s;
    const unsigned char* end = p + n;
    uint64_t j = 0;
    if (use_tab) {
        const unsigned char* st = p;
        for (;; ++p) {
            if (p == end || *p == '\t') {
                if (j >= skip) out[j - skip] = annot_cell((const char*)st, (size_t)(p - st));
                if (++j == ncol || p == end) break;
                st = p + 1;
            }
        }
    } else {
        while (j < ncol) {
            for (size_t k; p < end && (k = ws_fwd(p, end)) != 0;) p += k;
            if (p == end) break;
            const unsigned char* st = p;
            while (p < end && ws_fwd(p, end) == 0) ++p;
            if (j >= skip) out[j - skip] = annot_cell((const char*)st, (size_t)(p - st));
            ++j;
        }
    }
    for (; j < ncol; ++j)
        if (j >= skip) out[j - skip] = 0.0;
}

// A line of read_tsv: its bytes up to '\n' with every trailing '\r' / '\n'
// removed; empty lines are skipped.
inline size_t trimmed_len(const char* p, size_t l) {
    while (l && (p[l - 1] == '\r' || p[l - 1] == '\n')) --l;
    return l;
}

// Segment starts at line starts, one segment per thread.
std::vector<uint64_t> line_segments(const char* text, uint64_t len, int T) {
    std::vector<uint64_t> seg(T + 1, len);
    seg[0] = 0;
    for (int t = 1; t < T; ++t) {
        const uint64_t a = len * (uint64_t)t / (uint64_t)T;
        const char* q = (const char*)memchr(text + a - 1, '\n', (size_t)(len - a + 1));
        seg[t] = q ? (uint64_t)(q - text) + 1 : len;
        if (seg[t] < seg[t - 1]) seg[t] = seg[t - 1];
    }
    return seg;
}

uint64_t count_nonempty(const char* text, uint64_t len) {
    uint64_t n = 0;
    const char* p = text;
    const char* end = text + len;
    while (p < end) {
        const char* q = (const char*)memchr(p, '\n', (size_t)(end - p));
        const size_t l = q ? (size_t)(q - p) : (size_t)(end - p);
        n += trimmed_len(p, l) != 0;
        p = q ? q + 1 : end;
    }
    return n;
}
}  // namespace

extern "C" uint64_t ldsc_io_count_lines(const char* text, uint64_t len) {
    if (!text || !len) return 0;
    uint64_t n = 0;
    const char* p = text;
    const char* end = text + len;
    while (p < end) {
        const char* q = (const char*)memchr(p, '\n', (size_t)(end - p));
        if (!q) return n + 1;
        ++n;
        p = q + 1;
    }
    return n;
}

extern "C" int ldsc_io_parse_bim(const char* text, uint64_t len, uint64_t n, uint8_t* chr,
                                 double* cm, uint32_t* bp, uint64_t* name_off,
                                 uint32_t* name_len, int threads, char* err, uint64_t err_len) {
    if ((len && !text) || (n && (!chr || !cm || !bp || !name_off || !name_len))) {
        set_err(err, err_len, "ldsc_io_parse_bim: null argument");
        return kErrInvalid;
    }
    if (ldsc_io_count_lines(text, len) != n) {
        set_err(err, err_len, "ldsc_io_parse_bim: n differs from the text's line count");
        return kErrInvalid;
    }
    if (!n) return 0;
    int T = (int)std::min<uint64_t>(threads_or_default(threads), std::max<uint64_t>(1, len >> 16));
    const std::vector<uint64_t> seg = line_segments(text, len, T);
    std::vector<uint64_t> lines(T + 1, 0);
#pragma omp parallel for num_threads(T)
    for (int t = 0; t < T; ++t)
        lines[t + 1] = ldsc_io_count_lines(text + seg[t], seg[t + 1] - seg[t]);
    for (int t = 0; t < T; ++t) lines[t + 1] += lines[t];
    std::vector<uint64_t> bad_line(T, UINT64_MAX);
    std::vector<int> bad_kind(T, 0);
#pragma omp parallel for num_threads(T)
    for (int t = 0; t < T; ++t) {
        uint64_t li = lines[t];
        const char* p = text + seg[t];
        const char* end = text + seg[t + 1];
        while (p < end) {
            const char* q = (const char*)memchr(p, '\n', (size_t)(end - p));
            size_t l = q ? (size_t)(q - p) : (size_t)(end - p);
            if (q && l && p[l - 1] == '\r') --l;  // "\r\n" ends a line; a bare final '\r' stays
            int k = parse_bim_line(p, l, (uint64_t)(p - text), chr + li, cm + li, bp + li,
                                   name_off + li, name_len + li);
            if (k) {
                bad_line[t] = li;
                bad_kind[t] = k;
                break;
            }
            ++li;
            p = q ? q + 1 : end;
        }
    }
    for (int t = 0; t < T; ++t) {
        if (bad_line[t] == UINT64_MAX) continue;
        uint64_t ln = bad_line[t] + 1;
        std::string msg;
        switch (bad_kind[t]) {
            case 1: msg = "reading BIM line " + std::to_string(ln) +
                          ": stream did not contain valid UTF-8"; break;
            case 2: msg = "BIM line " + std::to_string(ln) + ": expected SNP column"; break;
            case 3: msg = "BIM line " + std::to_string(ln) + ": missing A1 col"; break;
            default: msg = "BIM line " + std::to_string(ln) + ": missing A2 col"; break;
        }
        set_err(err, err_len, msg);
        return kErrInvalid;
    }
    return 0;
}

extern "C" uint64_t ldsc_io_count_nonempty_lines(const char* text, uint64_t len) {
    return (text && len) ? count_nonempty(text, len) : 0;
}

extern "C" int ldsc_io_parse_annot(const char* text, uint64_t len, int use_tab, uint64_t ncol,
                                   uint64_t skip, uint64_t n_rows, double* out, int threads,
                                   char* err, uint64_t err_len) {
    if ((len && !text) || skip >= ncol || (n_rows && !out)) {
        set_err(err, err_len, "ldsc_io_parse_annot: bad arguments");
        return kErrInvalid;
    }
    const uint64_t ne = ldsc_io_count_nonempty_lines(text, len);
    if (ne == 0 || ne - 1 != n_rows) {
        set_err(err, err_len, "ldsc_io_parse_annot: n_rows differs from the text's body rows");
        return kErrInvalid;
    }
    const uint64_t width = ncol - skip;
    const int T = (int)std::min<uint64_t>(threads_or_default(threads),
                                          std::max<uint64_t>(1, len >> 16));
    const std::vector<uint64_t> seg = line_segments(text, len, T);
    std::vector<uint64_t> before(T + 1, 0);
#pragma omp parallel for num_threads(T)
    for (int t = 0; t < T; ++t) before[t + 1] = count_nonempty(text + seg[t], seg[t + 1] - seg[t]);
    for (int t = 0; t < T; ++t) before[t + 1] += before[t];
    std::vector<uint64_t> bad(T, UINT64_MAX);
#pragma omp parallel for num_threads(T)
    for (int t = 0; t < T; ++t) {
        uint64_t k = before[t];  // non-empty lines before this segment (the header is k = 0)
        const char* p = text + seg[t];
        const char* end = text + seg[t + 1];
        while (p < end) {
            const char* q = (const char*)memchr(p, '\n', (size_t)(end - p));
            const size_t l0 = q ? (size_t)(q - p) : (size_t)(end - p);
            const size_t l = trimmed_len(p, l0);
            bool ascii = true;
            for (size_t i = 0; i < l0; ++i) ascii &= (unsigned char)p[i] < 0x80;
            if (!ascii && !valid_utf8((const unsigned char*)p, l0)) {
                bad[t] = (uint64_t)(p - text);
                break;
            }
            if (l) {
                if (k > 0) annot_row(p, l, use_tab != 0, ncol, skip, out + (k - 1) * width);
                ++k;
            }
            p = q ? q + 1 : end;
        }
    }
    for (int t = 0; t < T; ++t)
        if (bad[t] != UINT64_MAX) {
            set_err(err, err_len, "stream did not contain valid UTF-8 (line " +
                                      std::to_string(ldsc_io_count_lines(text, bad[t]) + 1) + ")");
            return kErrInvalid;
        }
    return 0;
}

extern "C" uint64_t ldsc_io_format_fixed3(double v, char* out, uint64_t cap) {
    char buf[400];
    size_t k = fmt3(v, buf);
    if (out && cap) {
        size_t c = std::min<size_t>(k, cap - 1);
        memcpy(out, buf, c);
        out[c] = 0;
    }
    return k;
}

extern "C" int ldsc_io_write_ldscore(const char* path, const char* header_tail, uint64_t n_snps,
                                     const uint8_t* chr, const uint32_t* bp, const char* names,
                                     uint64_t names_len, const uint64_t* name_off,
                                     const uint32_t* name_len, const double* l2, uint64_t K,
                                     const int64_t* rows, uint64_t n_rows, int threads,
                                     char* err, uint64_t err_len) {
    if (!path || !header_tail || (n_rows && (!chr || !bp || !names || !name_off || !name_len ||
                                             !rows || (K && !l2)))) {
        set_err(err, err_len, "ldsc_io_write_ldscore: null argument");
        return kErrInvalid;
    }
    int T = threads_or_default(threads);
    // every row index and name span must lie inside the arrays described
    int bad = 0;
#pragma omp parallel for reduction(| : bad) num_threads(T)
    for (uint64_t i = 0; i < n_rows; ++i) {
        const int64_t r = rows[i];
        if (r < 0 || (uint64_t)r >= n_snps) {
            bad |= 1;
        } else {
            const uint64_t o = name_off[r];
            bad |= (o > names_len || names_len - o < name_len[r]) ? 2 : 0;
        }
    }
    if (bad) {
        set_err(err, err_len, (bad & 1) ? "ldsc_io_write_ldscore: row index out of range"
                                        : "ldsc_io_write_ldscore: SNP ID span outside the names");
        return kErrInvalid;
    }
    GzOut g;
    g.f = fopen(path, "wb");
    if (!g.f) {
        set_err(err, err_len, std::string("creating output '") + path + "'");
        return kErrInvalid;
    }
    std::string why;
    bool ok = write_gz_header(g);
    const std::vector<std::string> head{std::string("CHR\tSNP\tBP") + header_tail + "\n"};
    // rows per formatting block (one deflate piece each) and per batch
    const uint64_t kBlock = 8192;
    const uint64_t kBatch = kBlock * 128;
    ok = ok && gz_append(g, head, n_rows == 0, T, why);
    std::vector<std::string> blocks;
    for (uint64_t b0 = 0; ok && b0 < n_rows; b0 += kBatch) {
        const uint64_t b1 = std::min(n_rows, b0 + kBatch);
        const uint64_t nb = (b1 - b0 + kBlock - 1) / kBlock;
        blocks.resize(nb);
#pragma omp parallel for schedule(dynamic, 1) num_threads(T)
        for (uint64_t bi = 0; bi < nb; ++bi) {
            std::string& s = blocks[bi];
            const uint64_t r0 = b0 + bi * kBlock, r1 = std::min(b1, r0 + kBlock);
            size_t used = 0;
            for (uint64_t i = r0; i < r1; ++i) {
                const int64_t r = rows[i];
                // a row is at most 3 + 1 + name + 1 + 10 + K * (1 + 330) + 1 bytes
                const size_t need = 16 + (size_t)name_len[r] + 16 + (size_t)K * 332;
                if (s.size() - used < need) s.resize(std::max(2 * s.size(), used + need));
                char* p = &s[used];
                p = put_u64(p, chr[r]);
                *p++ = '\t';
                memcpy(p, names + name_off[r], name_len[r]);
                p += name_len[r];
                *p++ = '\t';
                p = put_u64(p, bp[r]);
                for (uint64_t k = 0; k < K; ++k) {
                    *p++ = '\t';
                    p += fmt3(l2[(uint64_t)r * K + k], p);
                }
                *p++ = '\n';
                used = (size_t)(p - s.data());
            }
            s.resize(used);
        }
        ok = gz_append(g, blocks, b1 == n_rows, T, why);
        if (ok && b1 < n_rows)
            for (std::string& s : blocks) s.clear();
    }
    ok = ok && write_gz_trailer(g);
    if (fclose(g.f) != 0) ok = false;
    if (!ok) {
        set_err(err, err_len, std::string("writing '") + path + "': " +
                                  (why.empty() ? "I/O error" : why));
        return kErrInvalid;
    }
    return 0;
}
