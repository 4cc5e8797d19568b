/*
 * io.cpp - the .dat formats of dsmc_suite (src/common/fileio.c, include/fileio.h:8-44)
 * with the same names, signatures and printf formats, and the drop-in drivers
 * run_ballistic / run_advanced (ballistic.c:35-309, advanced.c:119-476) built on
 * the device path.
 *
 * Readers differ from the reference in one respect: a missing, short or
 * malformed file is reported on stderr and returns NULL / -1, where the
 * reference leaves zeros behind fscanf failures (fileio.c:11-21).
 *
 * The text work runs on host threads: a writer formats contiguous chunks of
 * the values concurrently and writes the chunks in order (the bytes of one
 * fprintf per value with the reference's formats; fixed-point conversions are
 * done exactly in integers, other values by snprintf); a reader loads the file,
 * splits it at whitespace and converts chunks concurrently (plain decimals by
 * Clinger's exact fast path, anything else by strtod: the values fscanf's %lf
 * gives).  The fileio functions use up to 16 threads; the drop-in drivers use
 * their `nthreads`.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <errno.h>
#include <limits.h>
#include <atomic>
#include <thread>
#include <vector>
#include <string>

#include "../../include/dsmc_hip.h"

/* ------------------------------------------------------------------ parallel text */
static std::atomic<int> g_io_threads{0};     /* 0: the machine's cores, at most 16 */

static int default_io_threads()
{
    int n = g_io_threads.load(std::memory_order_relaxed);
    if (n > 0) return n;
    unsigned h = std::thread::hardware_concurrency();
    return h == 0 ? 1 : (h > 16 ? 16 : (int)h);
}

extern "C" int dsmc_io_threads(int n)
{
    return g_io_threads.exchange(n < 0 ? 0 : (n > 256 ? 256 : n), std::memory_order_relaxed);
}

/* Runs body(t, lo, hi) for `threads` contiguous ranges of [0, n) (one range
 * below `min_per_thread` items per thread). */
template <class F>
static void for_chunks(size_t n, int threads, size_t min_per_thread, F body)
{
    size_t t = threads < 1 ? 1 : (size_t)threads;
    if (t > 1 && n / t < min_per_thread) t = n / min_per_thread > 0 ? n / min_per_thread : 1;
    if (t <= 1) { body(0, (size_t)0, n); return; }
    std::vector<std::thread> pool;
    for (size_t i = 1; i < t; i++) pool.emplace_back(body, i, n * i / t, n * (i + 1) / t);
    body(0, (size_t)0, n / t);
    for (auto &th : pool) th.join();
}

/* "%<width>.<prec>f" of v (prec <= 7), the bytes glibc's printf writes in the
 * default rounding mode: the exact binary value m 2^-k times 10^prec, rounded
 * to the nearest integer with ties to even (a double can sit exactly halfway,
 * e.g. 1/128 at six decimals), in 128-bit integers; then sign (kept for -0 and
 * for negatives that round to zero), digits and left padding.  Returns the
 * length, or -1 for values it leaves to snprintf (|v| >= 1e12, non-finite). */
static int fixed_fmt(char *out, double v, int prec, int width)
{
    static const uint64_t P10[8] = {1, 10, 100, 1000, 10000, 100000, 1000000, 10000000};
    if (!(fabs(v) < 1e12) || prec < 0 || prec > 7) return -1;
    const bool neg = signbit(v);
    const double a = fabs(v);
    uint64_t q = 0;
    if (a != 0.0) {
        int e;
        const double fr = frexp(a, &e);                   /* a = fr 2^e, fr in [0.5, 1) */
        const uint64_t m = (uint64_t)ldexp(fr, 53);        /* exact: 53-bit integer */
        const int k = 53 - e;                              /* a = m 2^-k, k > 0 below 2^53 */
        const unsigned __int128 num = (unsigned __int128)m * P10[prec];
        if (k >= 120) q = 0;                               /* num < 2^77 < half */
        else {
            q = (uint64_t)(num >> k);
            const unsigned __int128 rem = num & ((((unsigned __int128)1) << k) - 1);
            const unsigned __int128 half = ((unsigned __int128)1) << (k - 1);
            if (rem > half || (rem == half && (q & 1))) q++;
        }
    }
    char tmp[48];
    int n = 0;
    uint64_t ip = q / P10[prec], fp = q % P10[prec];
    for (int i = 0; i < prec; i++) { tmp[n++] = (char)('0' + fp % 10); fp /= 10; }
    if (prec > 0) tmp[n++] = '.';
    do { tmp[n++] = (char)('0' + ip % 10); ip /= 10; } while (ip);
    if (neg) tmp[n++] = '-';
    int len = 0;
    for (int i = n; i < width; i++) out[len++] = ' ';
    while (n) out[len++] = tmp[--n];
    return len;
}

static void append_fixed(std::string &o, double v, int prec, int width)
{
    char buf[64];
    int k = fixed_fmt(buf, v, prec, width);
    if (k < 0) {
        char fmt[16];
        snprintf(fmt, sizeof fmt, "%%%d.%df", width, prec);
        k = snprintf(buf, sizeof buf, fmt, v);
        if (k < 0) return;
        if ((size_t)k >= sizeof buf) {
            std::vector<char> big(k + 1);
            snprintf(big.data(), big.size(), fmt, v);
            o.append(big.data(), k);
            return;
        }
    }
    o.append(buf, k);
}

static void append_int(std::string &o, int v)
{
    char tmp[16];
    int n = 0;
    unsigned u = v < 0 ? 0u - (unsigned)v : (unsigned)v;
    do { tmp[n++] = (char)('0' + u % 10); u /= 10; } while (u);
    if (v < 0) tmp[n++] = '-';
    while (n) o.push_back(tmp[--n]);
}

/* isspace of the C locale (the locale the reference's fscanf runs in) */
static inline bool is_ws(char c) { return c == ' ' || (c >= '\t' && c <= '\r'); }

/* strtol base 10 into an int (fscanf "%d" of a plain token), false on overflow */
static bool parse_int(const char *s, char **end, int *out)
{
    const char *c = s;
    while (is_ws(*c)) c++;
    bool neg = false;
    if (*c == '+' || *c == '-') neg = *c++ == '-';
    if (!(*c >= '0' && *c <= '9')) { *end = (char *)s; return true; }
    int64_t v = 0;
    for (; *c >= '0' && *c <= '9'; c++) {
        v = v * 10 + (*c - '0');
        if (v > (int64_t)INT_MAX + 1) return false;
    }
    v = neg ? -v : v;
    if (v < INT_MIN || v > INT_MAX) return false;
    *out = (int)v;
    *end = (char *)c;
    return true;
}

/* strtod of a plain decimal token ([+-]digits[.digits], at most 15 significant
 * digits, at most 22 after the point): the digits as an exact integer divided
 * by an exact power of ten, one correctly rounded operation, hence strtod's
 * value (Clinger's fast path); anything else goes to strtod. */
static double fast_strtod(const char *s, char **end)
{
    static const double P10[23] = {1e0, 1e1, 1e2, 1e3, 1e4, 1e5, 1e6, 1e7, 1e8, 1e9, 1e10, 1e11,
                                   1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
    const char *c = s;
    while (is_ws(*c)) c++;
    bool neg = false;
    if (*c == '+' || *c == '-') neg = *c++ == '-';
    uint64_t m = 0;
    int sig = 0, frac = 0, digits = 0;
    bool point = false;
    for (;; c++) {
        if (*c >= '0' && *c <= '9') {
            digits++;
            if (m || *c != '0') { m = m * 10 + (uint64_t)(*c - '0'); if (++sig > 15) break; }
            if (point) frac++;
        } else if (*c == '.' && !point) point = true;
        else break;
    }
    const char nx = *c;
    if (digits == 0 || sig > 15 || frac > 22 || nx == 'e' || nx == 'E' || nx == 'x' || nx == 'X' ||
        nx == '.' || (nx >= '0' && nx <= '9'))
        return strtod(s, end);
    *end = (char *)c;
    double v = (double)m / P10[frac];
    return neg ? -v : v;
}

/* Formats items [0, n) with item(i, out) on up to `threads` threads and writes
 * the chunks to fp in order.  Returns 0 or -1. */
template <class F>
static int write_items(FILE *fp, size_t n, int threads, size_t bytes_per_item, F item)
{
    const size_t t = threads < 1 ? 1 : (size_t)threads;
    std::vector<std::string> part(t);
    std::vector<char> ok(t, 1);
    for_chunks(n, (int)t, 16384, [&](size_t c, size_t lo, size_t hi) {
        std::string &o = part[c];
        o.reserve((hi - lo) * bytes_per_item + 16);
        for (size_t i = lo; i < hi; i++)
            if (!item(i, o)) { ok[c] = 0; return; }
    });
    for (size_t c = 0; c < t; c++) {
        if (!ok[c]) return -1;
        if (!part[c].empty() && fwrite(part[c].data(), 1, part[c].size(), fp) != part[c].size()) return -1;
    }
    return 0;
}

/* A whole file in memory, NUL-terminated (strtod stops there). */
static bool slurp(const char *filename, std::string &text)
{
    FILE *fp = fopen(filename, "rb");
    if (!fp) return false;
    char buf[1 << 16];
    size_t k;
    while ((k = fread(buf, 1, sizeof buf, fp)) > 0) text.append(buf, k);
    bool err = ferror(fp) != 0;
    fclose(fp);
    return !err;
}

/* Reads `count` leading integers of text at *pos (fscanf "%d"), advancing *pos. */
static bool parse_ints(const std::string &text, size_t *pos, int *out, int count)
{
    const char *s = text.c_str() + *pos;
    for (int i = 0; i < count; i++) {
        char *e;
        errno = 0;
        long v = strtol(s, &e, 10);
        if (e == s || errno == ERANGE || v < INT_MIN || v > INT_MAX) return false;
        out[i] = (int)v;
        s = e;
    }
    *pos = (size_t)(s - text.c_str());
    return true;
}

/* Converts the first `count` whitespace-separated tokens of text[pos..] with
 * conv(k, token, &end) (k = token index; end must reach the token's end) on up
 * to `threads` threads: pass 1 counts each chunk's tokens (chunks split at
 * whitespace), pass 2 converts them at their global indices.  Returns the
 * number converted in order before the first missing or malformed token. */
template <class F>
static size_t parse_tokens(const std::string &text, size_t pos, size_t count, int threads, F conv)
{
    const char *base = text.c_str();
    const size_t len = text.size();
    size_t t = threads < 1 ? 1 : (size_t)threads;
    if (len - pos < t * 65536) t = (len - pos) / 65536 > 0 ? (len - pos) / 65536 : 1;
    std::vector<size_t> cut(t + 1);
    cut[0] = pos;
    cut[t] = len;
    for (size_t c = 1; c < t; c++) {
        size_t q = pos + (len - pos) * c / t;
        if (q < cut[c - 1]) q = cut[c - 1];
        while (q < len && !is_ws(base[q])) q++;
        cut[c] = q;
    }
    std::vector<size_t> ntok(t, 0), good(t, 0);
    for_chunks(t, (int)t, 1, [&](size_t, size_t lo, size_t hi) {
        for (size_t c = lo; c < hi; c++) {
            size_t k = 0;
            bool in = false;
            for (size_t i = cut[c]; i < cut[c + 1]; i++) {
                bool w = is_ws(base[i]);
                if (!w && !in) k++;
                in = !w;
            }
            ntok[c] = k;
        }
    });
    std::vector<size_t> first(t + 1, 0);
    for (size_t c = 0; c < t; c++) first[c + 1] = first[c] + ntok[c];
    for_chunks(t, (int)t, 1, [&](size_t, size_t lo, size_t hi) {
        for (size_t c = lo; c < hi; c++) {
            size_t i = cut[c], k = first[c];
            while (k < count && k < first[c + 1]) {
                while (i < cut[c + 1] && is_ws(base[i])) i++;
                size_t j = i;
                while (j < cut[c + 1] && !is_ws(base[j])) j++;
                char *e = nullptr;
                if (!conv(k, base + i, &e) || e != base + j) break;
                good[c]++;
                k++;
                i = j;
            }
        }
    });
    size_t done = 0;
    for (size_t c = 0; c < t && done < count; c++) {
        done += good[c];
        if (good[c] < ntok[c] && first[c] + good[c] < count) break;
    }
    return done < count ? done : count;
}

/* `count` doubles after the header (fscanf "%lf" each). */
static size_t parse_doubles(const std::string &text, size_t pos, double *out, size_t count, int threads)
{
    return parse_tokens(text, pos, count, threads, [out](size_t k, const char *tok, char **end) {
        out[k] = fast_strtod(tok, end);
        return true;
    });
}

static double *read_density_n(const char *filename, int *ftd, int *ffd, int *hd, int has_hd, int threads)
{
    std::string text;
    if (!slurp(filename, text)) { fprintf(stderr, "Cannot open %s\n", filename); return NULL; }
    size_t pos = 0;
    int h[3] = {0, 0, 0};
    if (!parse_ints(text, &pos, h, has_hd ? 3 : 2) || h[0] <= 0 || h[1] <= 0 || (has_hd && h[2] <= 0)) {
        fprintf(stderr, "%s: malformed density header\n", filename);
        return NULL;
    }
    *ftd = h[0];
    *ffd = h[1];
    if (has_hd) *hd = h[2];
    if (*hd <= 0) {
        fprintf(stderr, "%s: malformed density header\n", filename);
        return NULL;
    }
    size_t total = (size_t)(*hd) * (*ftd) * (*ffd);
    double *data = (double *)calloc(total, sizeof(double));
    size_t got = parse_doubles(text, pos, data, total, threads);
    if (got < total) {
        fprintf(stderr, "%s: short density file (%zu of %zu values)\n", filename, got, total);
        free(data);
        return NULL;
    }
    return data;
}

extern "C" double *read_density_dat(const char *filename, int *ftd, int *ffd, int *hd, int has_hd)
{
    return read_density_n(filename, ftd, ffd, hd, has_hd, default_io_threads());
}

/* n doubles, one "%<width>.<prec>f\n" each. */
static int write_values(FILE *fp, int width, int prec, const double *v, size_t n, int threads)
{
    return write_items(fp, n, threads, 12, [&](size_t i, std::string &o) {
        append_fixed(o, v[i], prec, width);
        o.push_back('\n');
        return true;
    });
}

static int write_density_n(const char *filename, const double *den, int ftd, int ffd, int hd, int threads)
{
    FILE *fp = fopen(filename, "w");
    if (!fp) return -1;
    fprintf(fp, "%d\n%d\n%d\n", ftd, ffd, hd);
    int rc = write_values(fp, 9, 6, den, (size_t)hd * ftd * ffd, threads);
    if (fclose(fp)) rc = -1;
    return rc;
}

extern "C" int write_density_dat(const char *filename, const double *den, int ftd, int ffd, int hd)
{
    return write_density_n(filename, den, ftd, ffd, hd, default_io_threads());
}

extern "C" double *read_temperature_dat(const char *filename, int *ftd, int *ffd)
{
    std::string text;
    if (!slurp(filename, text)) { fprintf(stderr, "Cannot open %s\n", filename); return NULL; }
    size_t pos = 0;
    int h[2];
    if (!parse_ints(text, &pos, h, 2) || h[0] <= 0 || h[1] <= 0) {
        fprintf(stderr, "%s: malformed temperature header\n", filename);
        return NULL;
    }
    *ftd = h[0];
    *ffd = h[1];
    size_t total = (size_t)(*ftd) * (*ffd);
    double *data = (double *)calloc(total, sizeof(double));
    size_t got = parse_doubles(text, pos, data, total, default_io_threads());
    if (got < total) {
        fprintf(stderr, "%s: short temperature file (%zu of %zu values)\n", filename, got, total);
        free(data);
        return NULL;
    }
    return data;
}

/* fileio.c:64-109: ffd copies of the field, each shifted one column further. */
extern "C" double *read_temperature_time_dat(const char *filename, int *ftd, int *ffd)
{
    int a, b;
    double *T = read_temperature_dat(filename, &a, &b);
    if (!T) return NULL;
    *ftd = a;
    *ffd = b;
    double *out = (double *)calloc((size_t)b * a * b, sizeof(double));
    for (int t = 0; t < b; t++)
        for (int td = 0; td < a; td++)
            for (int fd = 0; fd < b; fd++)
                out[(size_t)t * a * b + td * b + fd] = T[td * b + (fd + t) % b];
    free(T);
    return out;
}

static int write_ttime_n(const char *filename, const double *ttime, int nrows, int totaln, int threads)
{
    FILE *fp = fopen(filename, "w");
    if (!fp) return -1;
    int rc = write_values(fp, 9, 5, ttime, (size_t)nrows * totaln, threads);
    if (fclose(fp)) rc = -1;
    return rc;
}

extern "C" int write_ttime_dat(const char *filename, const double *ttime, int nrows, int totaln)
{
    return write_ttime_n(filename, ttime, nrows, totaln, default_io_threads());
}

static int write_v_record_n(const char *filename, const double *v_record, const int *td_list, const int *fd_list,
                            const int *ht_list, int nrecords, int threads)
{
    FILE *fp = fopen(filename, "w");
    if (!fp) return -1;
    fprintf(fp, "%d\n", nrecords);
    /* "%f %f %f\n%d\n%d\n%d\n" (fileio.c) */
    int rc = write_items(fp, nrecords < 0 ? 0 : (size_t)nrecords, threads, 48, [&](size_t i, std::string &o) {
        for (int j = 0; j < 3; j++) {
            append_fixed(o, v_record[i * 3 + j], 6, 0);
            o.push_back(j < 2 ? ' ' : '\n');
        }
        append_int(o, td_list[i]); o.push_back('\n');
        append_int(o, fd_list[i]); o.push_back('\n');
        append_int(o, ht_list[i]); o.push_back('\n');
        return true;
    });
    if (fclose(fp)) rc = -1;
    return rc;
}

extern "C" int write_v_record_dat(const char *filename, const double *v_record, const int *td_list,
                                  const int *fd_list, const int *ht_list, int nrecords)
{
    return write_v_record_n(filename, v_record, td_list, fd_list, ht_list, nrecords, default_io_threads());
}

static int read_v_record_n(const char *filename, double **v_record, int **td_list, int **fd_list, int **ht_list,
                           int *nrecords, int threads)
{
    std::string text;
    if (!slurp(filename, text)) return -1;
    size_t pos = 0;
    int n;
    if (!parse_ints(text, &pos, &n, 1) || n < 0) return -1;
    *nrecords = n;
    double *v = (double *)calloc((size_t)n * 3 + 1, sizeof(double));
    int *td = (int *)calloc((size_t)n + 1, sizeof(int));
    int *fd = (int *)calloc((size_t)n + 1, sizeof(int));
    int *ht = (int *)calloc((size_t)n + 1, sizeof(int));
    /* per record: three "%lf" then three "%d" */
    size_t got = parse_tokens(text, pos, (size_t)n * 6, threads, [=](size_t k, const char *tok, char **end) {
        const size_t r = k / 6, f = k % 6;
        if (f < 3) { v[r * 3 + f] = fast_strtod(tok, end); return true; }
        return parse_int(tok, end, &(f == 3 ? td : f == 4 ? fd : ht)[r]);
    });
    if (got < (size_t)n * 6) {
        fprintf(stderr, "%s: short v_record file (%zu of %d records)\n", filename, got / 6, n);
        free(v); free(td); free(fd); free(ht);
        *v_record = NULL; *td_list = *fd_list = *ht_list = NULL;
        return -1;
    }
    *v_record = v; *td_list = td; *fd_list = fd; *ht_list = ht;
    return 0;
}

extern "C" int read_v_record_dat(const char *filename, double **v_record, int **td_list,
                                 int **fd_list, int **ht_list, int *nrecords)
{
    return read_v_record_n(filename, v_record, td_list, fd_list, ht_list, nrecords, default_io_threads());
}

extern "C" double *read_flux_dat(const char *filename, int *nface, int *tt)
{
    std::string text;
    if (!slurp(filename, text)) return NULL;
    size_t pos = 0;
    int h[2];
    if (!parse_ints(text, &pos, h, 2) || h[0] <= 0 || h[1] <= 0) return NULL;
    *nface = h[0];
    *tt = h[1];
    size_t total = (size_t)(*nface) * (*tt);
    double *F = (double *)calloc(total, sizeof(double));
    if (parse_doubles(text, pos, F, total, default_io_threads()) < total) { free(F); return NULL; }
    return F;
}

extern "C" int write_flux_dat(const char *filename, const double *F, int nface, int tt)
{
    FILE *fp = fopen(filename, "w");
    if (!fp) return -1;
    fprintf(fp, "%d\n%d\n", nface, tt);
    int rc = write_values(fp, 9, 7, F, (size_t)nface * tt, default_io_threads());
    if (fclose(fp)) rc = -1;
    return rc;
}

static int write_horizon_n(const char *filename, const double *horizon, int ftd, int ffd, int threads)
{
    FILE *fp = fopen(filename, "w");
    if (!fp) return -1;
    fprintf(fp, "%d\n%d\n", ftd, ffd);
    int rc = write_values(fp, 9, 6, horizon, (size_t)ftd * ffd, threads);
    if (fclose(fp)) rc = -1;
    return rc;
}

extern "C" int write_horizon_dat(const char *filename, const double *horizon, int ftd, int ffd)
{
    return write_horizon_n(filename, horizon, ftd, ffd, default_io_threads());
}

/* ------------------------------------------------------------------ drop-in drivers */
static uint64_t env_seed()
{
    const char *s = getenv("DSMC_SEED");
    return s ? strtoull(s, NULL, 10) : 0ull;
}

/* The drivers' nthreads (host threads in the reference, ballistic.c:97) sets
 * the threads of their file work; the device runs the particles. */
static int driver_threads(int nthreads) { return nthreads < 1 ? 1 : (nthreads > 64 ? 64 : nthreads); }

/* What a ballistic run hands to the file writers (and, in the pipeline, to
 * phase 2). */
struct BallisticOut {
    int ftd = 0, ffd = 0, hd = 0;
    int64_t nrec = 0;
    std::vector<double> den, vrec;
    std::vector<int32_t> vtd, vfd, vht;
};

/* run_ballistic from a temperature field in memory: the run, its messages and
 * its four files (ballistic.c:35-309).  Returns 0 or a dsmc_status. */
static int ballistic_core(int totaln, const double *T, int ftd, int ffd, const char *output_prefix, int io,
                          BallisticOut *keep)
{
    dsmc_params p;
    dsmc_params_preset(&p, DSMC_PRESET_COMPAT_BALLISTIC, ftd, ffd);
    p.nparticles = totaln < 0 ? 0 : totaln;
    p.seed = env_seed();
    dsmc_surface s;
    memset(&s, 0, sizeof s);
    s.ftd = ftd; s.ffd = ffd; s.T_latlon = T;
    const int hd = p.hd;
    size_t cap = (size_t)p.record_cap_total;
    std::vector<double> den((size_t)hd * ftd * ffd), hor((size_t)ftd * ffd), ttime((size_t)5 * p.nparticles);
    std::vector<double> vrec(cap * 3);
    std::vector<int32_t> vtd(cap), vfd(cap), vht(cap);
    dsmc_result r;
    memset(&r, 0, sizeof r);
    r.density = den.data(); r.horizon = hor.data(); r.ttime = ttime.data();
    r.vrec = vrec.data(); r.vrec_td = vtd.data(); r.vrec_fd = vfd.data(); r.vrec_ht = vht.data();
    printf("Starting ballistic simulation: %d particles\n", totaln);
    printf("  ftd=%d, ffd=%d, hd=%d\n", ftd, ffd, hd);
    printf("  Escape velocity = %.2f m/s\n", p.v_esc);
    printf("  Lifetime = %.1f s\n", p.lifetime);
    int rc = dsmc_run_ballistic(&p, &s, &r);
    if (rc) {
        fprintf(stderr, "dsmc_run_ballistic failed (%d)\n", rc);
        return rc;
    }
    printf("Simulation complete. Escaped: %d/%d\n", (int)r.n_escaped, totaln);
    std::string pre(output_prefix);
    write_density_n((pre + "_den.dat").c_str(), den.data(), ftd, ffd, hd, io);
    write_horizon_n((pre + "_horizon.dat").c_str(), hor.data(), ftd, ffd, io);
    write_v_record_n((pre + "_v_record.dat").c_str(), vrec.data(), vtd.data(), vfd.data(), vht.data(),
                     (int)r.vrec_n, io);
    write_ttime_n((pre + "_ttime.dat").c_str(), ttime.data(), 5, (int)p.nparticles, io);
    if (keep) {
        keep->ftd = ftd; keep->ffd = ffd; keep->hd = hd; keep->nrec = r.vrec_n;
        keep->den.swap(den);
        vrec.resize((size_t)r.vrec_n * 3); vtd.resize(r.vrec_n); vfd.resize(r.vrec_n); vht.resize(r.vrec_n);
        keep->vrec.swap(vrec); keep->vtd.swap(vtd); keep->vfd.swap(vfd); keep->vht.swap(vht);
    }
    return DSMC_OK;
}

extern "C" void run_ballistic(int totaln, const char *temp_file, const char *output_prefix, int nthreads)
{
    int ftd, ffd;
    double *T = read_temperature_dat(temp_file, &ftd, &ffd);
    if (!T) {
        fprintf(stderr, "Failed to read temperature file\n");
        return;
    }
    (void)ballistic_core(totaln, T, ftd, ffd, output_prefix, driver_threads(nthreads), nullptr);
    free(T);
}

/* run_advanced with its inputs in memory (advanced.c:119-476): the run, its
 * messages and its three files. */
static int advanced_core(int totaln, const double *T, int ftd, int ffd, const dsmc_phase2_input *in,
                         const char *output_prefix, int io)
{
    dsmc_params p;
    dsmc_params_preset(&p, DSMC_PRESET_COMPAT_ADVANCED, ftd, ffd);
    p.nparticles = totaln < 0 ? 0 : totaln;
    p.seed = env_seed();
    dsmc_surface s;
    memset(&s, 0, sizeof s);
    s.ftd = ftd; s.ffd = ffd; s.T_latlon = T;
    const int hd = p.hd;
    std::vector<double> den((size_t)hd * ftd * ffd), hor((size_t)ftd * ffd), ttime((size_t)9 * p.nparticles);
    dsmc_result r;
    memset(&r, 0, sizeof r);
    r.density = den.data(); r.horizon = hor.data(); r.ttime = ttime.data();
    printf("Building polynomial velocity fits...\n");
    int rc = dsmc_run_advanced(&p, &s, in, &r);
    if (rc) {
        fprintf(stderr, "dsmc_run_advanced failed (%d)\n", rc);
        return rc;
    }
    printf("Advanced DSMC complete. Escaped: %d/%d, Total collisions: %d\n", (int)r.n_escaped, totaln,
           (int)r.n_collisions);
    std::string pre(output_prefix);
    write_density_n((pre + "_den.dat").c_str(), den.data(), ftd, ffd, hd, io);
    write_horizon_n((pre + "_horizon.dat").c_str(), hor.data(), ftd, ffd, io);
    write_ttime_n((pre + "_ttime.dat").c_str(), ttime.data(), 9, (int)p.nparticles, io);
    return DSMC_OK;
}

extern "C" void run_advanced(int totaln, const char *temp_file, const char *prev_den_file,
                             const char *prev_v_record_file, const char *output_prefix, int nthreads)
{
    const int io = driver_threads(nthreads);
    int ftd, ffd;
    double *T = read_temperature_dat(temp_file, &ftd, &ffd);
    if (!T) {
        fprintf(stderr, "Failed to read temperature file\n");
        return;
    }
    dsmc_phase2_input in;
    memset(&in, 0, sizeof in);
    int pf = 0, pd = 0, ph = 0;
    double *atm = prev_den_file ? read_density_n(prev_den_file, &pf, &pd, &ph, 1, io) : NULL;
    if (atm) { in.copy_atm = atm; in.atm_ftd = pf; in.atm_ffd = pd; in.atm_hd = ph; }
    double *vr = NULL;
    int *vtd = NULL, *vfd = NULL, *vht = NULL, nrec = 0;
    if (prev_v_record_file && read_v_record_n(prev_v_record_file, &vr, &vtd, &vfd, &vht, &nrec, io) == 0) {
        printf("Read %d velocity records\n", nrec);
        in.vrec = vr; in.vrec_td = vtd; in.vrec_fd = vfd; in.vrec_ht = vht; in.nrec = nrec;
    }
    (void)advanced_core(totaln, T, ftd, ffd, &in, output_prefix, io);
    free(T); free(atm); free(vr); free(vtd); free(vfd); free(vht);
}

/* ------------------------------------------------------------------ pipeline */
/* v as the file holds it and a reader gets it back: the "%<prec>f" text
 * (fixed_fmt / snprintf, as the writers) parsed as the readers parse it. */
static double through_text(double v, int prec)
{
    std::string t;
    append_fixed(t, v, prec, 0);
    char *end;
    return fast_strtod(t.c_str(), &end);
}

extern "C" void grid_free(dsmc_sph_grid *g)
{
    if (!g) return;
    free(g->density);
    free(g->temperature);
    free(g->volume);
    free(g->horizon);
    free(g);
}

extern "C" dsmc_sph_grid *grid_create(int ftd, int ffd, int hd)
{
    if (ftd < 0 || ffd < 0 || hd < 0) return nullptr;
    if (hd > 0 && (size_t)ftd * ffd > SIZE_MAX / sizeof(double) / (size_t)hd) return nullptr;
    dsmc_sph_grid *g = (dsmc_sph_grid *)calloc(1, sizeof(dsmc_sph_grid));
    if (!g) return nullptr;
    g->ftd = ftd;
    g->ffd = ffd;
    g->hd = hd;
    const size_t surf = (size_t)ftd * ffd, cells = surf * hd;
    /* calloc(0) may return NULL: ask for at least one element */
    g->density = (double *)calloc(cells ? cells : 1, sizeof(double));
    g->temperature = (double *)calloc(surf ? surf : 1, sizeof(double));
    g->volume = (double *)calloc(cells ? cells : 1, sizeof(double));
    g->horizon = (double *)calloc(surf ? surf : 1, sizeof(double));
    if (!g->density || !g->temperature || !g->volume || !g->horizon) {
        grid_free(g);
        return nullptr;
    }
    return g;
}

extern "C" int dsmc_pipeline(int totaln, const char *output_prefix, int nthreads)
{
    if (!output_prefix) return DSMC_ERR_ARG;
    const int io = driver_threads(nthreads);
    const std::string pre(output_prefix);
    /* the thermal model on the reference CLI's 36 x 72 grid (main.c:166-187) */
    const int ftd = 36, ffd = 72;
    dsmc_thermal_params tp;
    thermal_params_default(&tp);
    std::vector<double> T((size_t)ftd * ffd);
    dsmc_sph_grid g;
    memset(&g, 0, sizeof g);
    g.ftd = ftd; g.ffd = ffd; g.hd = 1; g.temperature = T.data();
    thermal_solve(&g, &tp, nullptr, nthreads);
    const std::string tfile = pre + "_temperature.dat";
    FILE *fp = fopen(tfile.c_str(), "w");
    if (!fp) return DSMC_ERR_IO;
    fprintf(fp, "%d\n%d\n", ftd, ffd);
    int rc = write_values(fp, 8, 6, T.data(), T.size(), io);
    if (fclose(fp) || rc) return DSMC_ERR_IO;
    for (double &v : T) v = through_text(v, 6);
    /* phase 1 (main.c:189-193) */
    BallisticOut b;
    if ((rc = ballistic_core(totaln, T.data(), ftd, ffd, (pre + "_phase1").c_str(), io, &b))) return rc;
    /* phase 2 on phase 1's outputs as its files hold them (main.c:195-202) */
    std::vector<double> atm(b.den.size()), vr(b.vrec.size());
    for_chunks(atm.size(), io, 16384, [&](size_t, size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; i++) atm[i] = through_text(b.den[i], 6);
    });
    for_chunks(vr.size(), io, 16384, [&](size_t, size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; i++) vr[i] = through_text(b.vrec[i], 6);
    });
    dsmc_phase2_input in;
    memset(&in, 0, sizeof in);
    in.copy_atm = atm.data(); in.atm_ftd = b.ftd; in.atm_ffd = b.ffd; in.atm_hd = b.hd;
    printf("Read %d velocity records\n", (int)b.nrec);
    in.vrec = vr.data(); in.vrec_td = b.vtd.data(); in.vrec_fd = b.vfd.data(); in.vrec_ht = b.vht.data();
    in.nrec = b.nrec;
    return advanced_core(totaln, T.data(), ftd, ffd, &in, (pre + "_phase2").c_str(), io);
}
