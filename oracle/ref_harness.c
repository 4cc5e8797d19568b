/*
 * ref_harness.c - drives the COMPILED REFERENCE (src/common/*.c, src/dsmc/*.c of
 * dsmc_suite, built from their sources by oracle/Makefile into oracle/_ref/) to
 * produce golden fixtures.  Runs only where the reference sources exist; its
 * outputs are small binary files that tests/golden/make_golden.py turns into
 * committed .npz fixtures.  Test infrastructure, never shipped.
 *
 * Link-time wrapping (-Wl,--wrap=...) intercepts, without touching the
 * reference sources:
 *   rand_r            - replaced by the Philox draw stream (philox.h) keyed by
 *                       the particle id the harness is processing, or passed
 *                       through to glibc (mode "randr");
 *   verlet_step       - counted, last state kept, optional per-step trace;
 *   write_*_dat       - full-precision arrays captured before text rounding
 *                       (the text files are still written by __real_*).
 *
 * run_ballistic/run_advanced are called with totaln = 1 once per particle so
 * that every particle's draws come from its own keyed stream (the reference's
 * seeds are per thread, not per particle).
 *
 * usage:
 *   ref_harness ballistic_keyed T.dat N seed out.bin [trace_pid trace_steps]
 *   ref_harness advanced_keyed  T.dat den.dat vrec.dat N seed out.bin
 *   ref_harness ballistic_seq   T.dat N seed out.bin     (one call, sequential stream)
 *   ref_harness ballistic_randr T.dat N nthreads prefix  (unwrapped draws, text files)
 *   ref_harness advanced_randr  T.dat den vrec N nthreads prefix
 *   ref_harness vtable out.bin
 *   ref_harness polyfit in.bin out.bin                   (x,y sets -> coeffs)
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#include <math.h>

#include "types.h"
#include "constants.h"
#include "integrator.h"
#include "velocity.h"
#include "polyfit.h"
#include "grid.h"
#include "thermal.h"
#include "../dsmc_suite_amd/csrc/philox.h"

void run_ballistic(int totaln, const char *temp_file, const char *output_prefix, int nthreads);
void run_advanced(int totaln, const char *temp_file, const char *prev_den_file,
                  const char *prev_v_record_file, const char *output_prefix, int nthreads);

/* ------------------------------------------------------------- wrap state */
enum { M_PASS = 0, M_KEYED = 1 };
static int g_mode = M_PASS;
static uint64_t g_key, g_pid, g_k;
static uint32_t g_stream;
static int64_t g_verlet_calls;
static Particle g_last;
static int64_t g_trace_left;
static FILE *g_trace;
static int g_capture_files = 0;

static double *cap_den; static int cap_den_n, cap_hd, cap_ftd, cap_ffd;
static double *cap_hor; static int cap_hor_n;
static double *cap_tt; static int cap_tt_rows, cap_tt_n;
static double *cap_vr; static int *cap_vtd, *cap_vfd, *cap_vht; static int cap_vn;

int __real_rand_r(unsigned int *seed);
int __wrap_rand_r(unsigned int *seed)
{
    if (g_mode == M_PASS) return __real_rand_r(seed);
    return (int)dsmc_draw31(g_key, g_pid, g_stream, g_k++);
}

VerletStatus __real_verlet_step(Particle *p, const PhysicsParams *params, double h, Vec3 *oa,
                                double *ds_out);
VerletStatus __wrap_verlet_step(Particle *p, const PhysicsParams *params, double h, Vec3 *oa,
                                double *ds_out)
{
    VerletStatus st = __real_verlet_step(p, params, h, oa, ds_out);
    g_verlet_calls++;
    g_last = *p;
    if (g_trace && g_trace_left > 0) {
        double rec[7] = {p->pos.x, p->pos.y, p->pos.z, p->vel.x, p->vel.y, p->vel.z, (double)st};
        fwrite(rec, sizeof(double), 7, g_trace);
        g_trace_left--;
    }
    return st;
}

int __real_write_density_dat(const char *f, const double *den, int ftd, int ffd, int hd);
int __wrap_write_density_dat(const char *f, const double *den, int ftd, int ffd, int hd)
{
    int n = hd * ftd * ffd;
    free(cap_den);
    cap_den = malloc(sizeof(double) * n);
    memcpy(cap_den, den, sizeof(double) * n);
    cap_den_n = n; cap_hd = hd; cap_ftd = ftd; cap_ffd = ffd;
    return g_capture_files ? __real_write_density_dat(f, den, ftd, ffd, hd) : 0;
}
int __real_write_horizon_dat(const char *f, const double *h, int ftd, int ffd);
int __wrap_write_horizon_dat(const char *f, const double *h, int ftd, int ffd)
{
    free(cap_hor);
    cap_hor = malloc(sizeof(double) * ftd * ffd);
    memcpy(cap_hor, h, sizeof(double) * ftd * ffd);
    cap_hor_n = ftd * ffd;
    return g_capture_files ? __real_write_horizon_dat(f, h, ftd, ffd) : 0;
}
int __real_write_ttime_dat(const char *f, const double *tt, int nrows, int totaln);
int __wrap_write_ttime_dat(const char *f, const double *tt, int nrows, int totaln)
{
    free(cap_tt);
    cap_tt = malloc(sizeof(double) * nrows * totaln);
    memcpy(cap_tt, tt, sizeof(double) * nrows * totaln);
    cap_tt_rows = nrows; cap_tt_n = totaln;
    return g_capture_files ? __real_write_ttime_dat(f, tt, nrows, totaln) : 0;
}
int __real_write_v_record_dat(const char *f, const double *vr, const int *td, const int *fd,
                              const int *ht, int n);
int __wrap_write_v_record_dat(const char *f, const double *vr, const int *td, const int *fd,
                              const int *ht, int n)
{
    free(cap_vr); free(cap_vtd); free(cap_vfd); free(cap_vht);
    cap_vr = malloc(sizeof(double) * 3 * (n + 1));
    cap_vtd = malloc(sizeof(int) * (n + 1));
    cap_vfd = malloc(sizeof(int) * (n + 1));
    cap_vht = malloc(sizeof(int) * (n + 1));
    memcpy(cap_vr, vr, sizeof(double) * 3 * n);
    memcpy(cap_vtd, td, sizeof(int) * n);
    memcpy(cap_vfd, fd, sizeof(int) * n);
    memcpy(cap_vht, ht, sizeof(int) * n);
    cap_vn = n;
    return g_capture_files ? __real_write_v_record_dat(f, vr, td, fd, ht, n) : 0;
}

/* ------------------------------------------------------------- output */
static void put_i64(FILE *f, int64_t v) { fwrite(&v, 8, 1, f); }
static void put_arr(FILE *f, const char *name, const void *data, int64_t count, int elsize,
                    char kind)
{
    char nm[32] = {0};
    strncpy(nm, name, 31);
    fwrite(nm, 1, 32, f);
    fputc(kind, f);
    put_i64(f, elsize);
    put_i64(f, count);
    if (count) fwrite(data, elsize, count, f);
}

static void quiet(void) { if (!freopen("/dev/null", "w", stdout)) perror("freopen"); }

/* ------------------------------------------------------------- modes */
static int keyed_ballistic(const char *T, int N, uint64_t seed, const char *out, int trace_pid,
                           int trace_steps, int advanced, const char *den, const char *vrec)
{
    quiet();
    int rows = advanced ? 9 : 5;
    double *ttime = calloc((size_t)rows * N, sizeof(double));
    int64_t *steps = calloc(N, sizeof(int64_t)), *draws = calloc(N, sizeof(int64_t));
    double *endst = calloc((size_t)N * 6, sizeof(double));
    double *den_sum = NULL, *hor_sum = NULL;
    int den_n = 0, hor_n = 0;
    /* records: concatenation over particles of each single-particle run */
    int64_t vcap = 1 << 20, vn = 0;
    double *vr = malloc(sizeof(double) * 3 * vcap);
    int *vtd = malloc(sizeof(int) * vcap), *vfd = malloc(sizeof(int) * vcap),
        *vht = malloc(sizeof(int) * vcap), *vpid = malloc(sizeof(int) * vcap);
    char tracefile[1024];
    snprintf(tracefile, sizeof tracefile, "%s.trace", out);
    g_mode = M_KEYED;
    g_key = seed;
    g_stream = advanced ? 2 : 1;
    for (int n = 0; n < N; n++) {
        g_pid = (uint64_t)n;
        g_k = 0;
        g_verlet_calls = 0;
        memset(&g_last, 0, sizeof g_last);
        if (n == trace_pid) { g_trace = fopen(tracefile, "wb"); g_trace_left = trace_steps; }
        if (advanced) run_advanced(1, T, den, vrec, "/tmp/ref_harness_adv", 1);
        else run_ballistic(1, T, "/tmp/ref_harness_bal", 1);
        if (g_trace) { fclose(g_trace); g_trace = NULL; }
        for (int r = 0; r < rows; r++) ttime[(size_t)r * N + n] = cap_tt[r];
        steps[n] = g_verlet_calls;
        draws[n] = (int64_t)g_k;
        endst[n * 6 + 0] = g_last.pos.x; endst[n * 6 + 1] = g_last.pos.y;
        endst[n * 6 + 2] = g_last.pos.z; endst[n * 6 + 3] = g_last.vel.x;
        endst[n * 6 + 4] = g_last.vel.y; endst[n * 6 + 5] = g_last.vel.z;
        if (!den_sum) { den_n = cap_den_n; den_sum = calloc(den_n, sizeof(double)); }
        if (!hor_sum) { hor_n = cap_hor_n; hor_sum = calloc(hor_n, sizeof(double)); }
        for (int i = 0; i < den_n; i++) den_sum[i] += cap_den[i];
        for (int i = 0; i < hor_n; i++) hor_sum[i] += cap_hor[i];
        if (!advanced) {
            for (int i = 0; i < cap_vn && vn < vcap; i++, vn++) {
                vr[vn * 3] = cap_vr[i * 3]; vr[vn * 3 + 1] = cap_vr[i * 3 + 1];
                vr[vn * 3 + 2] = cap_vr[i * 3 + 2];
                vtd[vn] = cap_vtd[i]; vfd[vn] = cap_vfd[i]; vht[vn] = cap_vht[i]; vpid[vn] = n;
            }
        }
    }
    FILE *f = fopen(out, "wb");
    if (!f) return 2;
    int dims[4] = {cap_hd, cap_ftd, cap_ffd, N};
    put_arr(f, "dims", dims, 4, 4, 'i');
    put_arr(f, "ttime", ttime, (int64_t)rows * N, 8, 'f');
    put_arr(f, "steps", steps, N, 8, 'i');
    put_arr(f, "draws", draws, N, 8, 'i');
    put_arr(f, "end_state", endst, (int64_t)N * 6, 8, 'f');
    put_arr(f, "density", den_sum, den_n, 8, 'f');
    put_arr(f, "horizon", hor_sum, hor_n, 8, 'f');
    put_arr(f, "vrec", vr, vn * 3, 8, 'f');
    put_arr(f, "vrec_td", vtd, vn, 4, 'i');
    put_arr(f, "vrec_fd", vfd, vn, 4, 'i');
    put_arr(f, "vrec_ht", vht, vn, 4, 'i');
    put_arr(f, "vrec_pid", vpid, vn, 4, 'i');
    fclose(f);
    return 0;
}

static int seq_ballistic(const char *T, int N, uint64_t seed, const char *out)
{
    quiet();
    g_mode = M_KEYED;
    g_key = seed; g_pid = 0; g_k = 0; g_stream = 1;
    g_verlet_calls = 0;
    run_ballistic(N, T, "/tmp/ref_harness_seq", 1);
    FILE *f = fopen(out, "wb");
    if (!f) return 2;
    int dims[4] = {cap_hd, cap_ftd, cap_ffd, N};
    put_arr(f, "dims", dims, 4, 4, 'i');
    put_arr(f, "ttime", cap_tt, (int64_t)cap_tt_rows * cap_tt_n, 8, 'f');
    put_arr(f, "density", cap_den, cap_den_n, 8, 'f');
    put_arr(f, "horizon", cap_hor, cap_hor_n, 8, 'f');
    put_arr(f, "vrec", cap_vr, (int64_t)cap_vn * 3, 8, 'f');
    put_arr(f, "vrec_td", cap_vtd, cap_vn, 4, 'i');
    put_arr(f, "vrec_fd", cap_vfd, cap_vn, 4, 'i');
    put_arr(f, "vrec_ht", cap_vht, cap_vn, 4, 'i');
    int64_t tot[2] = {g_verlet_calls, (int64_t)g_k};
    put_arr(f, "totals", tot, 2, 8, 'i');
    fclose(f);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 2) { fprintf(stderr, "usage: see source header\n"); return 1; }
    const char *m = argv[1];
    if (!strcmp(m, "ballistic_keyed") && argc >= 6) {
        int tp = argc >= 8 ? atoi(argv[6]) : -1, ts = argc >= 8 ? atoi(argv[7]) : 0;
        return keyed_ballistic(argv[2], atoi(argv[3]), strtoull(argv[4], 0, 10), argv[5], tp, ts,
                               0, NULL, NULL);
    }
    if (!strcmp(m, "advanced_keyed") && argc >= 8)
        return keyed_ballistic(argv[2], atoi(argv[5]), strtoull(argv[6], 0, 10), argv[7], -1, 0, 1,
                               argv[3], argv[4]);
    if (!strcmp(m, "ballistic_seq") && argc >= 6)
        return seq_ballistic(argv[2], atoi(argv[3]), strtoull(argv[4], 0, 10), argv[5]);
    if (!strcmp(m, "ballistic_randr") && argc >= 6) {
        g_mode = M_PASS; g_capture_files = 1; quiet();
        run_ballistic(atoi(argv[3]), argv[2], argv[5], atoi(argv[4]));
        return 0;
    }
    if (!strcmp(m, "advanced_randr") && argc >= 8) {
        g_mode = M_PASS; g_capture_files = 1; quiet();
        run_advanced(atoi(argv[5]), argv[2], argv[3], argv[4], argv[7], atoi(argv[6]));
        return 0;
    }
    if (!strcmp(m, "thermal") && argc >= 3) {
        /* the T field of `dsmc_suite thermal` (main.c:118-144), written the same way */
        int ftd = 36, ffd = 72;
        SphericalGrid *grid = grid_create(ftd, ffd, 1);
        ThermalParams tp;
        thermal_params_default(&tp);
        quiet();
        thermal_solve(grid, &tp, NULL, 1);
        FILE *fp = fopen(argv[2], "w");
        if (!fp) return 2;
        fprintf(fp, "%d\n", ftd);
        fprintf(fp, "%d\n", ffd);
        for (int td = 0; td < ftd; td++)
            for (int fd = 0; fd < ffd; fd++) fprintf(fp, "%8.6f\n", grid->temperature[td * ffd + fd]);
        fclose(fp);
        return 0;
    }
    if (!strcmp(m, "vtable") && argc >= 3) {
        static VelocityTable vt;
        velocity_table_init(&vt, CONST_MMASS_H2O, CONST_MOLE, CONST_K_B);
        FILE *f = fopen(argv[2], "wb");
        if (!f) return 2;
        put_arr(f, "temps", vt.temps, VEL_TABLE_NTEMP, 8, 'f');
        put_arr(f, "speeds", vt.speeds, VEL_TABLE_NTEMP * VEL_TABLE_NVEL, 8, 'f');
        put_arr(f, "count", vt.count, VEL_TABLE_NTEMP, 4, 'i');
        /* test_verlet.c:76-86 known answer: mean of 10000 draws at 200 K, seed 42 */
        unsigned int seed = 42;
        double sum = 0;
        for (int i = 0; i < 10000; i++) sum += velocity_sample(&vt, 200.0, &seed);
        double mean = sum / 10000;
        put_arr(f, "mean200", &mean, 1, 8, 'f');
        fclose(f);
        return 0;
    }
    if (!strcmp(m, "polyfit") && argc >= 4) {
        /* in: int64 nsets, then per set: int64 n, int64 degree, x[n], y[n] */
        FILE *fi = fopen(argv[2], "rb"), *fo = fopen(argv[3], "wb");
        if (!fi || !fo) return 2;
        int64_t ns;
        if (fread(&ns, 8, 1, fi) != 1) return 3;
        for (int64_t s = 0; s < ns; s++) {
            int64_t n, d;
            if (fread(&n, 8, 1, fi) != 1 || fread(&d, 8, 1, fi) != 1) return 3;
            double *x = malloc(8 * n), *y = malloc(8 * n), c[8] = {0};
            if (fread(x, 8, n, fi) != (size_t)n || fread(y, 8, n, fi) != (size_t)n) return 3;
            int rc = polyfit(x, y, (int)n, (int)d, c);
            int64_t rc64 = rc;
            fwrite(&rc64, 8, 1, fo);
            fwrite(c, 8, 7, fo);
            free(x); free(y);
        }
        fclose(fi); fclose(fo);
        return 0;
    }
    fprintf(stderr, "bad arguments\n");
    return 1;
}
