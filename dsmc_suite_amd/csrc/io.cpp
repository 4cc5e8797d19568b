/*
 * io.cpp - the .dat formats of dsmc_suite (src/common/fileio.c, include/fileio.h:8-44)
 * with the same names, signatures and printf formats, and the drop-in drivers
 * run_ballistic / run_advanced (ballistic.c:35-309, advanced.c:119-476) built on
 * the device path.
 *
 * Readers differ from the reference in one respect: a missing, short or
 * malformed file is reported on stderr and returns NULL / -1, where the
 * reference leaves zeros behind fscanf failures (fileio.c:11-21).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <vector>
#include <string>

#include "../../include/dsmc_hip.h"

static int scan_int(FILE *fp, int *v) { return fscanf(fp, "%d", v) == 1; }

extern "C" double *read_density_dat(const char *filename, int *ftd, int *ffd, int *hd, int has_hd)
{
    FILE *fp = fopen(filename, "r");
    if (!fp) { fprintf(stderr, "Cannot open %s\n", filename); return NULL; }
    if (!scan_int(fp, ftd) || !scan_int(fp, ffd) || (has_hd && !scan_int(fp, hd)) || *ftd <= 0 ||
        *ffd <= 0 || *hd <= 0) {
        fprintf(stderr, "%s: malformed density header\n", filename);
        fclose(fp);
        return NULL;
    }
    size_t total = (size_t)(*hd) * (*ftd) * (*ffd);
    double *data = (double *)calloc(total, sizeof(double));
    for (size_t i = 0; i < total; i++)
        if (fscanf(fp, "%lf", &data[i]) != 1) {
            fprintf(stderr, "%s: short density file (%zu of %zu values)\n", filename, i, total);
            free(data);
            fclose(fp);
            return NULL;
        }
    fclose(fp);
    return data;
}

/* Formats n doubles "%fmt\n" into one buffer and writes it (same bytes as a
 * per-value fprintf, fewer library calls). */
static int write_values(FILE *fp, const char *fmt, const double *v, size_t n)
{
    std::string buf;
    buf.reserve(n * 12 + 16);
    char tmp[64];
    for (size_t i = 0; i < n; i++) {
        int k = snprintf(tmp, sizeof tmp, fmt, v[i]);
        if (k < 0) return -1;
        if ((size_t)k < sizeof tmp) buf.append(tmp, k);
        else {
            std::vector<char> big(k + 1);
            snprintf(big.data(), big.size(), fmt, v[i]);
            buf.append(big.data(), k);
        }
    }
    return fwrite(buf.data(), 1, buf.size(), fp) == buf.size() ? 0 : -1;
}

extern "C" int write_density_dat(const char *filename, const double *den, int ftd, int ffd, int hd)
{
    FILE *fp = fopen(filename, "w");
    if (!fp) return -1;
    fprintf(fp, "%d\n%d\n%d\n", ftd, ffd, hd);
    int rc = write_values(fp, "%9.6f\n", den, (size_t)hd * ftd * ffd);
    fclose(fp);
    return rc;
}

extern "C" double *read_temperature_dat(const char *filename, int *ftd, int *ffd)
{
    FILE *fp = fopen(filename, "r");
    if (!fp) { fprintf(stderr, "Cannot open %s\n", filename); return NULL; }
    if (!scan_int(fp, ftd) || !scan_int(fp, ffd) || *ftd <= 0 || *ffd <= 0) {
        fprintf(stderr, "%s: malformed temperature header\n", filename);
        fclose(fp);
        return NULL;
    }
    size_t total = (size_t)(*ftd) * (*ffd);
    double *data = (double *)calloc(total, sizeof(double));
    for (size_t i = 0; i < total; i++)
        if (fscanf(fp, "%lf", &data[i]) != 1) {
            fprintf(stderr, "%s: short temperature file (%zu of %zu values)\n", filename, i, total);
            free(data);
            fclose(fp);
            return NULL;
        }
    fclose(fp);
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

extern "C" int write_ttime_dat(const char *filename, const double *ttime, int nrows, int totaln)
{
    FILE *fp = fopen(filename, "w");
    if (!fp) return -1;
    int rc = write_values(fp, "%9.5f\n", ttime, (size_t)nrows * totaln);
    fclose(fp);
    return rc;
}

extern "C" int write_v_record_dat(const char *filename, const double *v_record, const int *td_list,
                                  const int *fd_list, const int *ht_list, int nrecords)
{
    FILE *fp = fopen(filename, "w");
    if (!fp) return -1;
    std::string buf;
    buf.reserve((size_t)nrecords * 48 + 16);
    char tmp[512];
    snprintf(tmp, sizeof tmp, "%d\n", nrecords);
    buf += tmp;
    for (int i = 0; i < nrecords; i++) {
        int k = snprintf(tmp, sizeof tmp, "%f %f %f\n%d\n%d\n%d\n", v_record[i * 3], v_record[i * 3 + 1],
                         v_record[i * 3 + 2], td_list[i], fd_list[i], ht_list[i]);
        buf.append(tmp, k < (int)sizeof tmp ? k : (int)sizeof tmp - 1);
    }
    int rc = fwrite(buf.data(), 1, buf.size(), fp) == buf.size() ? 0 : -1;
    fclose(fp);
    return rc;
}

extern "C" int read_v_record_dat(const char *filename, double **v_record, int **td_list,
                                 int **fd_list, int **ht_list, int *nrecords)
{
    FILE *fp = fopen(filename, "r");
    if (!fp) return -1;
    int n;
    if (!scan_int(fp, &n) || n < 0) { fclose(fp); return -1; }
    *nrecords = n;
    *v_record = (double *)calloc((size_t)n * 3 + 1, sizeof(double));
    *td_list = (int *)calloc((size_t)n + 1, sizeof(int));
    *fd_list = (int *)calloc((size_t)n + 1, sizeof(int));
    *ht_list = (int *)calloc((size_t)n + 1, sizeof(int));
    for (int i = 0; i < n; i++) {
        if (fscanf(fp, "%lf %lf %lf", &(*v_record)[i * 3], &(*v_record)[i * 3 + 1],
                   &(*v_record)[i * 3 + 2]) != 3 ||
            !scan_int(fp, &(*td_list)[i]) || !scan_int(fp, &(*fd_list)[i]) || !scan_int(fp, &(*ht_list)[i])) {
            fprintf(stderr, "%s: short v_record file (%d of %d records)\n", filename, i, n);
            free(*v_record); free(*td_list); free(*fd_list); free(*ht_list);
            *v_record = NULL; *td_list = *fd_list = *ht_list = NULL;
            fclose(fp);
            return -1;
        }
    }
    fclose(fp);
    return 0;
}

extern "C" double *read_flux_dat(const char *filename, int *nface, int *tt)
{
    FILE *fp = fopen(filename, "r");
    if (!fp) return NULL;
    if (!scan_int(fp, nface) || !scan_int(fp, tt) || *nface <= 0 || *tt <= 0) { fclose(fp); return NULL; }
    size_t total = (size_t)(*nface) * (*tt);
    double *F = (double *)calloc(total, sizeof(double));
    for (size_t i = 0; i < total; i++)
        if (fscanf(fp, "%lf", &F[i]) != 1) { free(F); fclose(fp); return NULL; }
    fclose(fp);
    return F;
}

extern "C" int write_flux_dat(const char *filename, const double *F, int nface, int tt)
{
    FILE *fp = fopen(filename, "w");
    if (!fp) return -1;
    fprintf(fp, "%d\n%d\n", nface, tt);
    int rc = write_values(fp, "%9.7f\n", F, (size_t)nface * tt);
    fclose(fp);
    return rc;
}

extern "C" int write_horizon_dat(const char *filename, const double *horizon, int ftd, int ffd)
{
    FILE *fp = fopen(filename, "w");
    if (!fp) return -1;
    fprintf(fp, "%d\n%d\n", ftd, ffd);
    int rc = write_values(fp, "%9.6f\n", horizon, (size_t)ftd * ffd);
    fclose(fp);
    return rc;
}

/* ------------------------------------------------------------------ drop-in drivers */
static uint64_t env_seed()
{
    const char *s = getenv("DSMC_SEED");
    return s ? strtoull(s, NULL, 10) : 0ull;
}

extern "C" void run_ballistic(int totaln, const char *temp_file, const char *output_prefix, int nthreads)
{
    (void)nthreads;
    int ftd, ffd;
    double *T = read_temperature_dat(temp_file, &ftd, &ffd);
    if (!T) {
        fprintf(stderr, "Failed to read temperature file\n");
        return;
    }
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
    free(T);
    if (rc) {
        fprintf(stderr, "dsmc_run_ballistic failed (%d)\n", rc);
        return;
    }
    printf("Simulation complete. Escaped: %d/%d\n", (int)r.n_escaped, totaln);
    std::string pre(output_prefix);
    write_density_dat((pre + "_den.dat").c_str(), den.data(), ftd, ffd, hd);
    write_horizon_dat((pre + "_horizon.dat").c_str(), hor.data(), ftd, ffd);
    write_v_record_dat((pre + "_v_record.dat").c_str(), vrec.data(), vtd.data(), vfd.data(), vht.data(),
                       (int)r.vrec_n);
    write_ttime_dat((pre + "_ttime.dat").c_str(), ttime.data(), 5, (int)p.nparticles);
}

extern "C" void run_advanced(int totaln, const char *temp_file, const char *prev_den_file,
                             const char *prev_v_record_file, const char *output_prefix, int nthreads)
{
    (void)nthreads;
    int ftd, ffd;
    double *T = read_temperature_dat(temp_file, &ftd, &ffd);
    if (!T) {
        fprintf(stderr, "Failed to read temperature file\n");
        return;
    }
    dsmc_params p;
    dsmc_params_preset(&p, DSMC_PRESET_COMPAT_ADVANCED, ftd, ffd);
    p.nparticles = totaln < 0 ? 0 : totaln;
    p.seed = env_seed();
    dsmc_surface s;
    memset(&s, 0, sizeof s);
    s.ftd = ftd; s.ffd = ffd; s.T_latlon = T;
    dsmc_phase2_input in;
    memset(&in, 0, sizeof in);
    int pf = 0, pd = 0, ph = 0;
    double *atm = prev_den_file ? read_density_dat(prev_den_file, &pf, &pd, &ph, 1) : NULL;
    if (atm) { in.copy_atm = atm; in.atm_ftd = pf; in.atm_ffd = pd; in.atm_hd = ph; }
    double *vr = NULL;
    int *vtd = NULL, *vfd = NULL, *vht = NULL, nrec = 0;
    if (prev_v_record_file && read_v_record_dat(prev_v_record_file, &vr, &vtd, &vfd, &vht, &nrec) == 0) {
        printf("Read %d velocity records\n", nrec);
        in.vrec = vr; in.vrec_td = vtd; in.vrec_fd = vfd; in.vrec_ht = vht; in.nrec = nrec;
    }
    const int hd = p.hd;
    std::vector<double> den((size_t)hd * ftd * ffd), hor((size_t)ftd * ffd), ttime((size_t)9 * p.nparticles);
    dsmc_result r;
    memset(&r, 0, sizeof r);
    r.density = den.data(); r.horizon = hor.data(); r.ttime = ttime.data();
    printf("Building polynomial velocity fits...\n");
    int rc = dsmc_run_advanced(&p, &s, &in, &r);
    free(T); free(atm); free(vr); free(vtd); free(vfd); free(vht);
    if (rc) {
        fprintf(stderr, "dsmc_run_advanced failed (%d)\n", rc);
        return;
    }
    printf("Advanced DSMC complete. Escaped: %d/%d, Total collisions: %d\n", (int)r.n_escaped, totaln,
           (int)r.n_collisions);
    std::string pre(output_prefix);
    write_density_dat((pre + "_den.dat").c_str(), den.data(), ftd, ffd, hd);
    write_horizon_dat((pre + "_horizon.dat").c_str(), hor.data(), ftd, ffd);
    write_ttime_dat((pre + "_ttime.dat").c_str(), ttime.data(), 9, (int)p.nparticles);
}
