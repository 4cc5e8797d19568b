/*
 * dsmc_cli.c - a command-line driver for the GPU tests, linked against
 * libdsmc_hip.so alone.  It makes the calls the dsmc_suite CLI makes
 * (src/main.c:104-210) with the same commands and flags, so the tests can run
 * every drop-in entry point the way that CLI drives it without the reference's
 * own code on the GPU machine:
 *
 *   dsmc      -n -T -o -t              run_ballistic            (main.c:146-147)
 *   dsmc-adv  -n -T -d -v -o -t        run_advanced             (main.c:149-151)
 *   thermal   --za -o -t               grid_create, thermal_params_default,
 *                                      thermal_solve, the T file (main.c:117-144)
 *   raytrace  -m -f --npoint --nface --za --tt --nmonte -o -t
 *                                      run_raytracer            (main.c:104-116)
 *   pipeline  -n -o -t [mesh flags]    the four steps of main.c:153-204
 *
 * The two drivers are declared here exactly as main.c:18-23 declares them (the
 * reference ships no header for them); everything else comes from
 * include/dsmc_hip.h.  Test infrastructure: not part of the library.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#include "../../include/dsmc_hip.h"

extern void run_ballistic(int totaln, const char *temp_file, const char *output_prefix, int nthreads);
extern void run_advanced(int totaln, const char *temp_file, const char *prev_den_file,
                         const char *prev_v_record_file, const char *output_prefix, int nthreads);

typedef struct {
    int totaln, nthreads, npoint, nface, tt, nmonte;
    double za_deg;
    const char *temp_file, *den_file, *vrec_file, *prefix, *vert_file, *face_file;
} cli_opts;

/* Flags with one value each; an unknown flag or a flag without its value is
 * skipped, as the reference's parser skips it. */
static void parse(int argc, char **argv, cli_opts *o)
{
    for (int i = 2; i + 1 < argc; i++) {
        const char *f = argv[i], *v = argv[i + 1];
        int used = 1;
        if (!strcmp(f, "-n")) o->totaln = atoi(v);
        else if (!strcmp(f, "-t")) o->nthreads = atoi(v);
        else if (!strcmp(f, "-T")) o->temp_file = v;
        else if (!strcmp(f, "-d")) o->den_file = v;
        else if (!strcmp(f, "-v")) o->vrec_file = v;
        else if (!strcmp(f, "-o")) o->prefix = v;
        else if (!strcmp(f, "-m")) o->vert_file = v;
        else if (!strcmp(f, "-f")) o->face_file = v;
        else if (!strcmp(f, "--npoint")) o->npoint = atoi(v);
        else if (!strcmp(f, "--nface")) o->nface = atoi(v);
        else if (!strcmp(f, "--za")) o->za_deg = atof(v);
        else if (!strcmp(f, "--tt")) o->tt = atoi(v);
        else if (!strcmp(f, "--nmonte")) o->nmonte = atoi(v);
        else used = 0;
        i += used;
    }
}

/* The 36 x 72 thermal model as the CLI runs it, T written "%8.6f" one per line
 * after the two dimensions (main.c:119-142); za_deg NULL keeps the default
 * obliquity, as the pipeline's thermal step does (main.c:168-171).  0 or -1. */
static int thermal_file(const char *path, const double *za_deg, int nthreads)
{
    const int ftd = 36, ffd = 72;
    dsmc_sph_grid *g = grid_create(ftd, ffd, 1);
    if (!g) return -1;
    dsmc_thermal_params tp;
    thermal_params_default(&tp);
    if (za_deg) tp.z_angle = *za_deg / 180.0 * M_PI;
    thermal_solve(g, &tp, NULL, nthreads);
    FILE *fp = fopen(path, "w");
    if (fp) {
        fprintf(fp, "%d\n%d\n", ftd, ffd);
        for (int i = 0; i < ftd * ffd; i++) fprintf(fp, "%8.6f\n", g->temperature[i]);
        fclose(fp);
    }
    grid_free(g);
    return fp ? 0 : -1;
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: %s dsmc|dsmc-adv|thermal|raytrace|pipeline [flags]\n", argv[0]);
        return 1;
    }
    cli_opts o = {10000, 4, 0, 0, 24, 50, 12.3, "Ceres_T_VP.dat", NULL, NULL, "output", NULL, NULL};
    parse(argc, argv, &o);
    const char *cmd = argv[1];
    char a[512], b[512], c[512], d[512];
    const int have_mesh = o.vert_file && o.face_file && o.npoint > 0 && o.nface > 0;
    printf("Using %d OpenMP threads\n", o.nthreads);
    if (!strcmp(cmd, "dsmc")) {
        run_ballistic(o.totaln, o.temp_file, o.prefix, o.nthreads);
    } else if (!strcmp(cmd, "dsmc-adv")) {
        run_advanced(o.totaln, o.temp_file, o.den_file, o.vrec_file, o.prefix, o.nthreads);
    } else if (!strcmp(cmd, "thermal")) {
        snprintf(a, sizeof a, "%s_temperature.dat", o.prefix);
        if (thermal_file(a, &o.za_deg, o.nthreads)) return 1;
        printf("Temperature written to %s\n", a);
    } else if (!strcmp(cmd, "raytrace")) {
        if (!have_mesh) {
            fprintf(stderr, "raytrace requires: -m <vertex> -f <face> --npoint <n> --nface <n>\n");
            return 1;
        }
        snprintf(a, sizeof a, "%s_flux.dat", o.prefix);
        run_raytracer(o.vert_file, o.face_file, o.npoint, o.nface, o.za_deg, o.tt, o.nmonte, a, o.nthreads);
    } else if (!strcmp(cmd, "pipeline")) {
        printf("=== FULL PIPELINE ===\n\n");
        if (have_mesh) {
            printf("--- Step 1: Ray Tracing ---\n");
            snprintf(a, sizeof a, "%s_flux.dat", o.prefix);
            run_raytracer(o.vert_file, o.face_file, o.npoint, o.nface, o.za_deg, o.tt, o.nmonte, a, o.nthreads);
        } else {
            printf("--- Step 1: Skipped (no mesh provided) ---\n");
        }
        printf("\n--- Step 2: Thermal Model ---\n");
        snprintf(a, sizeof a, "%s_temperature.dat", o.prefix);
        thermal_file(a, NULL, o.nthreads);
        printf("\n--- Step 3: Basic DSMC ---\n");
        snprintf(b, sizeof b, "%s_phase1", o.prefix);
        run_ballistic(o.totaln, a, b, o.nthreads);
        printf("\n--- Step 4: Advanced DSMC ---\n");
        snprintf(c, sizeof c, "%s_den.dat", b);
        snprintf(d, sizeof d, "%s_v_record.dat", b);
        char adv[512];
        snprintf(adv, sizeof adv, "%s_phase2", o.prefix);
        run_advanced(o.totaln, a, c, d, adv, o.nthreads);
        printf("\n=== PIPELINE COMPLETE ===\n");
    } else {
        fprintf(stderr, "Unknown command: %s\n", cmd);
        return 1;
    }
    return 0;
}
