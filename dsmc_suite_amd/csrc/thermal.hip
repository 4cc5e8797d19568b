/*
 * thermal.hip - dsmc_suite's thermal model (src/thermal/thermal_model.c:
 * thermal_params_default, thermal_solve) on the device.
 *
 * Every surface cell (td, fd) is an independent 1-D problem: a depth profile of
 * h points (467 for the default parameters), n_substep substeps of [surface
 * temperature by bisection on the energy balance, then one Crank-Nicolson step
 * solved by the Thomas algorithm].  One lane per cell; the profiles live in HBM
 * as [depth][cell], so a wavefront's lanes read and write consecutive cells at
 * each depth.  The tridiagonal matrix is the same for every cell and substep,
 * so its elimination coefficients (the Thomas algorithm's c'_i and
 * denominators, thermal_model.c:86-92) are computed once on the host with the
 * same operations; each lane then runs the forward sweep with the right-hand
 * side built on the fly from its old profile, and the back substitution.
 * Arithmetic follows thermal_model.c operation for operation (no contraction);
 * log10 and 10^y come from dsmc_math.h, so the result equals the oracle's
 * device-math build bit for bit.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <vector>

#include "../../include/dsmc_hip.h"
#include "dsmc_math.h"
#include "dsmc_host.h"

#define HIP_OK(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) {                                                           \
            fprintf(stderr, "dsmc_hip: %s failed: %s (%s:%d)\n", #expr,                  \
                    hipGetErrorString(e_), __FILE__, __LINE__);                           \
            rc = DSMC_ERR_DEVICE;                                                         \
            goto done;                                                                    \
        }                                                                                 \
    } while (0)

/* include/constants.h */
#define TH_K_B 1.38e-23
#define TH_MMOL (18e-3 / 6.02e23)
#define TH_SIGMA 5.6704e-8
#define TH_SOLAR 1368.0
#define TH_PERIOD (9.07417 * 3600.0)
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

struct ThermalGeom {
    int ftd, ffd, h, n_sub;
    double dr, R, k_over_dr, dF;
};

static int thermal_geom(int ftd, int ffd, const dsmc_thermal_params *tp, ThermalGeom *g)
{
    if (ftd <= 0 || ffd <= 0 || !tp || tp->n_substep <= 0 || !(tp->conductivity > 0) || !(tp->density > 0) ||
        !(tp->Cp > 0) || !(tp->depth > 0) || !(tp->distance_AU > 0))
        return DSMC_ERR_ARG;
    g->ftd = ftd;
    g->ffd = ffd;
    g->n_sub = tp->n_substep;
    const double dt = TH_PERIOD / ffd / g->n_sub;                   /* thermal_model.c:109-117 */
    const double kk = tp->conductivity, lo = tp->density, Cp = tp->Cp;
    g->dr = sqrt(0.1 * 2.0 * dt * kk / lo / Cp);
    const double hd = ceil(tp->depth / g->dr);
    if (!(hd < 1e6)) return DSMC_ERR_ARG;
    g->h = hd < 3 ? 3 : (int)hd;
    g->R = kk * dt / lo / 2.0 / Cp / (g->dr * g->dr);
    g->k_over_dr = kk / g->dr;
    const double Si = TH_SOLAR * (1.0 / tp->distance_AU) * (1.0 / tp->distance_AU);   /* :131-134 */
    g->dF = Si * (1.0 - tp->albedo);
    return DSMC_OK;
}

/* thermal_model.c:52-75 */
__device__ static double th_balance(double T, double F, double k_over_dr, double oT, const dsmc_thermal_params &tp)
{
    const double *a = tp.p_mat, *b = tp.LH_mat;
    double lg = a[0] / T + a[1] + a[2] * T + a[3] * T * T + a[4] * dm_log10(T);
    double p = tp.tran_unit_p * dm_pow10(lg);
    double LH = tp.tran_unit_LH * (b[0] / T + b[1] + b[2] * T + b[3] * T * T + b[4] * dm_log10(T));
    double sublim = LH * sqrt(TH_MMOL / (2.0 * M_PI * TH_K_B * T)) * p;
    return F + k_over_dr * oT - k_over_dr * T - TH_SIGMA * T * T * T * T - sublim;
}

/* One lane per cell.  q, dp: [h][ncell]; a, den, cp: [h] (row i: a_i x_{i-1} +
 * b_i x_i + c_i x_{i+1} = d_i, den_i = b_i - a_i c'_{i-1}, c'_i = c_i / den_i). */
__global__ void __launch_bounds__(64) thermal_kernel(ThermalGeom g, dsmc_thermal_params tp, const double *flux,
                                                     const double *a, const double *den, const double *cp, double *q,
                                                     double *dp, double *T_out)
{
    const int ncell = g.ftd * g.ffd;
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= ncell) return;
    const int h = g.h;
    const size_t N = (size_t)ncell;
    double F;
    if (flux) F = flux[cell];
    else {                                                            /* thermal_model.c:136-145 */
        const int td = cell / g.ffd, fd = cell % g.ffd;
        const double thi = (td + 0.5) / g.ftd * M_PI - M_PI / 2.0;
        const double fi = (fd + 0.5) / g.ffd * 2.0 * M_PI;
        F = fi >= M_PI ? 0.0 : g.dF * dm_cos(thi) * dm_sin(fi);
    }
    F = fabs(F);
    const double T_init = F > 0 ? dm_root4(F / TH_SIGMA) : 100.0;    /* :170-178 */
    for (int i = 0; i < h; i++) q[i * N + cell] = T_init;
    const double R = g.R, omr = 1.0 - 2.0 * R;
    for (int hh = 0; hh < g.n_sub; hh++) {
        const double oT = q[N + cell];                                 /* :199-222 */
        double T0 = 50.0, T1 = 300.0;
        double f0 = th_balance(T0, F, g.k_over_dr, oT, tp), f1 = th_balance(T1, F, g.k_over_dr, oT, tp);
        for (int m = 0; m < 30; m++) {
            if (T1 >= 10000.0) break;
            if (f0 * f1 > 0) { T1 *= 5.0; f1 = th_balance(T1, F, g.k_over_dr, oT, tp); continue; }
            const double T2 = (T0 + T1) / 2.0;
            const double f2 = th_balance(T2, F, g.k_over_dr, oT, tp);
            if (f0 * f2 <= 0) { T1 = T2; f1 = f2; }
            else { T0 = T2; f0 = f2; }
            if (fabs(T0 - T1) <= 0.01) break;
        }
        /* forward sweep: d_0 = surface temperature, d_i = R q_{i-1} + (1-2R) q_i +
         * R q_{i+1} from the old profile (q_0 already the new surface value),
         * d_{h-1} = 0 (:224-234, :86-92) */
        double qm = (T0 + T1) / 2.0;
        q[cell] = qm;
        double qc = q[N + cell];
        double d_prev = qm / 1.0;                                      /* dp_0 = d_0 / b_0, b_0 = 1 */
        dp[cell] = d_prev;
        for (int i = 1; i < h - 1; i++) {
            const double qn = q[(size_t)(i + 1) * N + cell];
            const double d = R * qm + omr * qc + R * qn;
            d_prev = (d - a[i] * d_prev) / den[i];
            dp[(size_t)i * N + cell] = d_prev;
            qm = qc;
            qc = qn;
        }
        d_prev = (0.0 - a[h - 1] * d_prev) / den[h - 1];
        /* back substitution (:95-98) */
        double x = d_prev;
        q[(size_t)(h - 1) * N + cell] = x;
        for (int i = h - 2; i >= 0; i--) {
            x = dp[(size_t)i * N + cell] - cp[i] * x;
            q[(size_t)i * N + cell] = x;
        }
    }
    T_out[cell] = q[cell];
}

static int thermal_run(int ftd, int ffd, const dsmc_thermal_params *tp, const double *flux, double *T_out,
                       ThermalGeom *geom_out)
{
    ThermalGeom g;
    int rc = thermal_geom(ftd, ffd, tp, &g);
    if (rc) return rc;
    if (!T_out) return DSMC_ERR_ARG;
    if (geom_out) *geom_out = g;
    if ((rc = dsmc::ensure_device())) return rc;
    const int h = g.h;
    const size_t nc = (size_t)ftd * ffd;
    /* the tridiagonal rows (:237-249) and their elimination (:86-92) */
    std::vector<double> a(h), b(h), c(h), den(h), cp(h);
    a[0] = 0.0; b[0] = 1.0; c[0] = 0.0;
    for (int i = 1; i < h - 1; i++) { a[i] = -g.R; b[i] = 1.0 + 2.0 * g.R; c[i] = -g.R; }
    a[h - 1] = -1.0 / g.dr; b[h - 1] = 1.0 / g.dr; c[h - 1] = 0.0;
    den[0] = b[0];
    cp[0] = c[0] / b[0];
    for (int i = 1; i < h; i++) {
        den[i] = b[i] - a[i] * cp[i - 1];
        cp[i] = c[i] / den[i];
    }
    double *d_flux = nullptr, *d_tab = nullptr, *d_q = nullptr, *d_dp = nullptr, *d_T = nullptr;
    HIP_OK(hipMalloc((void **)&d_tab, sizeof(double) * 3 * h));
    HIP_OK(hipMalloc((void **)&d_q, sizeof(double) * h * nc));
    HIP_OK(hipMalloc((void **)&d_dp, sizeof(double) * h * nc));
    HIP_OK(hipMalloc((void **)&d_T, sizeof(double) * nc));
    HIP_OK(hipMemcpy(d_tab, a.data(), sizeof(double) * h, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d_tab + h, den.data(), sizeof(double) * h, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(d_tab + 2 * h, cp.data(), sizeof(double) * h, hipMemcpyHostToDevice));
    if (flux) {
        HIP_OK(hipMalloc((void **)&d_flux, sizeof(double) * nc));
        HIP_OK(hipMemcpy(d_flux, flux, sizeof(double) * nc, hipMemcpyHostToDevice));
    }
    hipLaunchKernelGGL(thermal_kernel, dim3((unsigned)((nc + 63) / 64)), dim3(64), 0, 0, g, *tp, d_flux, d_tab,
                       d_tab + h, d_tab + 2 * h, d_q, d_dp, d_T);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpy(T_out, d_T, sizeof(double) * nc, hipMemcpyDeviceToHost));
done:
    for (double *p : {d_flux, d_tab, d_q, d_dp, d_T})
        if (p) (void)hipFree(p);
    return rc;
}

extern "C" int dsmc_thermal_solve(int ftd, int ffd, const dsmc_thermal_params *tp, const double *flux,
                                  double *T_surface)
{
    return thermal_run(ftd, ffd, tp, flux, T_surface, nullptr);
}

/* thermal_model.c:23-49 */
extern "C" void thermal_params_default(dsmc_thermal_params *tp)
{
    memset(tp, 0, sizeof(*tp));
    tp->density = 730.0;
    tp->conductivity = 6e-4;
    tp->Cp = 1300.0;
    tp->albedo = 0.02;
    tp->distance_AU = 2.6;
    tp->depth = 0.05;
    tp->n_substep = 5;
    tp->z_angle = 12.3 / 180.0 * M_PI;
    const double pm[5] = {-2445.5646, -6.757169, -0.01677006, 1.20514e-5, 8.2312};
    const double lh[5] = {0.0, 12420.0, -4.8, 0.0, 0.0};
    for (int i = 0; i < 5; i++) { tp->p_mat[i] = pm[i]; tp->LH_mat[i] = lh[i]; }
    tp->tran_unit_p = 133.32237;
    tp->tran_unit_LH = 4.2;
}

/* thermal_model.c:105-283, the reference's entry point and progress lines */
extern "C" void thermal_solve(dsmc_sph_grid *grid, const dsmc_thermal_params *tp, const double *flux, int nthreads)
{
    (void)nthreads;
    ThermalGeom g;
    if (!grid || thermal_geom(grid->ftd, grid->ffd, tp, &g)) {
        fprintf(stderr, "thermal_solve: invalid arguments\n");
        return;
    }
    printf("Thermal model: dr=%.6f m, depth points=%d\n", g.dr, g.h);
    printf("Running thermal diffusion...\n");
    int rc = thermal_run(grid->ftd, grid->ffd, tp, flux, grid->temperature, nullptr);
    if (rc) {
        fprintf(stderr, "thermal_solve: device run failed (%d)\n", rc);
        return;
    }
    printf("Thermal model done.\n");
}
