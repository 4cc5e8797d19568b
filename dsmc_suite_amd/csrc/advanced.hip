/*
 * advanced.hip - phase 2 of the collisional path (src/dsmc/advanced.c:119-476).
 *
 * Setup on the device, from the phase-1 velocity records:
 *   1. cell of every record, [td][fd][ht] (advanced.c:187-194), histogram;
 *   2. exclusive prefix scan of the histogram (one block: per-lane serial sums,
 *      then a wavefront scan with DPP-free shuffles, then across waves in LDS);
 *   3. scatter record ids into per-cell buckets (counting sort);
 *   4. one lane per cell: keep the first max_per_cell records in record order
 *      (rank = number of smaller record ids in the bucket), their speeds, sort
 *      them, histogram into max(4, (nn+1)/2) bins, reverse cumulative counts and
 *      the least-squares polynomial of polyfit.c:10-101 (normal equations,
 *      Gaussian elimination with partial pivoting), exactly as advanced.c:32-117.
 * Then the particle kernel runs with the collision branch (dsmc_lane.h).
 */
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>
#include <vector>

#include "../../include/dsmc_hip.h"
#include "dsmc_lane.h"
#include "dsmc_polyfit.h"
#include "dsmc_kernels.h"
#include "dsmc_host.h"

#define HIP_OK(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess) {                                                           \
            fprintf(stderr, "dsmc_hip: %s failed: %s (%s:%d)\n", #expr,                  \
                    hipGetErrorString(e_), __FILE__, __LINE__);                           \
            return DSMC_ERR_DEVICE;                                                       \
        }                                                                                 \
    } while (0)

#define FIT_MAXN 64      /* max_per_cell limit of the fit kernel */
#define FIT_BLOCK 64

__global__ void rec_cell_kernel(const int32_t *td, const int32_t *fd, const int32_t *ht, int64_t n,
                                int ftd, int ffd, int hd, int32_t *cell, int32_t *hist)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int a = td[i], b = fd[i], c = ht[i];
    int k = -1;
    if (a >= 0 && a < ftd && b >= 0 && b < ffd && c >= 0 && c < hd) k = a * ffd * hd + b * hd + c;
    cell[i] = k;
    if (k >= 0) atomicAdd(&hist[k], 1);
}

/* Exclusive scan of hist[0..n) into start[0..n) with one block of 1024 lanes. */
__global__ void __launch_bounds__(1024) scan_kernel(const int32_t *hist, int32_t *start, int n)
{
    __shared__ int32_t wave_sum[16];
    const int t = threadIdx.x, nt = blockDim.x;
    const int per = (n + nt - 1) / nt;
    const int lo = t * per, hi = lo + per < n ? lo + per : n;
    int32_t s = 0;
    for (int i = lo; i < hi; i++) s += hist[i];
    /* inclusive scan of s across the wave */
    const int lane = t & 63, w = t >> 6;
    int32_t v = s;
    for (int off = 1; off < 64; off <<= 1) {
        int32_t u = __shfl_up(v, off);
        if (lane >= off) v += u;
    }
    if (lane == 63) wave_sum[w] = v;
    __syncthreads();
    int32_t wbase = 0;
    for (int j = 0; j < w; j++) wbase += wave_sum[j];
    int32_t run = wbase + v - s;        /* exclusive prefix of this lane's chunk */
    for (int i = lo; i < hi; i++) { start[i] = run; run += hist[i]; }
}

__global__ void scatter_kernel(const int32_t *cell, int64_t n, const int32_t *start, int32_t *fill,
                               int32_t *bucket)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int k = cell[i];
    if (k < 0) return;
    int pos = start[k] + atomicAdd(&fill[k], 1);
    bucket[pos] = (int32_t)i;
}

__global__ void __launch_bounds__(FIT_BLOCK)
fit_kernel(const int32_t *hist, const int32_t *start, const int32_t *bucket, const double *vrec, int ncell,
           int max_per_cell, int degree, int min_samples, int32_t *count_num, double *coeffs,
           double *pmin, double *pmax, int32_t *valid)
{
    __shared__ double vd_s[FIT_BLOCK][FIT_MAXN];
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    double *vd = vd_s[threadIdx.x];
    const int m = hist[c], s0 = start[c];
    /* first max_per_cell records of this cell in record order (advanced.c:195-201) */
    int nn = 0;
    for (int j = 0; j < m; j++) {
        int id = bucket[s0 + j];
        int rank = 0;
        if (m > max_per_cell)
            for (int k = 0; k < m; k++) rank += bucket[s0 + k] < id;
        if (rank < max_per_cell) {
            const double *v = vrec + 3 * (size_t)id;
            double sp = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            /* insertion keeps vd sorted (advanced.c:55-64 sorts before use) */
            int i = nn++;
            while (i > 0 && vd[i - 1] > sp) { vd[i] = vd[i - 1]; i--; }
            vd[i] = sp;
        }
    }
    count_num[c] = nn;
    valid[c] = 0;
    if (nn <= min_samples) return;
    double min_v = vd[0], max_v = vd[nn - 1];
    int nbins = (nn + 1) / 2;
    if (nbins < 4) nbins = 4;
    double xx[FIT_MAXN / 2 + 4], yy[FIT_MAXN / 2 + 4], yy2[FIT_MAXN / 2 + 4];
    double dv = (max_v - min_v) / nbins;
    for (int i = 0; i < nbins; i++) { xx[i] = min_v + (i + 0.5) * dv; yy[i] = 0.0; }
    for (int i = 0; i < nn; i++) {
        int bin = (int)((vd[i] - min_v) / dv);
        if (bin >= nbins) bin = nbins - 1;
        if (bin < 0) bin = 0;
        yy[bin] += 1.0;
    }
    for (int i = 0; i < nbins; i++) {
        yy2[i] = 0.0;
        for (int j = i; j < nbins; j++) yy2[i] += yy[j];
    }
    double cf[7];
    if (lsq_poly(xx, yy2, nbins, degree, cf) == 0) {
        for (int i = 0; i <= degree; i++) coeffs[(size_t)c * (degree + 1) + i] = cf[i];
        pmin[c] = min_v;
        pmax[c] = max_v;
        valid[c] = 1;
    }
}

/* Physical collisional runs: partner-speed fits from phase-1 speed histograms,
 * one lane per cell (cells [ht][td][fd]).  P at the occupied bins lo..hi is 50 x
 * (steps in bins >= k) / (steps of the cell), fitted in the normalised speed
 * x = (k - lo + 0.5) / (hi - lo + 1); valid when the cell has more than
 * min_samples steps and more occupied-range bins than coefficients. */
__global__ void hist_fit_kernel(const unsigned long long *vhist, int ncell, int bins, double dv, int degree,
                                int min_samples, double *coeffs, double *pmin, double *pmax, int32_t *valid,
                                int32_t *count_num)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const unsigned long long *h = vhist + (size_t)c * bins;
    unsigned long long N = 0;
    int lo = -1, hi = -1;
    for (int b = 0; b < bins; b++)
        if (h[b]) { N += h[b]; if (lo < 0) lo = b; hi = b; }
    count_num[c] = DSMC_VIRTUAL_RECORDS;
    valid[c] = 0;
    if (lo < 0 || N <= (unsigned long long)(min_samples > 0 ? min_samples : 0)) return;
    const int npts = hi - lo + 1;
    if (npts < degree + 2) return;
    double x[64], y[64], cf[7];
    unsigned long long acc = 0;
    for (int k = hi; k >= lo; k--) {
        acc += h[k];
        x[k - lo] = ((double)(k - lo) + 0.5) / (double)npts;
        y[k - lo] = (double)DSMC_VIRTUAL_RECORDS * (double)acc / (double)N;
    }
    if (lsq_poly(x, y, npts, degree, cf) != 0) return;
    for (int i = 0; i <= degree; i++) coeffs[(size_t)c * (degree + 1) + i] = cf[i];
    pmin[c] = (double)lo * dv;
    pmax[c] = (double)(hi + 1) * dv;
    valid[c] = 1;
}

/* n sigma per cell where a partner fit exists (collisions need a partner). */
__global__ void nsig_kernel(const double *ndens, const int32_t *valid, double sigma, int ncell, double *nsig)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    nsig[c] = valid[c] ? ndens[c] * sigma : 0.0;
}

extern "C" hipError_t dsmc_launch_hist_fits(const unsigned long long *vhist, int ncell, int bins, double dv, int degree,
                                            int min_samples, double *coeffs, double *pmin, double *pmax,
                                            int32_t *valid, int32_t *count_num, hipStream_t st)
{
    hipLaunchKernelGGL(hist_fit_kernel, dim3((ncell + 63) / 64), dim3(64), 0, st, vhist, ncell, bins, dv, degree,
                       min_samples, coeffs, pmin, pmax, valid, count_num);
    return hipGetLastError();
}

extern "C" hipError_t dsmc_launch_nsig(const double *ndens, const int32_t *valid, double sigma, int ncell, double *nsig,
                                       hipStream_t st)
{
    hipLaunchKernelGGL(nsig_kernel, dim3((ncell + 255) / 256), dim3(256), 0, st, ndens, valid, sigma, ncell, nsig);
    return hipGetLastError();
}

using namespace dsmc;

/* Physical phase-2 tables: fits from the histograms (vhist_dev: device array,
 * or NULL to upload in->vhist), n sigma from in->ndens. */
static int stage_phase2_phys(const dsmc_params *p, const dsmc_phase2_input *in, const unsigned long long *vhist_dev,
                             RunDev *rd, hipStream_t st)
{
    const int ncell = p->ftd * p->ffd * p->hd, deg = p->fit_degree, bins = p->vhist_bins;
    int rc;
    DevBuf vh, nd;
    if (!vhist_dev) {
        if ((rc = vh.upload(in->vhist, sizeof(int64_t) * (size_t)ncell * bins, st))) return rc;
        vhist_dev = (const unsigned long long *)vh.ptr;
    }
    if ((rc = nd.upload(in->ndens, sizeof(double) * ncell, st))) return rc;
    if ((rc = rd->count_num.zeros(sizeof(int32_t) * ncell, st))) return rc;
    if ((rc = rd->coeffs.zeros(sizeof(double) * (size_t)ncell * (deg + 1), st))) return rc;
    if ((rc = rd->pmin.zeros(sizeof(double) * ncell, st))) return rc;
    if ((rc = rd->pmax.zeros(sizeof(double) * ncell, st))) return rc;
    if ((rc = rd->pvalid.zeros(sizeof(int32_t) * ncell, st))) return rc;
    if ((rc = rd->nsig.alloc(sizeof(double) * ncell))) return rc;
    const double dv = p->vhist_max / p->vhist_bins;
    hipLaunchKernelGGL(hist_fit_kernel, dim3((ncell + 63) / 64), dim3(64), 0, st, vhist_dev, ncell, bins, dv, deg,
                       p->min_fit_samples, (double *)rd->coeffs.ptr, (double *)rd->pmin.ptr, (double *)rd->pmax.ptr,
                       (int32_t *)rd->pvalid.ptr, (int32_t *)rd->count_num.ptr);
    HIP_OK(hipGetLastError());
    hipLaunchKernelGGL(nsig_kernel, dim3((ncell + 255) / 256), dim3(256), 0, st, (const double *)nd.ptr,
                       (const int32_t *)rd->pvalid.ptr, p->cross_section, ncell, (double *)rd->nsig.ptr);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(st));   /* temporaries are freed on return */
    DsmcEnv &E = rd->E;
    E.count_num = (const int32_t *)rd->count_num.ptr;
    E.coeffs = (const double *)rd->coeffs.ptr;
    E.pmin = (const double *)rd->pmin.ptr;
    E.pmax = (const double *)rd->pmax.ptr;
    E.pvalid = (const int32_t *)rd->pvalid.ptr;
    E.nsig = (const double *)rd->nsig.ptr;
    E.fit_deg = deg;
    E.copy_atm = nullptr;
    E.atm_total = 0;
    return DSMC_OK;
}

static int check_phys_coll(const dsmc_params *p)
{
    if (!p || p->mode != DSMC_MODE_PHYSICAL || p->fit_degree < 1 || p->fit_degree > 6 || p->vhist_bins <= 0 ||
        p->vhist_bins > 64 || !(p->vhist_max > 0) || !(p->cross_section >= 0))
        return DSMC_ERR_ARG;
    return DSMC_OK;
}

/* Stages the phase-2 tables of `in` into rd (device) and points rd->E at them. */
static int stage_phase2(const dsmc_params *p, const dsmc_phase2_input *in, RunDev *rd, hipStream_t st)
{
    const int ftd = p->ftd, ffd = p->ffd, hd = p->hd;
    const int ncell = ftd * ffd * hd;
    const int deg = p->fit_degree;
    int rc;
    if ((rc = rd->count_num.zeros(sizeof(int32_t) * ncell, st))) return rc;
    if ((rc = rd->coeffs.zeros(sizeof(double) * (size_t)ncell * (deg + 1), st))) return rc;
    if ((rc = rd->pmin.zeros(sizeof(double) * ncell, st))) return rc;
    if ((rc = rd->pmax.zeros(sizeof(double) * ncell, st))) return rc;
    if ((rc = rd->pvalid.zeros(sizeof(int32_t) * ncell, st))) return rc;
    if (in->vrec && in->nrec > 0) {
        if (!in->vrec_td || !in->vrec_fd || !in->vrec_ht || in->nrec > INT32_MAX) return DSMC_ERR_ARG;
        const int64_t n = in->nrec;
        DevBuf dv, dtd, dfd, dht, dcell, hist, start, fill, bucket;
        if ((rc = dv.upload(in->vrec, sizeof(double) * 3 * n, st))) return rc;
        if ((rc = dtd.upload(in->vrec_td, sizeof(int32_t) * n, st))) return rc;
        if ((rc = dfd.upload(in->vrec_fd, sizeof(int32_t) * n, st))) return rc;
        if ((rc = dht.upload(in->vrec_ht, sizeof(int32_t) * n, st))) return rc;
        if ((rc = dcell.alloc(sizeof(int32_t) * n))) return rc;
        if ((rc = hist.zeros(sizeof(int32_t) * ncell, st))) return rc;
        if ((rc = start.alloc(sizeof(int32_t) * ncell))) return rc;
        if ((rc = fill.zeros(sizeof(int32_t) * ncell, st))) return rc;
        if ((rc = bucket.alloc(sizeof(int32_t) * n))) return rc;
        int nb = (int)((n + 255) / 256);
        hipLaunchKernelGGL(rec_cell_kernel, dim3(nb), dim3(256), 0, st, (const int32_t *)dtd.ptr,
                           (const int32_t *)dfd.ptr, (const int32_t *)dht.ptr, n, ftd, ffd, hd,
                           (int32_t *)dcell.ptr, (int32_t *)hist.ptr);
        HIP_OK(hipGetLastError());
        hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(1024), 0, st, (const int32_t *)hist.ptr,
                           (int32_t *)start.ptr, ncell);
        HIP_OK(hipGetLastError());
        hipLaunchKernelGGL(scatter_kernel, dim3(nb), dim3(256), 0, st, (const int32_t *)dcell.ptr, n,
                           (const int32_t *)start.ptr, (int32_t *)fill.ptr, (int32_t *)bucket.ptr);
        HIP_OK(hipGetLastError());
        hipLaunchKernelGGL(fit_kernel, dim3((ncell + FIT_BLOCK - 1) / FIT_BLOCK), dim3(FIT_BLOCK), 0, st,
                           (const int32_t *)hist.ptr, (const int32_t *)start.ptr, (const int32_t *)bucket.ptr,
                           (const double *)dv.ptr, ncell, p->max_per_cell, deg, p->min_fit_samples,
                           (int32_t *)rd->count_num.ptr, (double *)rd->coeffs.ptr, (double *)rd->pmin.ptr,
                           (double *)rd->pmax.ptr, (int32_t *)rd->pvalid.ptr);
        HIP_OK(hipGetLastError());
        HIP_OK(hipStreamSynchronize(st));   /* the temporaries above are freed on return */
    }
    DsmcEnv &E = rd->E;
    E.count_num = (const int32_t *)rd->count_num.ptr;
    E.coeffs = (const double *)rd->coeffs.ptr;
    E.pmin = (const double *)rd->pmin.ptr;
    E.pmax = (const double *)rd->pmax.ptr;
    E.pvalid = (const int32_t *)rd->pvalid.ptr;
    E.copy_atm = nullptr;
    E.atm_total = 0;
    if (in->copy_atm) {
        int64_t total = (int64_t)in->atm_hd * in->atm_ftd * in->atm_ffd;
        if (total <= 0) return DSMC_ERR_ARG;
        std::vector<double> atm((size_t)total);
        for (int64_t i = 0; i < total; i++) atm[i] = in->copy_atm[i] * p->density_scale;
        if ((rc = rd->atm.upload(atm.data(), sizeof(double) * total, st))) return rc;
        E.copy_atm = (const double *)rd->atm.ptr;
        E.atm_total = total;
    }
    return DSMC_OK;
}

static int check_phase2(const dsmc_params *p, const dsmc_phase2_input *in)
{
    if (p && in && p->mode == DSMC_MODE_PHYSICAL)
        return check_phys_coll(p) ? DSMC_ERR_ARG : (in->ndens && in->vhist ? DSMC_OK : DSMC_ERR_ARG);
    if (!p || !in || p->mode != DSMC_MODE_COMPAT || p->fit_degree < 1 || p->fit_degree > 6 ||
        p->max_per_cell <= 0 || p->max_per_cell > FIT_MAXN)
        return DSMC_ERR_ARG;
    return DSMC_OK;
}

extern "C" int dsmc_build_fits(const dsmc_params *p, const dsmc_phase2_input *in, int32_t *count_num,
                               double *coeffs, double *min_v, double *max_v, int32_t *valid)
{
    int rc = check_phase2(p, in);
    if (rc) return rc;
    if ((rc = ensure_device())) return rc;
    RunDev rd;
    memset(&rd.E, 0, sizeof rd.E);
    hipStream_t st;
    HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    rc = stage_phase2(p, in, &rd, st);
    const size_t ncell = (size_t)p->ftd * p->ffd * p->hd;
    if (!rc) {
        if (count_num) HIP_OK(hipMemcpy(count_num, rd.count_num.ptr, 4 * ncell, hipMemcpyDeviceToHost));
        if (coeffs) HIP_OK(hipMemcpy(coeffs, rd.coeffs.ptr, rd.coeffs.bytes, hipMemcpyDeviceToHost));
        if (min_v) HIP_OK(hipMemcpy(min_v, rd.pmin.ptr, 8 * ncell, hipMemcpyDeviceToHost));
        if (max_v) HIP_OK(hipMemcpy(max_v, rd.pmax.ptr, 8 * ncell, hipMemcpyDeviceToHost));
        if (valid) HIP_OK(hipMemcpy(valid, rd.pvalid.ptr, 4 * ncell, hipMemcpyDeviceToHost));
    }
    (void)hipStreamDestroy(st);
    return rc;
}

extern "C" int dsmc_run_advanced(const dsmc_params *p, const dsmc_surface *s, const dsmc_phase2_input *in,
                                 dsmc_result *r)
{
    int rc = check_phase2(p, in);
    if (rc || !s || !r || p->nparticles < 0) return rc ? rc : DSMC_ERR_ARG;
    if ((rc = ensure_device())) return rc;
    TimedStream ts;
    if (ts.create()) return DSMC_ERR_DEVICE;
    hipStream_t st = ts.st;
    hipEvent_t e0 = ts.e0, e1 = ts.e1;
    RunDev rd;
    rc = stage_run(p, s, r, 1, &rd, st);
    if (!rc) rc = p->mode == DSMC_MODE_PHYSICAL ? stage_phase2_phys(p, in, nullptr, &rd, st) : stage_phase2(p, in, &rd, st);
    if (!rc) {
        HIP_OK(hipEventRecord(e0, st));
        rc = launch_run(&rd, nullptr, st);
        HIP_OK(hipEventRecord(e1, st));
    }
    if (!rc) rc = collect_run(&rd, p, r, st);
    if (!rc) {
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, e0, e1));
        r->elapsed_ms = ms;
    }
    return rc;
}

extern "C" int dsmc_build_hist_fits(const dsmc_params *p, const int64_t *vhist, double *coeffs, double *min_v,
                                    double *max_v, int32_t *valid)
{
    int rc = check_phys_coll(p);
    if (rc || !vhist) return DSMC_ERR_ARG;
    if ((rc = ensure_device())) return rc;
    const size_t ncell = (size_t)p->ftd * p->ffd * p->hd;
    std::vector<double> zeros(ncell, 0.0);
    dsmc_phase2_input in;
    memset(&in, 0, sizeof in);
    in.vhist = vhist;
    in.ndens = zeros.data();
    RunDev rd;
    memset(&rd.E, 0, sizeof rd.E);
    hipStream_t st;
    HIP_OK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    rc = stage_phase2_phys(p, &in, nullptr, &rd, st);
    if (!rc) {
        if (coeffs) HIP_OK(hipMemcpy(coeffs, rd.coeffs.ptr, rd.coeffs.bytes, hipMemcpyDeviceToHost));
        if (min_v) HIP_OK(hipMemcpy(min_v, rd.pmin.ptr, 8 * ncell, hipMemcpyDeviceToHost));
        if (max_v) HIP_OK(hipMemcpy(max_v, rd.pmax.ptr, 8 * ncell, hipMemcpyDeviceToHost));
        if (valid) HIP_OK(hipMemcpy(valid, rd.pvalid.ptr, 4 * ncell, hipMemcpyDeviceToHost));
    }
    (void)hipStreamDestroy(st);
    return rc;
}

extern "C" int dsmc_number_density(const dsmc_params *p1, const dsmc_surface *s, const int64_t *vcount,
                                   double *ndens)
{
    if (!p1 || !s || !vcount || !ndens || p1->mode != DSMC_MODE_PHYSICAL || !s->face_w || s->levels < 0 ||
        s->levels > 10 || p1->nparticles <= 0 || p1->ftd <= 0 || p1->ffd <= 0 || p1->hd <= 0)
        return DSMC_ERR_ARG;
    const int nfine = 20 << (2 * s->levels);
    double F = 0.0;
    for (int f = 0; f < nfine; f++) F += s->face_w[f];
    number_density(p1, F, vcount, ndens);
    return DSMC_OK;
}

namespace dsmc {
/* ndens from the sum of the face weights (in face order). */
void number_density(const dsmc_params *p1, double wsum, const int64_t *vcount, double *ndens)
{
    const double R = p1->radius;
    const double F = (p1->production_rate > 0 ? p1->production_rate : wsum * R * R) * p1->sublimation_scale;
    const int ftd = p1->ftd, ffd = p1->ffd, hd = p1->hd;
    std::vector<double> redge(hd + 1), sl(ftd + 1);
    for (int k = 0; k <= hd; k++)
        redge[k] = p1->exp_grid ? R + p1->unih * (pow(1.0 + p1->alpha2, k) - 1.0) / p1->alpha2 : R + p1->unih * k;
    for (int t = 0; t <= ftd; t++) sl[t] = sin(-M_PI / 2.0 + M_PI * t / ftd);
    const double dphi = 2.0 * M_PI / ffd, N = (double)p1->nparticles;
    for (int ht = 0; ht < hd; ht++) {
        const double r0 = redge[ht], r1 = redge[ht + 1];
        const double shell = (r1 * r1 * r1 - r0 * r0 * r0) / 3.0;
        for (int td = 0; td < ftd; td++) {
            const double V = shell * (sl[td + 1] - sl[td]) * dphi;
            for (int fd = 0; fd < ffd; fd++) {
                const size_t c = ((size_t)ht * ftd + td) * ffd + fd;
                ndens[c] = (double)vcount[c] * p1->dt * F / (N * V);
            }
        }
    }
}
}  // namespace dsmc

extern "C" int dsmc_run_collisional(const dsmc_params *p1, const dsmc_params *p2, const dsmc_surface *s,
                                    dsmc_result *r1, dsmc_result *r2)
{
    int rc = check_phys_coll(p1);
    if (rc || check_phys_coll(p2) || !s || !r2 || p1->ftd != p2->ftd || p1->ffd != p2->ffd || p1->hd != p2->hd ||
        p1->vhist_bins != p2->vhist_bins || p1->vhist_max != p2->vhist_max || p1->nparticles <= 0 ||
        p2->nparticles < 0)
        return DSMC_ERR_ARG;
    if ((rc = ensure_device())) return rc;
    const size_t ncell = (size_t)p1->ftd * p1->ffd * p1->hd;
    TimedStream ts;
    if (ts.create()) return DSMC_ERR_DEVICE;
    hipStream_t st = ts.st;
    hipEvent_t e0 = ts.e0, e1 = ts.e1;
    /* phase 1: the caller's outputs plus counts and histograms */
    dsmc_result spec1;
    if (r1) spec1 = *r1;
    else memset(&spec1, 0, sizeof spec1);
    std::vector<int64_t> cnt(ncell);
    spec1.vcount = cnt.data();
    int64_t dummy = 0;
    if (!spec1.vhist) spec1.vhist = &dummy;            /* non-NULL: keep histograms */
    {
        RunDev rd1;
        rc = stage_run(p1, s, &spec1, 0, &rd1, st);
        if (!rc) HIP_OK(hipEventRecord(e0, st));
        if (!rc) rc = launch_run(&rd1, nullptr, st);
        if (!rc) {
            dsmc_result out = spec1;
            if (out.vhist == &dummy) out.vhist = nullptr;
            rc = collect_run(&rd1, p1, &out, st);
            if (!rc && r1) {
                int64_t *keep_cnt = r1->vcount;
                *r1 = out;
                r1->vcount = keep_cnt;
                if (keep_cnt) memcpy(keep_cnt, cnt.data(), ncell * sizeof(int64_t));
            }
        }
        std::vector<double> nd(ncell);
        if (!rc) rc = dsmc_number_density(p1, s, cnt.data(), nd.data());
        /* phase 2, with the phase-1 histograms still on the device */
        RunDev rd2;
        if (!rc) rc = stage_run(p2, s, r2, 1, &rd2, st);
        dsmc_phase2_input in;
        memset(&in, 0, sizeof in);
        in.ndens = nd.data();
        if (!rc) rc = stage_phase2_phys(p2, &in, (const unsigned long long *)rd1.vhist.ptr, &rd2, st);
        if (!rc) rc = launch_run(&rd2, nullptr, st);
        if (!rc) HIP_OK(hipEventRecord(e1, st));
        if (!rc) rc = collect_run(&rd2, p2, r2, st);
    }
    if (!rc) {
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, e0, e1));
        r2->elapsed_ms = ms;
    }
    return rc;
}
