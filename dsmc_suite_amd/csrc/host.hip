/*
 * host.hip - the C ABI of libdsmc_hip.so (include/dsmc_hip.h).
 *
 * Host side of a run: derive the run's scalars and exact index tables on the CPU
 * (dsmc_tables.h), upload them with the surface, launch the persistent particle
 * kernel on one HIP stream, copy the results back.  The drop-in run_ballistic /
 * run_advanced wrap this with the reference's parameters and .dat files.
 */
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <time.h>
#include <limits.h>
#include <vector>
#include <mutex>
#include <chrono>

#include "../../include/dsmc_hip.h"
#include "dsmc_tables.h"
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

namespace dsmc {

/* ------------------------------------------------------------------ device buffers */
DevBuf::~DevBuf() { release(); }
void DevBuf::release()
{
    if (ptr && !borrowed) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
    borrowed = false;
}
int DevBuf::alloc(size_t n)
{
    release();
    if (n == 0) return DSMC_OK;
    if (hipMalloc(&ptr, n) != hipSuccess) { ptr = nullptr; return DSMC_ERR_NOMEM; }
    bytes = n;
    return DSMC_OK;
}
int DevBuf::upload(const void *src, size_t n, hipStream_t s)
{
    int rc = alloc(n);
    if (rc || n == 0) return rc;
    HIP_OK(hipMemcpyAsync(ptr, src, n, hipMemcpyHostToDevice, s));
    return DSMC_OK;
}
int DevBuf::zeros(size_t n, hipStream_t s)
{
    int rc = alloc(n);
    if (rc || n == 0) return rc;
    HIP_OK(hipMemsetAsync(ptr, 0, n, s));
    return DSMC_OK;
}

int DevBuf::take(Arena &a, size_t n)
{
    release();
    if (n == 0) return DSMC_OK;
    ptr = a.take(n);
    if (!ptr) return DSMC_ERR_BUDGET;
    borrowed = true;
    bytes = n;
    return DSMC_OK;
}
int DevBuf::upload(Arena &a, const void *src, size_t n, hipStream_t s)
{
    int rc = take(a, n);
    if (rc || n == 0) return rc;
    HIP_OK(hipMemcpyAsync(ptr, src, n, hipMemcpyHostToDevice, s));
    return DSMC_OK;
}
int DevBuf::zeros(Arena &a, size_t n, hipStream_t s)
{
    int rc = take(a, n);
    if (rc || n == 0) return rc;
    HIP_OK(dsmc_launch_fill0(ptr, (n + 7) / 8, s));     /* the arena rounds blocks up to 256 bytes */
    return DSMC_OK;
}

int device_props(int *cus)
{
    static int cached = -1;
    static std::mutex mu;
    std::lock_guard<std::mutex> g(mu);
    if (cached < 0) {
        int dev = 0;
        HIP_OK(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        HIP_OK(hipGetDeviceProperties(&prop, dev));
        if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
            fprintf(stderr, "dsmc_hip: warning: device is %s, kernels are built for gfx950\n",
                    prop.gcnArchName);
        cached = prop.multiProcessorCount;
    }
    *cus = cached;
    return DSMC_OK;
}

/* ------------------------------------------------------------------ run staging */
/* Uploads everything a particle kernel reads and allocates what it writes. */
int stage_run(const dsmc_params *p, const dsmc_surface *s, const dsmc_result *r, int advanced,
              RunDev *rd, hipStream_t st)
{
    int rc = build_env(p, s, advanced, &rd->E, &rd->T);
    if (rc) return rc;
    /* device buffers: hipMalloc'ed, or taken from the run's arena (cleared by a kernel) */
    Arena *ar = rd->arena;
    auto up = [&](DevBuf &b, const void *src, size_t n, hipStream_t s_) {
        return ar ? b.upload(*ar, src, n, s_) : b.upload(src, n, s_);
    };
    auto zz = [&](DevBuf &b, size_t n, hipStream_t s_) { return ar ? b.zeros(*ar, n, s_) : b.zeros(n, s_); };
    auto al = [&](DevBuf &b, size_t n) { return ar ? b.take(*ar, n) : b.alloc(n); };
    DsmcEnv &E = rd->E;
    HostTables &T = rd->T;
    const size_t cells = (size_t)p->hd * p->ftd * p->ffd;
    const int64_t n = p->nparticles;
#define UP(buf, vec) do { if ((rc = up((buf), (vec).data(), (vec).size() * sizeof((vec)[0]), st))) return rc; } while (0)
    UP(rd->tq, T.tq);
    UP(rd->hr, T.hr);
    UP(rd->cs, T.cs);
    E.tq = (const double *)rd->tq.ptr;
    E.hr = (const double *)rd->hr.ptr;
    E.cs = (const double *)rd->cs.ptr;
    if (p->mode == DSMC_MODE_COMPAT) {
        UP(rd->vt_speeds, T.vt_speeds);
        UP(rd->vt_count, T.vt_count);
        if ((rc = up(rd->Tsurf, s->T_latlon, sizeof(double) * p->ftd * p->ffd, st))) return rc;
        E.vt_speeds = (const double *)rd->vt_speeds.ptr;
        E.vt_count = (const int32_t *)rd->vt_count.ptr;
        E.T = (const double *)rd->Tsurf.ptr;
    } else {
        UP(rd->cdf, T.cdf);
        UP(rd->enorm, T.enorm);
        UP(rd->fnorm, T.fnorm);
        UP(rd->face_sig, T.face_sig);
        E.fmap_g = face_map_grid(T.nfine);
        build_face_map(E, T, E.fmap_g, T.fmap, &T.fmap_lat);
        UP(rd->fmap, T.fmap);
        if ((rc = up(rd->vert, s->vert, sizeof(double) * 3 * s->nvert, st))) return rc;
        if ((rc = up(rd->faces, s->faces, sizeof(int32_t) * 3 * T.nface_all, st))) return rc;
        if ((rc = up(rd->face_T, s->face_T, sizeof(double) * T.nfine, st))) return rc;
        E.cdf = (const double *)rd->cdf.ptr;
        E.enorm = (const double *)rd->enorm.ptr;
        E.fnorm = (const double *)rd->fnorm.ptr;
        E.fmap = (const int32_t *)rd->fmap.ptr;
        E.vert = (const double *)rd->vert.ptr;
        E.faces = (const int32_t *)rd->faces.ptr;
        E.face_T = (const double *)rd->face_T.ptr;
        E.face_sig = (const double *)rd->face_sig.ptr;
        /* short hops: the body holds its gas close (escape speed more than twice
         * the mean thermal speed at the hottest face), so a particle lands every
         * few dozen steps and the landing service belongs in the flight kernel */
        double tmax = 0.0;
        for (size_t f = 0; f < T.nfine; f++) tmax = s->face_T[f] > tmax ? s->face_T[f] : tmax;
        const double v_th = sqrt(2.0 * p->k_B * tmax / p->mmass);
        rd->short_hops = sqrt(2.0 * p->G * p->M / p->radius) > 2.0 * v_th ? 1 : 0;
    }
#undef UP
    if ((rc = zz(rd->density, cells * sizeof(double), st))) return rc;
    if ((rc = zz(rd->horizon, sizeof(double) * p->ftd * p->ffd, st))) return rc;
    E.density = (double *)rd->density.ptr;
    E.horizon = (double *)rd->horizon.ptr;
    if (r->vmom) {
        if ((rc = zz(rd->vmom, cells * 3 * sizeof(double), st))) return rc;
        E.vmom = (double *)rd->vmom.ptr;
        /* physical runs: deposits go to per-cell records (folded at the end) */
        if (p->mode == DSMC_MODE_PHYSICAL) {
            if ((rc = zz(rd->cellacc, cells * 4 * sizeof(double), st))) return rc;
            E.cellacc = (double *)rd->cellacc.ptr;
        }
    }
    if (r->vcount || p->mode == DSMC_MODE_PHYSICAL) {
        /* physical particles keep W = 1, so the density is the deposit count: one
         * atomic per cell visit instead of two */
        if ((rc = zz(rd->vcount, cells * sizeof(unsigned long long), st))) return rc;
        E.vcount = (unsigned long long *)rd->vcount.ptr;
        E.density_from_count = p->mode == DSMC_MODE_PHYSICAL;
    }
    if (r->vhist && p->mode == DSMC_MODE_PHYSICAL && !advanced) {
        if (p->vhist_bins <= 0 || p->vhist_bins > 64 || !(p->vhist_max > 0)) return DSMC_ERR_ARG;
        if ((rc = zz(rd->vhist, cells * p->vhist_bins * sizeof(unsigned long long), st))) return rc;
        E.vhist = (unsigned long long *)rd->vhist.ptr;
        E.vhist_bins = p->vhist_bins;
        E.vhist_inv = (double)p->vhist_bins / p->vhist_max;
    }
    int rows = advanced ? 9 : 5;
    E.ttime_rows = rows;
    if (r->ttime && p->mode == DSMC_MODE_COMPAT) {
        if ((rc = zz(rd->ttime, sizeof(double) * rows * n, st))) return rc;
        E.ttime = (double *)rd->ttime.ptr;
    }
    if (r->end_state) {
        if ((rc = zz(rd->end_state, sizeof(double) * 6 * n, st))) return rc;
        E.end_state = (double *)rd->end_state.ptr;
    }
    if (r->steps) {
        if ((rc = zz(rd->steps, sizeof(int64_t) * n, st))) return rc;
        E.steps_out = (int64_t *)rd->steps.ptr;
    }
    if (r->fate) {
        if ((rc = zz(rd->fate, sizeof(int32_t) * n, st))) return rc;
        E.fate_out = (int32_t *)rd->fate.ptr;
    }
    if (r->hops) {
        if ((rc = zz(rd->hops, sizeof(int32_t) * n, st))) return rc;
        E.hops_out = (int32_t *)rd->hops.ptr;
    }
    if (!advanced && p->mode == DSMC_MODE_COMPAT && p->record_cap_cell > 0 && r->vrec &&
        r->vrec_td && r->vrec_fd && r->vrec_ht && p->record_cap_total > 0) {
        size_t cap = (size_t)p->record_cap_total;
        if ((rc = zz(rd->vrec, cap * 3 * sizeof(double), st))) return rc;
        if ((rc = zz(rd->vrec_i, cap * 3 * sizeof(int32_t), st))) return rc;
        if ((rc = zz(rd->count_nn, sizeof(int32_t) * p->ftd * p->hd, st))) return rc;
        E.vrec = (double *)rd->vrec.ptr;
        E.vrec_td = (int32_t *)rd->vrec_i.ptr;
        E.vrec_fd = E.vrec_td + cap;
        E.vrec_ht = E.vrec_fd + cap;
        E.count_nn = (int32_t *)rd->count_nn.ptr;
    } else {
        E.rec_cap_cell = 0;
    }
    if ((rc = zz(rd->rec_total, sizeof(unsigned long long), st))) return rc;
    if ((rc = zz(rd->counters, sizeof(unsigned long long) * 8, st))) return rc;
    E.rec_total = (unsigned long long *)rd->rec_total.ptr;
    E.counters = (unsigned long long *)rd->counters.ptr;
    size_t ns = (size_t)(n > 0 ? n : 1);
    if (n > INT32_MAX) return DSMC_ERR_ARG;
    if ((rc = al(rd->rec, ns * sizeof(DsmcRec)))) return rc;
    if ((rc = al(rd->flyq0, ns * sizeof(int32_t)))) return rc;
    if ((rc = al(rd->flyq1, ns * sizeof(int32_t)))) return rc;
    if ((rc = al(rd->svcq, ns * sizeof(int32_t)))) return rc;
    if ((rc = zz(rd->qctl, sizeof(unsigned int) * 16, st))) return rc;
    if (ar) rd->host_ctl = rd->ctl_local;
    else if (!rd->host_ctl) {
        if (hipHostMalloc((void **)&rd->host_ctl, sizeof(unsigned int) * 16, 0) != hipSuccess) return DSMC_ERR_NOMEM;
        rd->ctl_pinned = true;
    }
    /* room for the environment in device memory (written by launch_run) */
    if ((rc = al(rd->env_d, sizeof(DsmcEnv)))) return rc;
    return DSMC_OK;
}

void RunDev::drop_buffers()
{
    DevBuf *all[] = {&tq, &hr, &cs, &vt_speeds, &vt_count, &Tsurf, &cdf, &enorm, &fnorm, &fmap, &vert, &faces,
                     &face_T, &face_sig, &density, &horizon, &vmom, &vcount, &cellacc, &ttime, &end_state, &steps, &fate,
                     &hops, &vrec, &vrec_i, &count_nn, &ctl, &atm, &count_num, &coeffs, &pmin, &pmax, &pvalid,
                     &speeds, &nsig, &vhist, &rec, &flyq0, &flyq1, &svcq, &qctl, &counters, &rec_total, &env_d};
    for (DevBuf *b : all) b->release();
}

RunDev::~RunDev()
{
    if (host_ctl && ctl_pinned) (void)hipHostFree(host_ctl);
}

/* Lanes whose hop ended that a wavefront of the late inline launch gathers
 * before running the landing service (the default of long-hop runs below 4e6
 * particles, which take that kernel from their start). */
#define LATE_INLINE_SVC_BATCH 24

/* Runs a staged run to completion on stream st: the start kernel, then rounds of
 * flight + service until no particle is in flight.  Completion is checked on the
 * host every `chunk` rounds (doubling up to 128), so a run costs few syncs. */
int launch_run(RunDev *rd, const dsmc_launch_shape *shape, hipStream_t st, int timed)
{
    int cus = 256;
    int rc = device_props(&cus);
    if (rc) return rc;
    dsmc_launch_shape sh = shape ? *shape : default_shape(rd->E, rd->short_hops);
    if (!shape_ok(sh)) return DSMC_ERR_ARG;
    /* physical collisional runs always service inline (collisions are service
     * events); compat advanced runs never (no kernel of that kind) */
    const bool compat_adv = rd->E.mode == DSMC_MODE_COMPAT && rd->E.advanced;
    const bool inline_svc = (sh.service_inline && !compat_adv) || (rd->E.mode == DSMC_MODE_PHYSICAL && rd->E.advanced);
    if (!shape) {
        /* the default asks for up to blocks_per_cu resident blocks; fewer if the
         * kernel's registers or LDS do not allow them */
        int occ = dsmc_flight_occupancy(&rd->E, inline_svc ? 1 : 0, sh.block);
        if (occ > 0 && occ < sh.blocks_per_cu) sh.blocks_per_cu = occ;
        /* physical collisional runs (inline by necessity): 1024-step launches */
        if (rd->E.mode == DSMC_MODE_PHYSICAL && rd->E.advanced && inline_svc && sh.step_budget > 1024)
            sh.step_budget = 1024;
    }
    const int64_t n = rd->E.nparticles;
    DsmcQueues Q;
    Q.fly[0] = (int32_t *)rd->flyq0.ptr;
    Q.fly[1] = (int32_t *)rd->flyq1.ptr;
    Q.svc = (int32_t *)rd->svcq.ptr;
    Q.ctl = (unsigned int *)rd->qctl.ptr;
    DsmcRec *rec = (DsmcRec *)rd->rec.ptr;
    HIP_OK(dsmc_launch_fill0(Q.ctl, sizeof(unsigned int) * 16 / 8, st));     /* a kernel, not a memset */
    /* the run's environment as the kernels' service code reads it from device memory */
    HIP_OK(hipMemcpyAsync(rd->env_d.ptr, &rd->E, sizeof(DsmcEnv), hipMemcpyHostToDevice, st));
    rd->rounds = 0;
    rd->flight_ms = 0.0;
    rd->lean_launches = 0;
    rd->lean_ms = 0.0;
    rd->lean_steps = 0;
    const bool lean = !inline_svc && rd->E.mode == DSMC_MODE_PHYSICAL && !rd->E.advanced;
    std::vector<char> is_lean;          /* per timed flight launch */
    if (n == 0) return DSMC_OK;
    int64_t blocks = (n + 255) / 256;
    int start_grid = (int)(blocks < (int64_t)cus * 8 ? blocks : (int64_t)cus * 8);
    HIP_OK(dsmc_launch_start(&rd->E, 0, rec, 0, n, Q, 0, start_grid, st));
    int flight_grid = cus * sh.blocks_per_cu;
    if ((int64_t)flight_grid > blocks) flight_grid = (int)blocks;
    int service_grid = start_grid;
    /* per-launch step budget: long hops are split over rounds so the run's tail is
     * set by its longest particle, not by the sum of each round's longest hop.
     * The default shape of the split kernels runs its first two rounds with the
     * full budget and later ones with at most 1024 steps: by then few particles
     * fly, mostly multi-hop ones whose next hop waits for the round's end (Ceres
     * 1e7: 203.6 against 210.5 ms per run with 4096 throughout, 258-282 ms with
     * 8192-20000 in the later rounds) */
    const int64_t sh_budget = sh.step_budget;
    const int64_t late_budget = sh.late_step_budget > 0 ? sh.late_step_budget : sh_budget;
    /* late_inline = k: the rounds after the first k run as launches of the kernel
     * with the landing service inline and no step budget, so the particles still
     * flying (most of them multi-hop, each hop of which would otherwise wait for
     * a round's end) finish in the first such launch; it holds fewer blocks per
     * CU than the lean flight kernel */
    const bool late_inline = !inline_svc && sh.late_inline && rd->E.mode == DSMC_MODE_PHYSICAL && !rd->E.advanced;
    const int split_rounds = sh.late_inline;
    int late_grid = flight_grid;
    if (late_inline) {
        const int occ = dsmc_flight_occupancy(&rd->E, 1, sh.block);
        if (occ > 0 && (int64_t)cus * occ < (int64_t)late_grid) late_grid = cus * occ;
    }
    /* per-launch timing events, destroyed on every return path */
    struct Events {
        std::vector<hipEvent_t> v;
        ~Events() { for (hipEvent_t e : v) (void)hipEventDestroy(e); }
    } evs;
    std::vector<hipEvent_t> &ev = evs.v;
    int p = 0, chunk = 4, issued = 0;
    for (;;) {
        for (int r = 0; r < chunk; r++) {
            if (timed) {
                hipEvent_t a, b;
                HIP_OK(hipEventCreate(&a));
                ev.push_back(a);
                HIP_OK(hipEventCreate(&b));
                ev.push_back(b);
                HIP_OK(hipEventRecord(a, st));
                is_lean.push_back(0);
            }
            if (inline_svc) {
                HIP_OK(dsmc_launch_phys(&rd->E, (const DsmcEnv *)rd->env_d.ptr, nullptr, rec, Q, p, flight_grid,
                                        sh.block, sh.relaunch_batch, sh_budget, st));
                if (timed) HIP_OK(hipEventRecord(ev.back(), st));
            } else if (late_inline && issued >= split_rounds) {
                HIP_OK(dsmc_launch_phys(&rd->E, (const DsmcEnv *)rd->env_d.ptr, nullptr, rec, Q, p, late_grid,
                                        sh.block, LATE_INLINE_SVC_BATCH, INT_MAX, st));
                if (timed) HIP_OK(hipEventRecord(ev.back(), st));
            } else {
                HIP_OK(dsmc_launch_flight(&rd->E, rec, Q, p, flight_grid, sh.block, sh.relaunch_batch,
                                          issued < 2 ? sh_budget : late_budget, sh.park_batch, st));
                if (timed) HIP_OK(hipEventRecord(ev.back(), st));
                if (timed) is_lean.back() = lean ? 1 : 0;
                /* the round before the late inline launch leaves its landings to
                 * it: they queue behind the paused particles, which start first */
                if (late_inline && issued == split_rounds - 1) HIP_OK(dsmc_launch_handover(Q, p, service_grid, st));
                else HIP_OK(dsmc_launch_service(&rd->E, rec, Q, p, service_grid, st));
            }
            p ^= 1;
            issued++;
        }
        HIP_OK(hipMemcpyAsync(rd->host_ctl, Q.ctl, sizeof(unsigned int) * 16, hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        if (rd->host_ctl[2 * p] == 0) break;
        if (rd->host_ctl[8] > (1u << 24)) {
            fprintf(stderr, "dsmc_hip: run did not finish within %u rounds\n", rd->host_ctl[8]);
            return DSMC_ERR_DEVICE;
        }
        if (chunk < 128) chunk *= 2;
    }
    /* physical runs with moments deposited into per-cell records: into vmom, vcount */
    HIP_OK(dsmc_launch_fold(&rd->E, (int64_t)rd->E.hd * rd->E.ftd * rd->E.ffd, st));
    rd->rounds = (int)rd->host_ctl[8];
    rd->lean_steps = (int64_t)((uint64_t)rd->host_ctl[12] | ((uint64_t)rd->host_ctl[13] << 32));
    for (size_t i = 0; i + 1 < ev.size(); i += 2) {
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
        rd->flight_ms += ms;
        if (is_lean[i / 2]) {
            rd->lean_ms += ms;
            rd->lean_launches++;
        }
    }
    return DSMC_OK;
}

dsmc_launch_shape default_shape(const DsmcEnv &E, int short_hops)
{
    dsmc_launch_shape s;
    s.block = 256;
    s.late_step_budget = 0;
    s.park_batch = 0;
    s.late_inline = 0;
    if (E.mode == 0 && !E.advanced) {
        /* compat ballistic particles make hundreds of hops, many of one step:
         * with the service in its own kernel every hop is a round of launches
         * and a run lasts as many rounds as its most-hopping particle has hops
         * (11900 at 1e4 particles, ~0.14 ms each).  With the service inline and
         * no step budget a run is one launch: 1e4 particles 254 ms against
         * 1634 ms, 1e6 708 ms against 5320 ms */
        s.blocks_per_cu = 4;
        s.relaunch_batch = 48;
        s.step_budget = INT_MAX;
        s.service_inline = 1;
    } else if (E.mode == 0) {
        /* compat advanced (few, long hops at dt 0.5 s): the split kernels */
        s.blocks_per_cu = 4;
        s.relaunch_batch = 8;
        s.step_budget = 4096;
        s.service_inline = 0;
    } else if (short_hops) {
        /* Moon / Mercury / Europa single runs of 1e4-2e6 particles on MI355X:
         * landing service inline, 48-lane batches, 2048-step budget 2.3e10
         * particle-steps/s; 32-lane batches 1.5e10; the split kernels 5e9 */
        s.blocks_per_cu = 4;
        s.relaunch_batch = 48;
        s.step_budget = 2048;
        s.service_inline = 1;
    } else if (E.nparticles < 4000000) {
        /* long hops (Ceres), runs below 4e6 particles: the landing service inline
         * in batches of 24 lanes, one launch (measured with the full-step loop:
         * 3e6 particles 82 ms against 87 ms with the split kernels, 1e6 44 against
         * 47, 3e4 21 against 24; with lean steps 1e6 takes 32 ms) */
        s.blocks_per_cu = 4;
        s.relaunch_batch = 24;
        s.step_budget = INT_MAX;
        s.service_inline = 1;
    } else {
        /* long hops (Ceres), 4e6 particles and more: the lean flight kernel for
         * the first two rounds (a 4096-step budget, 5 blocks per CU, refills of 8
         * idle lanes), then the rest of the run in one launch with the landing
         * service inline.  1e7 particles, flight kernels per run: 120.4 ms
         * (refills of 12 / 16 lanes: 122.2 / 124.3 ms); 125.8 ms with rounds of
         * the split kernels to the end (1024-step budgets after the second) */
        s.blocks_per_cu = 5;
        s.relaunch_batch = 8;
        s.step_budget = 4096;
        s.service_inline = 0;
        s.late_step_budget = 1024;
        s.late_inline = 2;
    }
    return s;
}

bool shape_ok(const dsmc_launch_shape &s)
{
    return s.block >= 64 && s.block <= 256 && s.block % 64 == 0 && s.blocks_per_cu > 0 &&
           s.blocks_per_cu * s.block <= 2048 && s.relaunch_batch >= 1 && s.relaunch_batch <= 64 &&
           s.step_budget >= 16 && (s.service_inline == 0 || s.service_inline == 1) &&
           (s.late_step_budget == 0 || s.late_step_budget >= 16) && s.park_batch >= 0 && s.park_batch <= 64 &&
           s.late_inline >= 0 && s.late_inline <= 64;
}

int collect_run(RunDev *rd, const dsmc_params *p, dsmc_result *r, hipStream_t st)
{
    const size_t cells = (size_t)p->hd * p->ftd * p->ffd;
    const int64_t n = p->nparticles;
    unsigned long long ctl[8], rtot = 0;
    HIP_OK(hipMemcpyAsync(ctl, rd->counters.ptr, sizeof ctl, hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(&rtot, rd->rec_total.ptr, sizeof rtot, hipMemcpyDeviceToHost, st));
    if (r->density) HIP_OK(hipMemcpyAsync(r->density, rd->density.ptr, cells * 8, hipMemcpyDeviceToHost, st));
    if (r->horizon)
        HIP_OK(hipMemcpyAsync(r->horizon, rd->horizon.ptr, sizeof(double) * p->ftd * p->ffd,
                              hipMemcpyDeviceToHost, st));
    if (r->vmom && rd->vmom.ptr) HIP_OK(hipMemcpyAsync(r->vmom, rd->vmom.ptr, cells * 24, hipMemcpyDeviceToHost, st));
    if (r->vcount && rd->vcount.ptr)
        HIP_OK(hipMemcpyAsync(r->vcount, rd->vcount.ptr, cells * 8, hipMemcpyDeviceToHost, st));
    /* physical runs: the density is the deposit count, staged in the caller's
     * density array (8 bytes per cell either way) when no count array is given */
    if (rd->E.density_from_count && r->density && !r->vcount)
        HIP_OK(hipMemcpyAsync(r->density, rd->vcount.ptr, cells * 8, hipMemcpyDeviceToHost, st));
    if (r->ttime) {
        if (rd->ttime.ptr)
            HIP_OK(hipMemcpyAsync(r->ttime, rd->ttime.ptr, rd->ttime.bytes, hipMemcpyDeviceToHost, st));
    }
    /* per-particle outputs (a run of no particles has none: no device buffer) */
    if (r->end_state && rd->end_state.ptr)
        HIP_OK(hipMemcpyAsync(r->end_state, rd->end_state.ptr, 48 * n, hipMemcpyDeviceToHost, st));
    if (r->steps && rd->steps.ptr) HIP_OK(hipMemcpyAsync(r->steps, rd->steps.ptr, 8 * n, hipMemcpyDeviceToHost, st));
    if (r->fate && rd->fate.ptr) HIP_OK(hipMemcpyAsync(r->fate, rd->fate.ptr, 4 * n, hipMemcpyDeviceToHost, st));
    if (r->hops && rd->hops.ptr) HIP_OK(hipMemcpyAsync(r->hops, rd->hops.ptr, 4 * n, hipMemcpyDeviceToHost, st));
    if (r->vhist && rd->vhist.ptr)
        HIP_OK(hipMemcpyAsync(r->vhist, rd->vhist.ptr, rd->vhist.bytes, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    if (rd->E.density_from_count && r->density) {
        for (size_t i = 0; i < cells; i++) {
            int64_t c;
            if (r->vcount) c = r->vcount[i];
            else memcpy(&c, &r->density[i], sizeof c);
            r->density[i] = (double)c;
        }
    }
    r->vrec_n = 0;
    if (rd->E.vrec) {
        int64_t nrec = (int64_t)rtot;
        if (nrec > p->record_cap_total) nrec = p->record_cap_total;
        size_t cap = (size_t)p->record_cap_total;
        if (nrec > 0) {
            HIP_OK(hipMemcpy(r->vrec, rd->vrec.ptr, nrec * 24, hipMemcpyDeviceToHost));
            const int32_t *base = (const int32_t *)rd->vrec_i.ptr;
            HIP_OK(hipMemcpy(r->vrec_td, base, nrec * 4, hipMemcpyDeviceToHost));
            HIP_OK(hipMemcpy(r->vrec_fd, base + cap, nrec * 4, hipMemcpyDeviceToHost));
            HIP_OK(hipMemcpy(r->vrec_ht, base + 2 * cap, nrec * 4, hipMemcpyDeviceToHost));
        }
        r->vrec_n = nrec;
    }
    /* physical runs have no transit times: their ttime buffer (5 rows, ballistic
     * or advanced) is zeroed; a compat run has a device buffer unless n == 0 */
    if (r->ttime && !rd->ttime.ptr) memset(r->ttime, 0, sizeof(double) * 5 * n);
    r->n_escaped = (int64_t)ctl[0];
    r->n_frozen = (int64_t)ctl[1];
    r->n_photolysed = (int64_t)ctl[2];
    r->n_depleted = (int64_t)ctl[3];
    r->n_capped = (int64_t)ctl[4];
    r->n_collisions = (int64_t)ctl[5];
    r->total_steps = (int64_t)ctl[6];
    return DSMC_OK;
}

}  // namespace dsmc

using namespace dsmc;

/* ------------------------------------------------------------------ presets */
extern "C" int dsmc_params_preset(dsmc_params *p, int preset, int ftd, int ffd)
{
    if (!p || ftd <= 0 || ffd <= 0) return DSMC_ERR_ARG;
    memset(p, 0, sizeof(*p));
    p->ftd = ftd;
    p->ffd = ffd;
    p->max_hop_steps = 10000000;            /* ballistic.c:188 */
    p->w_min = 0.01;                        /* ballistic.c:124 */
    p->density_scale = 1.0;
    p->sublimation_scale = 1.0;
    p->fit_degree = 3;                      /* types.h:89 POLY_DEGREE */
    p->max_per_cell = 50;                   /* advanced.c:169 */
    p->min_fit_samples = 8;                 /* advanced.c:44 */
    if (preset == DSMC_PRESET_COMPAT_BALLISTIC || preset == DSMC_PRESET_COMPAT_ADVANCED) {
        /* include/constants.h:5-35, ballistic.c:49-67, advanced.c:131-152 */
        p->mode = DSMC_MODE_COMPAT;
        p->G = 6.67e-11;
        p->M = 9.39e20;
        p->radius = 476.2e3;
        p->k_B = 1.38e-23;
        p->mmass = 18e-3;
        p->mole = 6.02e23;
        p->lifetime = 1.0 / 1.02e-5 * (2.8 * 2.8);
        p->period = 9.07417 * 3600.0;
        p->alpha = 0.5;
        p->v_esc = sqrt(2.0 * p->G * p->M / p->radius);
        p->t_freeze = 100.0;
        p->cross_section = 1e-19;
        if (preset == DSMC_PRESET_COMPAT_BALLISTIC) {
            p->stream = DSMC_STREAM_BALLISTIC;
            p->dt = 5.0;
            p->hd = 100;
            p->exp_grid = 0;
            p->unih = p->radius * 0.1;
            p->alpha2 = 0.0;
            p->max_height = p->unih * p->hd + p->radius;
        } else {
            p->stream = DSMC_STREAM_ADVANCED;
            p->dt = 0.5;
            p->hd = 50;
            p->exp_grid = 1;
            p->max_height = 3e6 * sqrt(2.0);
            p->alpha2 = 0.04;
            p->unih = (p->max_height - p->radius) / (pow(1.0 + p->alpha2, p->hd) - 1.0) * p->alpha2;
        }
        p->record_cap_cell = 200;                               /* ballistic.c:216 */
        p->record_cap_total = (int64_t)100 * ftd * p->hd;       /* ballistic.c:77 */
        return DSMC_OK;
    }
    if (preset == DSMC_PRESET_CERES_PHYSICAL || preset == DSMC_PRESET_CERES_PHYSICAL_COLL) {
        p->mode = DSMC_MODE_PHYSICAL;
        p->stream = DSMC_STREAM_BALLISTIC;
        p->G = 1.0;
        p->M = 6.26e10;                      /* GM of Ceres [m^3/s^2] */
        p->radius = 470e3;
        p->k_B = 1.380649e-23;
        p->mmass = 2.99e-26;                 /* H2O molecule [kg] */
        p->mole = 1.0;
        p->lifetime = 6.4e5;
        p->period = 9.07417 * 3600.0;
        p->alpha = 1.0;
        p->v_esc = sqrt(2.0 * p->G * p->M / p->radius);
        p->dt = preset == DSMC_PRESET_CERES_PHYSICAL ? 10.0 : 5.0;
        p->max_hop_steps = 0;
        p->max_steps = 20000;
        p->t_freeze = 110.0;
        p->w_min = 0.0;
        p->subsolar_lon = 0.0;
        p->hd = 50;
        p->exp_grid = 1;
        p->alpha2 = 0.04;
        p->max_height = 20.0 * p->radius;
        p->unih = (p->max_height - p->radius) / (pow(1.0 + p->alpha2, p->hd) - 1.0) * p->alpha2;
        p->record_cap_cell = 0;
        p->record_cap_total = 0;
        p->cross_section = M_PI * 4.6e-10 * 4.6e-10;
        p->fit_degree = 6;
        p->vhist_bins = 64;                  /* phase-1 speed histograms: 31.25 m/s bins */
        p->vhist_max = 2000.0;
        if (preset == DSMC_PRESET_CERES_PHYSICAL_COLL) {
            /* ~6 kg/s of water: the order of Ceres' measured outgassing, x10 */
            p->production_rate = 2.0e26;
            p->sublimation_scale = 10.0;
            p->stream = DSMC_STREAM_ADVANCED;   /* phase 2 draws apart from phase 1 */
        }
        return DSMC_OK;
    }
    return DSMC_ERR_ARG;
}

/* ------------------------------------------------------------------ single runs */
extern "C" int dsmc_run_ballistic(const dsmc_params *p, const dsmc_surface *s, dsmc_result *r)
{
    if (!p || !s || !r || p->nparticles < 0) return DSMC_ERR_ARG;
    int dev_rc = ensure_device();
    if (dev_rc) return dev_rc;
    TimedStream ts;
    if (ts.create()) return DSMC_ERR_DEVICE;
    hipStream_t st = ts.st;
    hipEvent_t e0 = ts.e0, e1 = ts.e1;
    RunDev rd;
    int rc = stage_run(p, s, r, 0, &rd, st);
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

/* ------------------------------------------------------------------ allocation-free runs */
/* Every device buffer of dsmc_run_ballistic_noalloc comes from this block (one
 * copy per device, from module load until the process ends).  Host code refers
 * to it only through the address hipGetSymbolAddress returns for the current
 * device: the array's name is not a device address on the host. */
#define DSMC_NOALLOC_BYTES ((size_t)96 << 20)
static __device__ __align__(256) unsigned char scratch[DSMC_NOALLOC_BYTES];

enum { NOALLOC_MAX_DEVICES = 64 };
static std::mutex noalloc_mu;                      /* the process's calls take turns */
static void *noalloc_base[NOALLOC_MAX_DEVICES];    /* scratch's address on each device */

/* scratch's address on the current device, looked up on the device's first call. */
static int noalloc_scratch(void **out)
{
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return DSMC_ERR_DEVICE;
    if (!noalloc_base[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fprintf(stderr, "dsmc_hip: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
            return DSMC_ERR_DEVICE;
        }
        noalloc_base[dev] = p;
    }
    *out = noalloc_base[dev];
    return DSMC_OK;
}

extern "C" int dsmc_run_ballistic_noalloc(const dsmc_params *p, const dsmc_surface *s, dsmc_result *r)
{
    if (!p || !s || !r || p->nparticles < 0) return DSMC_ERR_ARG;
    int rc = ensure_device();
    if (rc) return rc;
    std::lock_guard<std::mutex> g(noalloc_mu);
    void *base = nullptr;
    if ((rc = noalloc_scratch(&base))) return rc;
    /* one RunDev for every call: its host tables keep their storage, its device
     * buffers are borrowed from the arena, its queue readback is pageable */
    static RunDev rd;
    rd.drop_buffers();                            /* nothing of the previous call's run carries over */
    Arena ar;
    ar.base = (char *)base;
    ar.cap = DSMC_NOALLOC_BYTES;
    rd.arena = &ar;
    const hipStream_t st = nullptr;               /* the null stream: no stream or event is created */
    const auto t0 = std::chrono::steady_clock::now();
    rc = stage_run(p, s, r, 0, &rd, st);
    if (rc == DSMC_ERR_BUDGET)
        fprintf(stderr, "dsmc_hip: dsmc_run_ballistic_noalloc: a run of %lld particles and %d x %d x %d cells needs "
                        "more than the %zu MiB of preallocated device memory\n", (long long)p->nparticles, p->hd,
                p->ftd, p->ffd, DSMC_NOALLOC_BYTES >> 20);
    if (!rc) rc = launch_run(&rd, nullptr, st);
    if (!rc) rc = collect_run(&rd, p, r, st);
    if (!rc) r->elapsed_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    rd.drop_buffers();
    rd.arena = nullptr;
    return rc;
}

/* ------------------------------------------------------------------ staged runs */
struct dsmc_run {
    dsmc_params p;
    RunDev rd;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int launched = 0;
};

extern "C" dsmc_run *dsmc_run_create(const dsmc_params *p, const dsmc_surface *s,
                                     const dsmc_result *outputs, int *status)
{
    int rc = (!p || !s || !outputs || p->nparticles < 0) ? DSMC_ERR_ARG : ensure_device();
    if (rc) { if (status) *status = rc; return nullptr; }
    dsmc_run *run = new dsmc_run();
    run->p = *p;
    if (hipStreamCreateWithFlags(&run->st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&run->e0) != hipSuccess || hipEventCreate(&run->e1) != hipSuccess) {
        rc = DSMC_ERR_DEVICE;
    } else {
        rc = stage_run(p, s, outputs, 0, &run->rd, run->st);
        if (!rc && hipStreamSynchronize(run->st) != hipSuccess) rc = DSMC_ERR_DEVICE;
    }
    if (rc) { dsmc_run_destroy(run); run = nullptr; }
    if (status) *status = rc;
    return run;
}

/* Clears every output buffer of a staged run on its stream. */
static int clear_outputs(RunDev &rd, hipStream_t st)
{
    DevBuf *bufs[] = {&rd.density, &rd.horizon, &rd.vmom, &rd.vcount, &rd.cellacc, &rd.ttime, &rd.end_state,
                      &rd.steps, &rd.fate, &rd.hops, &rd.count_nn, &rd.counters, &rd.rec_total, &rd.vhist};
    for (DevBuf *b : bufs)
        if (b->ptr) HIP_OK(hipMemsetAsync(b->ptr, 0, b->bytes, st));
    return DSMC_OK;
}

extern "C" int dsmc_run_launch(dsmc_run *run, const dsmc_launch_shape *shape)
{
    if (!run) return DSMC_ERR_ARG;
    int rc = clear_outputs(run->rd, run->st);
    if (rc) return rc;
    HIP_OK(hipEventRecord(run->e0, run->st));
    rc = launch_run(&run->rd, shape, run->st, 1);
    if (rc) return rc;
    HIP_OK(hipEventRecord(run->e1, run->st));
    run->launched = 1;
    return DSMC_OK;
}

extern "C" int dsmc_run_sync(dsmc_run *run)
{
    if (!run) return DSMC_ERR_ARG;
    HIP_OK(hipStreamSynchronize(run->st));
    return DSMC_OK;
}

extern "C" double dsmc_run_kernel_ms(dsmc_run *run)
{
    if (!run || !run->launched) return -1.0;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, run->e0, run->e1) != hipSuccess) return -1.0;
    return ms;
}

extern "C" int dsmc_run_timing(dsmc_run *run, double *flight_ms, int *rounds)
{
    if (!run || !run->launched) return DSMC_ERR_ARG;
    if (flight_ms) *flight_ms = run->rd.flight_ms;
    if (rounds) *rounds = run->rd.rounds;
    return DSMC_OK;
}

extern "C" int dsmc_run_lean_timing(dsmc_run *run, double *lean_ms, int *launches, int64_t *steps)
{
    if (!run || !run->launched) return DSMC_ERR_ARG;
    if (lean_ms) *lean_ms = run->rd.lean_ms;
    if (launches) *launches = run->rd.lean_launches;
    if (steps) *steps = run->rd.lean_steps;
    return DSMC_OK;
}

extern "C" int dsmc_run_collect(dsmc_run *run, dsmc_result *r)
{
    if (!run || !r) return DSMC_ERR_ARG;
    int rc = collect_run(&run->rd, &run->p, r, run->st);
    if (!rc) r->elapsed_ms = dsmc_run_kernel_ms(run);
    return rc;
}

extern "C" void dsmc_run_destroy(dsmc_run *run)
{
    if (!run) return;
    if (run->st) (void)hipStreamSynchronize(run->st);
    if (run->e0) (void)hipEventDestroy(run->e0);
    if (run->e1) (void)hipEventDestroy(run->e1);
    hipStream_t st = run->st;
    delete run;
    if (st) (void)hipStreamDestroy(st);
}

/* ------------------------------------------------------------------ diagnostics */
extern "C" int dsmc_verlet_steps(double *state, int n, int nsteps, double GM, double R,
                                 double max_height, double h, double *traj)
{
    if (!state || n <= 0 || nsteps < 0) return DSMC_ERR_ARG;
    int rc = ensure_device();
    if (rc) return rc;
    DevBuf ds, dt;
    if ((rc = ds.upload(state, sizeof(double) * 9 * n, 0))) return rc;
    if (traj && (rc = dt.alloc(sizeof(double) * 7 * (size_t)n * nsteps))) return rc;
    HIP_OK(dsmc_launch_verlet((double *)ds.ptr, n, nsteps, GM, R, max_height, h,
                              traj ? (double *)dt.ptr : nullptr, 0));
    HIP_OK(hipMemcpy(state, ds.ptr, sizeof(double) * 9 * n, hipMemcpyDeviceToHost));
    if (traj) HIP_OK(hipMemcpy(traj, dt.ptr, dt.bytes, hipMemcpyDeviceToHost));
    return DSMC_OK;
}

extern "C" int dsmc_elementary(int fn, const double *x, const double *y, double *out, int n)
{
    if (!x || !out || n <= 0) return DSMC_ERR_ARG;
    int rc = ensure_device();
    if (rc) return rc;
    DevBuf dx, dy, dout;
    if ((rc = dx.upload(x, 8 * (size_t)n, 0))) return rc;
    if (y && (rc = dy.upload(y, 8 * (size_t)n, 0))) return rc;
    if ((rc = dout.alloc(8 * (size_t)n))) return rc;
    HIP_OK(dsmc_launch_math(fn, (const double *)dx.ptr, y ? (const double *)dy.ptr : nullptr,
                            (double *)dout.ptr, n, 0));
    HIP_OK(hipMemcpy(out, dout.ptr, 8 * (size_t)n, hipMemcpyDeviceToHost));
    return DSMC_OK;
}

extern "C" const char *dsmc_version(void) { return "dsmc_hip 0.1 (gfx950)"; }

extern "C" int dsmc_device_info(char *buf, size_t len)
{
    if (!buf || len == 0) return DSMC_ERR_ARG;
    int rc = ensure_device();
    if (rc) return rc;
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIP_OK(hipGetDeviceProperties(&prop, dev));
    snprintf(buf, len, "%s, %d CUs, %.1f GB HBM, clock %d kHz", prop.gcnArchName,
             prop.multiProcessorCount, prop.totalGlobalMem / 1e9, prop.clockRate);
    return DSMC_OK;
}

namespace dsmc {
int ensure_device()
{
    static int state = 0;   /* 0 unknown, 1 ok, <0 error */
    static std::mutex mu;
    std::lock_guard<std::mutex> g(mu);
    if (state == 0) {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
            fprintf(stderr, "dsmc_hip: no HIP device available\n");
            state = DSMC_ERR_DEVICE;
        } else {
            const char *d = getenv("DSMC_DEVICE");
            int dev = d ? atoi(d) : 0;
            if (dev < 0 || dev >= n || hipSetDevice(dev) != hipSuccess) state = DSMC_ERR_DEVICE;
            else state = 1;
        }
    }
    return state == 1 ? DSMC_OK : state;
}
}  // namespace dsmc
