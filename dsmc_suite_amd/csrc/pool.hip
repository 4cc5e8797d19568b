/*
 * pool.hip - multi-run pools on one device.
 *
 * A pool owns max_slots particle records (256 B each) and three slot queues
 * shared by all its runs, max_runs grid slots (density counts, velocity moments,
 * horizon, counters), per-run launch tables (CDF, face temperatures) and a few
 * shared tables (latitude thresholds and sector directions of the pool's grid,
 * up to POOL_MAX_HR_TABS radial threshold tables - one per body - and
 * POOL_MAX_GEO icosphere geometries).  Nothing is allocated or freed while the
 * pool runs: admission sub-allocates slot ranges and grid slots, so runs can be
 * admitted while others are in flight (hipFree would wait for the device).
 *
 * Rounds of the persistent flight / service kernels are captured once in a
 * hipGraph (graph_rounds rounds, an even number, so that every replay starts on
 * fly queue 0) and replayed.  A newly admitted run's start kernel appends its
 * particles to fly queue 0 between two replays.  After each replay the run
 * counters are copied to pinned memory; a run is finished when its count of
 * finished particles equals its particle count.  Two replays are kept in flight
 * so the device never waits for the host's bookkeeping.
 */
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <iterator>
#include <math.h>
#include <stdint.h>

#include "pool.h"
#include "dsmc_kernels.h"

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

static const int POOL_RING = 2;

size_t pool_slot_bytes(int per_particle)
{
    return sizeof(DsmcRec) + 5 * sizeof(int32_t) + (per_particle ? 6 * 8 + 8 + 4 + 4 : 0);
}

size_t pool_run_bytes(const PoolClass &c)
{
    size_t cells = (size_t)c.ftd * c.ffd * c.hd;
    size_t nfine = (size_t)20 << (2 * c.levels);
    return cells * (8 + 32) + (size_t)c.ftd * c.ffd * 8 + 64 + nfine * 16 + sizeof(DsmcEnv) +
           sizeof(DsmcStepEnv);
}

int pool_class_of(const dsmc_params *p, const dsmc_surface *s, PoolClass *c)
{
    if (!p || !s || p->mode != DSMC_MODE_PHYSICAL) return DSMC_ERR_ARG;
    c->mode = p->mode;
    c->adv = 0;
    c->ftd = p->ftd;
    c->ffd = p->ffd;
    c->hd = p->hd;
    c->levels = s->levels;
    return DSMC_OK;
}

static bool same_class(const PoolClass &a, const PoolClass &b)
{
    return a.mode == b.mode && a.adv == b.adv && a.ftd == b.ftd && a.ffd == b.ffd && a.hd == b.hd &&
           a.levels == b.levels;
}

Pool::~Pool()
{
    if (st) (void)hipStreamSynchronize(st);
    if (cst) (void)hipStreamSynchronize(cst);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (envs_h) (void)hipHostFree(envs_h);
    if (steps_h) (void)hipHostFree(steps_h);
    if (stage_h) (void)hipHostFree(stage_h);
    if (cnt_h) (void)hipHostFree(cnt_h);
    if (ctl_h) (void)hipHostFree(ctl_h);
    if (out_h) (void)hipHostFree(out_h);
    for (auto &e : ring_ev) if (e) (void)hipEventDestroy(e);
    rec.release();
    if (st) (void)hipStreamDestroy(st);
    if (cst) (void)hipStreamDestroy(cst);
}

static uint64_t fnv(const void *data, size_t n, uint64_t h = 1469598103934665603ull)
{
    const unsigned char *b = (const unsigned char *)data;
    for (size_t i = 0; i < n; i++) { h ^= b[i]; h *= 1099511628211ull; }
    return h;
}

/* Slot queues of ballistic (kind 0) or collisional (kind 1) particles. */
static DsmcQueues queues(Pool *P, int kind = 0)
{
    DsmcQueues Q;
    Q.fly[0] = (int32_t *)(kind ? P->cflyq0.ptr : P->flyq0.ptr);
    Q.fly[1] = (int32_t *)(kind ? P->cflyq1.ptr : P->flyq1.ptr);
    Q.svc = (int32_t *)P->svcq.ptr;
    Q.ctl = (unsigned int *)(kind ? P->cqctl.ptr : P->qctl.ptr);
    return Q;
}

static DsmcPoolDev pool_dev(Pool *P)
{
    DsmcPoolDev D;
    D.rec = (DsmcRec *)P->rec.ptr;
    D.envs = (const DsmcEnv *)P->envs_d.ptr;
    D.steps = (const DsmcStepEnv *)P->steps_d.ptr;
    D.tq = (const double *)P->tq.ptr;
    D.cs = (const double *)P->cs.ptr;
    D.hr = (const double *)P->hr.ptr;
    D.ftd = P->cls.ftd;
    D.ffd = P->cls.ffd;
    D.hd = P->cls.hd;
    D.nhr = POOL_MAX_HR_TABS * (P->cls.hd + 2);
    D.mode = P->cls.mode;
    D.adv = P->cls.adv;
    /* the run table in LDS when it is small (a 64-run batch: 12.8 KB per block) */
    D.steps_lds = (size_t)P->max_runs * sizeof(DsmcStepEnv) <= POOL_STEPS_LDS_BYTES ? P->max_runs : 0;
    return D;
}

int pool_occupancy(Pool *P, int block)
{
    const DsmcPoolDev D = pool_dev(P);
    return dsmc_pool_occupancy(&D, block);
}

/* One round sequence of graph_rounds rounds on stream s. */
static int launch_rounds(Pool *P, hipStream_t s)
{
    DsmcPoolDev D = pool_dev(P);
    DsmcQueues Q = queues(P);
    int fgrid = P->cus * P->shape.blocks_per_cu;
    int sgrid = P->cus * 8;
    DsmcPoolDev D1 = D;
    D1.adv = 1;
    DsmcQueues Q1 = queues(P, 1);
    for (int r = 0; r < P->graph_rounds; r++) {
        int p = r & 1;
        if (P->hist_bins) {        /* collisional particles: their own queues and kernel */
            HIP_OK(dsmc_launch_phys(nullptr, nullptr, &D1, D1.rec, Q1, p, fgrid, P->shape.block,
                                    P->shape.relaunch_batch, P->late_unbounded && r > 0 ? 0 : P->shape.step_budget, s));
        }
        const int64_t budget = P->late_unbounded && r > 0 ? 0 : P->shape.step_budget;
        if (P->shape.service_inline) {
            HIP_OK(dsmc_launch_phys(nullptr, nullptr, &D, D.rec, Q, p, fgrid, P->shape.block,
                                    P->shape.relaunch_batch, budget, s));
        } else {
            HIP_OK(dsmc_launch_pool_flight(&D, Q, p, fgrid, P->shape.block, P->shape.relaunch_batch, budget, s));
            HIP_OK(dsmc_launch_pool_service(&D, Q, p, sgrid, s));
        }
    }
    return DSMC_OK;
}

static int capture(Pool *P)
{
    if (P->st) HIP_OK(hipStreamSynchronize(P->st));   /* no replay of the old graph in flight */
    if (P->exec) { (void)hipGraphExecDestroy(P->exec); P->exec = nullptr; }
    if (!P->use_graph) return DSMC_OK;
    hipStream_t cap;
    HIP_OK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    hipGraph_t g = nullptr;
    int rc = DSMC_OK;
    if (hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) != hipSuccess) rc = DSMC_ERR_DEVICE;
    if (!rc) {
        rc = launch_rounds(P, cap);
        hipError_t e = hipStreamEndCapture(cap, &g);
        if (!rc && e != hipSuccess) rc = DSMC_ERR_DEVICE;
    }
    if (!rc && hipGraphInstantiate(&P->exec, g, nullptr, nullptr, 0) != hipSuccess) {
        P->exec = nullptr;
        rc = DSMC_ERR_DEVICE;
    }
    if (g) (void)hipGraphDestroy(g);
    (void)hipStreamDestroy(cap);
    if (rc) fprintf(stderr, "dsmc_hip: hipGraph capture of the pool rounds failed\n");
    return rc;
}

int pool_set_shape(Pool *P, const dsmc_launch_shape &s)
{
    if (!shape_ok(s)) return DSMC_ERR_ARG;
    P->shape = s;
    return capture(P);
}

int pool_create(Pool *P, const PoolClass &c, int max_runs, int64_t max_slots, int per_particle,
                const dsmc_launch_shape *shape)
{
    if (max_runs <= 0 || max_slots <= 0 || max_slots > INT32_MAX || c.mode != DSMC_MODE_PHYSICAL || c.adv ||
        c.levels < 0 || c.levels > 8)
        return DSMC_ERR_ARG;
    int rc = ensure_device();
    if (rc) return rc;
    if ((rc = device_props(&P->cus))) return rc;
    P->cls = c;
    P->max_runs = max_runs;
    P->max_slots = max_slots;
    P->per_particle = per_particle;
    P->cells = (size_t)c.ftd * c.ffd * c.hd;
    P->nfine = (size_t)20 << (2 * c.levels);
    P->nvert = (size_t)10 * ((size_t)1 << (2 * c.levels)) + 2;
    P->nface_all = 0;
    for (int l = 0; l <= c.levels; l++) P->nface_all += (size_t)20 << (2 * l);
    HIP_OK(hipStreamCreateWithFlags(&P->st, hipStreamNonBlocking));
    HIP_OK(hipStreamCreateWithFlags(&P->cst, hipStreamNonBlocking));
    for (auto &e : P->ring_ev) HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    hipStream_t st = P->st;
    const size_t ms = (size_t)max_slots, mr = (size_t)max_runs;
    if ((rc = P->rec.alloc(ms * sizeof(DsmcRec)))) return rc;
    if ((rc = P->flyq0.alloc(ms * 4)) || (rc = P->flyq1.alloc(ms * 4)) || (rc = P->svcq.alloc(ms * 4))) return rc;
    if ((rc = P->cflyq0.alloc(ms * 4)) || (rc = P->cflyq1.alloc(ms * 4))) return rc;
    if ((rc = P->qctl.zeros(16 * sizeof(unsigned int), st))) return rc;
    if ((rc = P->cqctl.zeros(16 * sizeof(unsigned int), st))) return rc;
    if ((rc = P->envs_d.zeros(mr * sizeof(DsmcEnv), st))) return rc;
    if ((rc = P->steps_d.zeros(mr * sizeof(DsmcStepEnv), st))) return rc;
    if ((rc = P->hr.zeros((size_t)POOL_MAX_HR_TABS * (c.hd + 2) * 8, st))) return rc;
    if ((rc = P->geo_vert.alloc(POOL_MAX_GEO * P->nvert * 3 * 8))) return rc;
    if ((rc = P->geo_faces.alloc(POOL_MAX_GEO * P->nface_all * 3 * 4))) return rc;
    if ((rc = P->geo_enorm.alloc(POOL_MAX_GEO * P->nface_all * 9 * 8))) return rc;
    if ((rc = P->geo_fnorm.alloc(POOL_MAX_GEO * P->nfine * 9 * 8))) return rc;
    P->fmap_g = face_map_grid((int)P->nfine);
    const size_t fmap_n = (size_t)6 * P->fmap_g * P->fmap_g * DSMC_FMAP_K;
    if ((rc = P->geo_fmap.alloc(POOL_MAX_GEO * fmap_n * 4))) return rc;
    if ((rc = P->geo_fmapn.alloc(POOL_MAX_GEO * fmap_n / DSMC_FMAP_K * 36 * 8))) return rc;
    if ((rc = P->cdf.alloc(mr * P->nfine * 8)) || (rc = P->face_T.alloc(mr * P->nfine * 8)) ||
        (rc = P->face_sig.alloc(mr * P->nfine * 8)))
        return rc;
    if ((rc = P->vcount.alloc(mr * P->cells * 8)) || (rc = P->cellacc.alloc(mr * P->cells * 32))) return rc;
    if ((rc = P->horizon.alloc(mr * c.ftd * c.ffd * 8)) || (rc = P->counters.zeros(mr * 64, st))) return rc;
    if (per_particle) {
        if ((rc = P->end_state.alloc(ms * 48)) || (rc = P->steps_o.alloc(ms * 8)) || (rc = P->fate.alloc(ms * 4)) ||
            (rc = P->hops.alloc(ms * 4)))
            return rc;
    }
    /* shared latitude / sector tables: from a run environment of the pool's grid */
    {
        std::vector<double> tq(c.ftd + 1), cs(2 * (c.ffd + 1));
        tq[0] = -INFINITY;
        tq[c.ftd] = INFINITY;
        for (int k = 1; k < c.ftd; k++) {
            int ftd = c.ftd;
            double t = first_ge([ftd](double q) { return td_of_q(q, ftd); }, k, -1.0, 1.0);
            if (isnan(t)) return DSMC_ERR_ARG;
            tq[k] = t;
        }
        for (int k = 0; k <= c.ffd; k++) {
            double phi = 2.0 * M_PI * k / c.ffd;
            cs[2 * k] = cos(phi);
            cs[2 * k + 1] = sin(phi);
        }
        if ((rc = P->tq.upload(tq.data(), tq.size() * 8, st)) || (rc = P->cs.upload(cs.data(), cs.size() * 8, st)))
            return rc;
        HIP_OK(hipStreamSynchronize(st));
    }
    if (hipHostMalloc((void **)&P->envs_h, mr * sizeof(DsmcEnv), 0) != hipSuccess ||
        hipHostMalloc((void **)&P->steps_h, mr * sizeof(DsmcStepEnv), 0) != hipSuccess ||
        hipHostMalloc((void **)&P->stage_h, mr * 3 * P->nfine * 8, 0) != hipSuccess ||
        hipHostMalloc((void **)&P->cnt_h, POOL_RING * mr * 64, 0) != hipSuccess ||
        hipHostMalloc((void **)&P->ctl_h, POOL_RING * 16 * sizeof(unsigned int), 0) != hipSuccess)
        return DSMC_ERR_NOMEM;
    memset(P->cnt_h, 0, POOL_RING * mr * 64);
    P->out_bytes = P->cells * 40 + (size_t)c.ftd * c.ffd * 8 + 64;
    if (hipHostMalloc(&P->out_h, P->out_bytes, 0) != hipSuccess) return DSMC_ERR_NOMEM;
    P->runs.assign(mr, PoolRun());
    P->free_slots.clear();
    P->free_slots[0] = max_slots;
    P->hr_tabs.assign(POOL_MAX_HR_TABS, std::vector<double>());
    P->hr_ref.assign(POOL_MAX_HR_TABS, 0);
    P->geos.assign(POOL_MAX_GEO, PoolGeo());
    P->inflight = 0;
    P->rounds = 0;
    dsmc_launch_shape sh;
    sh.block = 256;
    sh.blocks_per_cu = 4;
    sh.relaunch_batch = 48;     /* lanes serviced together with the landing service inline */
    sh.step_budget = 1024;
    sh.service_inline = 1;
    sh.late_step_budget = 0;
    sh.park_batch = 0;
    sh.late_inline = 0;
    return pool_set_shape(P, shape ? *shape : sh);
}

/* First-fit slot range. */
static int64_t take_slots(Pool *P, int64_t n)
{
    for (auto it = P->free_slots.begin(); it != P->free_slots.end(); ++it) {
        if (it->second >= n) {
            int64_t s0 = it->first, len = it->second;
            P->free_slots.erase(it);
            if (len > n) P->free_slots[s0 + n] = len - n;
            return s0;
        }
    }
    return -1;
}

static void give_slots(Pool *P, int64_t s0, int64_t n)
{
    if (n <= 0) return;
    auto next = P->free_slots.lower_bound(s0);
    if (next != P->free_slots.end() && s0 + n == next->first) {
        n += next->second;
        P->free_slots.erase(next);
    }
    auto it = P->free_slots.lower_bound(s0);
    if (it != P->free_slots.begin()) {
        auto prev = std::prev(it);
        if (prev->first + prev->second == s0) { prev->second += n; return; }
    }
    P->free_slots[s0] = n;
}

int pool_take_slot(std::vector<int> &used)
{
    for (size_t i = 0; i < used.size(); i++)
        if (!used[i]) { used[i] = 1; return (int)i; }
    return -1;
}

int pool_enable_collisional(Pool *P, int bins, int degree)
{
    if (bins <= 0 || bins > 64 || degree < 1 || degree > 6) return DSMC_ERR_ARG;
    if (P->hist_bins) return (bins == P->hist_bins && degree == P->fit_deg) ? DSMC_OK : DSMC_ERR_ARG;
    const size_t c = P->cells;
    int rc;
    if ((rc = P->hist.alloc((size_t)POOL_MAX_HIST * c * bins * 8))) return rc;
    if ((rc = P->coll_cf.alloc((size_t)POOL_MAX_COLL * c * (degree + 1) * 8)) ||
        (rc = P->coll_lo.alloc((size_t)POOL_MAX_COLL * c * 8)) || (rc = P->coll_hi.alloc((size_t)POOL_MAX_COLL * c * 8)) ||
        (rc = P->coll_nsig.alloc((size_t)POOL_MAX_COLL * c * 8)) ||
        (rc = P->coll_valid.alloc((size_t)POOL_MAX_COLL * c * 4)) || (rc = P->coll_cnt.alloc((size_t)POOL_MAX_COLL * c * 4)))
        return rc;
    P->hist_used.assign(POOL_MAX_HIST, 0);
    P->coll_used.assign(POOL_MAX_COLL, 0);
    P->hist_bins = bins;
    P->fit_deg = degree;
    return capture(P);          /* the rounds now include the collisional kernel */
}

int pool_build_fits(Pool *P, int hist, int coll, const double *ndens, const dsmc_params *p2)
{
    if (!P->hist_bins || hist < 0 || hist >= POOL_MAX_HIST || coll < 0 || coll >= POOL_MAX_COLL ||
        p2->vhist_bins != P->hist_bins || p2->fit_degree != P->fit_deg)
        return DSMC_ERR_ARG;
    const size_t c = P->cells;
    const int deg = P->fit_deg;
    double *nsig = (double *)P->coll_nsig.ptr + (size_t)coll * c;
    int32_t *valid = (int32_t *)P->coll_valid.ptr + (size_t)coll * c;
    /* the slot is not read by queued work (its run is admitted afterwards) */
    HIP_OK(hipMemcpy(nsig, ndens, c * 8, hipMemcpyHostToDevice));
    HIP_OK(dsmc_launch_hist_fits((const unsigned long long *)P->hist.ptr + (size_t)hist * c * P->hist_bins, (int)c,
                                 P->hist_bins, p2->vhist_max / p2->vhist_bins, deg, p2->min_fit_samples,
                                 (double *)P->coll_cf.ptr + (size_t)coll * c * (deg + 1),
                                 (double *)P->coll_lo.ptr + (size_t)coll * c, (double *)P->coll_hi.ptr + (size_t)coll * c,
                                 valid, (int32_t *)P->coll_cnt.ptr + (size_t)coll * c, P->st));
    HIP_OK(dsmc_launch_nsig(nsig, valid, p2->cross_section, (int)c, nsig, P->st));
    return DSMC_OK;
}

int pool_prepare(const dsmc_params *p, const dsmc_surface *s, PoolPrep *out, int collisional)
{
    PoolClass c;
    int rc = pool_class_of(p, s, &c);
    if (rc) return rc;
    if (p->nparticles <= 0 || p->nparticles > INT32_MAX) return DSMC_ERR_ARG;
    if ((rc = build_env(p, s, collisional ? 1 : 0, &out->E, &out->T))) return rc;
    out->kind = collisional ? 1 : 0;
    out->p = *p;
    const size_t nf = (size_t)out->T.nfine, nfa = (size_t)out->T.nface_all;
    out->face_T.assign(s->face_T, s->face_T + nf);
    out->vert.assign(s->vert, s->vert + (size_t)s->nvert * 3);
    out->faces.assign(s->faces, s->faces + nfa * 3);
    uint64_t key = fnv(out->vert.data(), out->vert.size() * 8);
    out->geo_key = fnv(out->faces.data(), out->faces.size() * 4, key);
    return DSMC_OK;
}

int pool_admit(Pool *P, const PoolPrep &pr, uint64_t tag, int *g_out, int hist, int coll)
{
    const dsmc_params *p = &pr.p;
    const HostTables &T = pr.T;
    if (p->mode != P->cls.mode || p->ftd != P->cls.ftd || p->ffd != P->cls.ffd || p->hd != P->cls.hd ||
        T.levels != P->cls.levels || pr.vert.size() != P->nvert * 3)
        return DSMC_ERR_ARG;
    if ((pr.kind == 1) != (coll >= 0) || (hist >= 0 && pr.kind == 1)) return DSMC_ERR_ARG;
    if ((hist >= 0 || coll >= 0) && (!P->hist_bins || hist >= POOL_MAX_HIST || coll >= POOL_MAX_COLL ||
                                     p->vhist_bins != P->hist_bins || p->fit_degree != P->fit_deg))
        return DSMC_ERR_ARG;
    if (p->nparticles > P->max_slots) return DSMC_ERR_BUDGET;
    int g = -1;
    for (int i = 0; i < P->max_runs; i++) if (!P->runs[i].used) { g = i; break; }
    if (g < 0) return DSMC_ERR_BUSY;
    PoolRun &R = P->runs[g];
    /* radial table of this body */
    int tab = -1;
    for (int i = 0; i < POOL_MAX_HR_TABS; i++)
        if (P->hr_ref[i] > 0 && P->hr_tabs[i] == T.hr) { tab = i; break; }
    bool new_tab = false;
    if (tab < 0) {
        for (int i = 0; i < POOL_MAX_HR_TABS; i++) if (P->hr_ref[i] == 0) { tab = i; break; }
        if (tab < 0) return DSMC_ERR_BUSY;
        new_tab = true;
    }
    /* geometry */
    const uint64_t key = pr.geo_key;
    int geo = -1;
    for (int i = 0; i < POOL_MAX_GEO; i++)
        if (P->geos[i].valid && P->geos[i].key == key) { geo = i; break; }
    bool new_geo = false;
    if (geo < 0) {
        for (int i = 0; i < POOL_MAX_GEO; i++) if (P->geos[i].ref == 0) { geo = i; break; }
        if (geo < 0) return DSMC_ERR_BUSY;
        new_geo = true;
    }
    int64_t s0 = take_slots(P, p->nparticles);
    if (s0 < 0) return DSMC_ERR_BUSY;
    R.E = pr.E;

    hipStream_t st = P->st;
    const int hd2 = P->cls.hd + 2;
    double *hr_d = (double *)P->hr.ptr + (size_t)tab * hd2;
    if (new_tab) {
        /* a table slot no queued work reads; it is read by every flight launch,
         * so it is written before the next one is enqueued */
        P->hr_tabs[tab] = T.hr;
        HIP_OK(hipMemcpy(hr_d, P->hr_tabs[tab].data(), hd2 * 8, hipMemcpyHostToDevice));
    }
    double *vert_d = (double *)P->geo_vert.ptr + (size_t)geo * P->nvert * 3;
    int32_t *faces_d = (int32_t *)P->geo_faces.ptr + (size_t)geo * P->nface_all * 3;
    double *enorm_d = (double *)P->geo_enorm.ptr + (size_t)geo * P->nface_all * 9;
    double *fnorm_d = (double *)P->geo_fnorm.ptr + (size_t)geo * P->nfine * 9;
    const size_t fmap_n = (size_t)6 * P->fmap_g * P->fmap_g * DSMC_FMAP_K;
    int32_t *fmap_d = (int32_t *)P->geo_fmap.ptr + (size_t)geo * fmap_n;
    double *fmapn_d = (double *)P->geo_fmapn.ptr + (size_t)geo * (fmap_n / DSMC_FMAP_K) * 36;
    if (new_geo) {
        /* a slot no queued work reads (its last user has finished); synchronous
         * copies from the caller's pageable arrays, once per geometry */
        HIP_OK(hipMemcpy(vert_d, pr.vert.data(), P->nvert * 24, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(faces_d, pr.faces.data(), P->nface_all * 12, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(enorm_d, T.enorm.data(), P->nface_all * 72, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(fnorm_d, T.fnorm.data(), P->nfine * 72, hipMemcpyHostToDevice));
        std::vector<int32_t> fm;
        build_face_map(pr.E, T, P->fmap_g, fm);
        HIP_OK(hipMemcpy(fmap_d, fm.data(), fmap_n * 4, hipMemcpyHostToDevice));
        std::vector<double> fmn;
        build_face_map_normals(fm, T.fnorm, P->fmap_g, fmn);
        HIP_OK(hipMemcpy(fmapn_d, fmn.data(), fmn.size() * 8, hipMemcpyHostToDevice));
        P->geos[geo].key = key;
        P->geos[geo].valid = 1;
    }
    P->hr_ref[tab]++;
    P->geos[geo].ref++;

    const size_t nf = P->nfine, cells = P->cells;
    double *stage = P->stage_h + (size_t)g * 3 * nf;
    memcpy(stage, T.cdf.data(), nf * 8);
    memcpy(stage + nf, pr.face_T.data(), nf * 8);
    memcpy(stage + 2 * nf, T.face_sig.data(), nf * 8);
    double *cdf_d = (double *)P->cdf.ptr + (size_t)g * nf;
    double *fT_d = (double *)P->face_T.ptr + (size_t)g * nf;
    double *fs_d = (double *)P->face_sig.ptr + (size_t)g * nf;
    HIP_OK(hipMemcpyAsync(cdf_d, stage, nf * 8, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(fT_d, stage + nf, nf * 8, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(fs_d, stage + 2 * nf, nf * 8, hipMemcpyHostToDevice, st));

    DsmcEnv &E = R.E;
    E.tq = (const double *)P->tq.ptr;
    E.cs = (const double *)P->cs.ptr;
    E.hr = hr_d;
    E.cdf = cdf_d;
    E.face_T = fT_d;
    E.face_sig = fs_d;
    E.vert = vert_d;
    E.faces = faces_d;
    E.enorm = enorm_d;
    E.fnorm = fnorm_d;
    E.fmap = fmap_d;
    E.fmap_g = P->fmap_g;
    E.fmap_n = fmapn_d;
    E.density = nullptr;
    E.density_from_count = 1;
    E.vcount = (unsigned long long *)P->vcount.ptr + (size_t)g * cells;
    /* every run deposits into 32-byte cell records, staged and issued four lanes
     * per record (vcount, and vmom for runs with moments, at collection): a
     * density-only run deposits the same records and returns their counts (the
     * 64-run batch: 8.4e10 particle-steps/s against 7.7e10 with one count
     * atomic per visit) */
    E.vmom = nullptr;
    E.cellacc = (double *)P->cellacc.ptr + (size_t)g * cells * 4;
    E.horizon = (double *)P->horizon.ptr + (size_t)g * P->cls.ftd * P->cls.ffd;
    E.counters = (unsigned long long *)P->counters.ptr + (size_t)g * 8;
    E.ttime = nullptr;
    E.ttime_rows = 5;
    E.rec_cap_cell = 0;
    if (P->per_particle) {
        E.end_state = (double *)P->end_state.ptr + (size_t)s0 * 6;
        E.steps_out = (int64_t *)P->steps_o.ptr + s0;
        E.fate_out = (int32_t *)P->fate.ptr + s0;
        E.hops_out = (int32_t *)P->hops.ptr + s0;
    }
    E.vhist = nullptr;
    E.nsig = nullptr;
    if (hist >= 0) {
        E.vhist = (unsigned long long *)P->hist.ptr + (size_t)hist * cells * P->hist_bins;
        E.vhist_bins = P->hist_bins;
        E.vhist_inv = (double)P->hist_bins / p->vhist_max;
    }
    if (coll >= 0) {
        const int deg = P->fit_deg;
        E.coeffs = (const double *)P->coll_cf.ptr + (size_t)coll * cells * (deg + 1);
        E.pmin = (const double *)P->coll_lo.ptr + (size_t)coll * cells;
        E.pmax = (const double *)P->coll_hi.ptr + (size_t)coll * cells;
        E.pvalid = (const int32_t *)P->coll_valid.ptr + (size_t)coll * cells;
        E.count_num = (const int32_t *)P->coll_cnt.ptr + (size_t)coll * cells;
        E.nsig = (const double *)P->coll_nsig.ptr + (size_t)coll * cells;
        E.fit_deg = deg;
    }
    DsmcStepEnv S;
    memset(&S, 0, sizeof S);
    S.h = E.h; S.GM = E.GM; S.c1 = E.c1; S.c2 = E.c2; S.R = E.R; S.max_height = E.max_height;
    S.sun_x = E.sun_x; S.sun_y = E.sun_y; S.max_steps = E.max_steps;
    S.ftd = E.ftd; S.ffd = E.ffd; S.hd = E.hd; S.density_from_count = 1;
    S.density = nullptr; S.vmom = nullptr; S.vcount = E.vcount; S.cellacc = E.cellacc;
    S.hr_off = tab * hd2;
    S.nsig = E.nsig;
    S.env = hist >= 0 ? (const DsmcEnv *)P->envs_d.ptr + g : nullptr;
    P->envs_h[g] = E;
    P->steps_h[g] = S;
    HIP_OK(hipMemcpyAsync((DsmcEnv *)P->envs_d.ptr + g, &P->envs_h[g], sizeof(DsmcEnv), hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync((DsmcStepEnv *)P->steps_d.ptr + g, &P->steps_h[g], sizeof(DsmcStepEnv),
                          hipMemcpyHostToDevice, st));
    HIP_OK(dsmc_launch_fill0(E.vcount, cells, st));
    if (E.cellacc) HIP_OK(dsmc_launch_fill0(E.cellacc, cells * 4, st));
    HIP_OK(dsmc_launch_fill0(E.horizon, (size_t)P->cls.ftd * P->cls.ffd, st));
    HIP_OK(dsmc_launch_fill0(E.counters, 8, st));
    if (E.vhist) HIP_OK(dsmc_launch_fill0(E.vhist, cells * P->hist_bins, st));

    R.used = 1;
    R.done = 0;
    R.p = *p;
    R.slot0 = s0;
    R.n = p->nparticles;
    R.hr_tab = tab;
    R.geo = geo;
    R.tag = tag;
    R.kind = pr.kind;
    R.moments = pr.moments;
    R.hist = hist;
    R.coll = coll;
    R.first_seq = P->issued;
    int64_t blocks = (R.n + 255) / 256;
    int grid = (int)std::min<int64_t>(blocks, (int64_t)P->cus * 8);
    HIP_OK(dsmc_launch_start(&E, g, (DsmcRec *)P->rec.ptr, s0, R.n, queues(P, pr.kind), 0, grid, st));
    P->inflight += R.n;
    *g_out = g;
    return DSMC_OK;
}

int pool_advance(Pool *P, int chunks)
{
    for (int c = 0; c < chunks; c++) {
        if (P->pending >= POOL_RING) return DSMC_ERR_BUSY;
        if (P->exec) HIP_OK(hipGraphLaunch(P->exec, P->st));
        else {
            int rc = launch_rounds(P, P->st);
            if (rc) return rc;
        }
        int k = (int)(P->issued % POOL_RING);
        HIP_OK(hipMemcpyAsync(P->cnt_h + (size_t)k * P->max_runs * 8, P->counters.ptr, (size_t)P->max_runs * 64,
                              hipMemcpyDeviceToHost, P->st));
        HIP_OK(hipMemcpyAsync(P->ctl_h + k * 16, P->qctl.ptr, 16 * sizeof(unsigned int), hipMemcpyDeviceToHost,
                              P->st));
        HIP_OK(hipEventRecord(P->ring_ev[k], P->st));
        P->issued++;
        P->pending++;
    }
    return DSMC_OK;
}

int pool_poll(Pool *P, std::vector<int> *finished)
{
    if (P->pending == 0) return 0;
    int k = (int)(P->consumed % POOL_RING);
    const int64_t seq = P->consumed;
    HIP_OK(hipEventSynchronize(P->ring_ev[k]));
    P->consumed++;
    P->pending--;
    const unsigned long long *cnt = P->cnt_h + (size_t)k * P->max_runs * 8;
    P->rounds = P->ctl_h[k * 16 + 8];
    P->last_ev = P->ring_ev[k];
    int nf = 0;
    for (int g = 0; g < P->max_runs; g++) {
        PoolRun &R = P->runs[g];
        if (!R.used || R.done || seq < R.first_seq) continue;
        if ((int64_t)cnt[(size_t)g * 8 + 7] >= R.n) {
            R.done = 1;
            P->inflight -= R.n;
            if (finished) finished->push_back(g);
            nf++;
        }
    }
    return nf;
}

int pool_collect(Pool *P, int g, dsmc_result *r)
{
    if (g < 0 || g >= P->max_runs || !P->runs[g].used || !P->runs[g].done) return DSMC_ERR_ARG;
    PoolRun &R = P->runs[g];
    const DsmcEnv &E = R.E;
    const size_t cells = P->cells, nh = (size_t)P->cls.ftd * P->cls.ffd;
    hipStream_t cs = P->cst;
    /* the run finished within the replay whose counters were just read */
    HIP_OK(hipStreamWaitEvent(cs, P->last_ev, 0));
    char *o = (char *)P->out_h;
    unsigned long long cnt[8];
    /* a density-only run's counts: folded from its records on the device first */
    if (!R.moments) HIP_OK(dsmc_launch_fold_counts(E.cellacc, E.vcount, (int64_t)cells, cs));
    HIP_OK(hipMemcpyAsync(o, E.vcount, cells * 8, hipMemcpyDeviceToHost, cs));
    if (R.moments) HIP_OK(hipMemcpyAsync(o + cells * 8, E.cellacc, cells * 32, hipMemcpyDeviceToHost, cs));
    HIP_OK(hipMemcpyAsync(o + cells * 40, E.horizon, nh * 8, hipMemcpyDeviceToHost, cs));
    HIP_OK(hipMemcpyAsync(o + cells * 40 + nh * 8, E.counters, sizeof cnt, hipMemcpyDeviceToHost, cs));
    HIP_OK(hipStreamSynchronize(cs));
    memcpy(cnt, o + cells * 40 + nh * 8, sizeof cnt);
    int64_t *vc = (int64_t *)o;
    const double *acc = (const double *)(o + cells * 8);
    /* a moments run's counts and velocity sums are in its cell records: exact
     * integer counts (below 2^53) in the fourth word */
    if (R.moments) for (size_t i = 0; i < cells; i++) vc[i] += (int64_t)acc[4 * i + 3];
    if (r->vcount) memcpy(r->vcount, vc, cells * 8);
    if (r->density) for (size_t i = 0; i < cells; i++) r->density[i] = (double)vc[i];
    if (r->vmom) {
        if (R.moments) {
            for (size_t i = 0; i < cells; i++)
                for (int k = 0; k < 3; k++) r->vmom[3 * i + k] = acc[4 * i + k];
        } else {
            memset(r->vmom, 0, cells * 24);       /* a density-only run */
        }
    }
    if (r->horizon) memcpy(r->horizon, o + cells * 40, nh * 8);
    const int64_t n = R.n;
    if (P->per_particle) {
        if (r->end_state) HIP_OK(hipMemcpy(r->end_state, E.end_state, n * 48, hipMemcpyDeviceToHost));
        if (r->steps) HIP_OK(hipMemcpy(r->steps, E.steps_out, n * 8, hipMemcpyDeviceToHost));
        if (r->fate) HIP_OK(hipMemcpy(r->fate, E.fate_out, n * 4, hipMemcpyDeviceToHost));
        if (r->hops) HIP_OK(hipMemcpy(r->hops, E.hops_out, n * 4, hipMemcpyDeviceToHost));
    } else if (r->end_state || r->steps || r->fate || r->hops) {
        return DSMC_ERR_ARG;
    }
    if (r->ttime) memset(r->ttime, 0, sizeof(double) * 5 * n);
    r->vrec_n = 0;
    r->n_escaped = (int64_t)cnt[0];
    r->n_frozen = (int64_t)cnt[1];
    r->n_photolysed = (int64_t)cnt[2];
    r->n_depleted = (int64_t)cnt[3];
    r->n_capped = (int64_t)cnt[4];
    r->n_collisions = (int64_t)cnt[5];
    r->total_steps = (int64_t)cnt[6];
    return DSMC_OK;
}

int pool_release(Pool *P, int g)
{
    if (g < 0 || g >= P->max_runs || !P->runs[g].used || !P->runs[g].done) return DSMC_ERR_ARG;
    PoolRun &R = P->runs[g];
    give_slots(P, R.slot0, R.n);
    if (R.hr_tab >= 0) P->hr_ref[R.hr_tab]--;
    if (R.geo >= 0) P->geos[R.geo].ref--;
    R = PoolRun();
    return DSMC_OK;
}

}  // namespace dsmc
