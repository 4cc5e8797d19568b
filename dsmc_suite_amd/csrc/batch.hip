/*
 * batch.hip - many independent physical runs in one pool (dsmc_run_batch and
 * the staged dsmc_batch_* API), and the launch-shape autotuner.
 *
 * A batch admits every run into one pool (pool.hip): their particles share the
 * persistent flight / service kernels, whose rounds are replayed from one
 * captured hipGraph until every run has finished.  Uneven particle counts cost
 * nothing extra: a wavefront pulls whatever particle is next in the queue, of
 * whichever run.
 *
 * Autotuner: coordinate descent over the launch shape - waves per CU (blocks per
 * CU), block size, refill batch (idle lanes a wavefront gathers before pulling
 * new particles or servicing landings), landing service inside the flight kernel
 * or in a kernel of its own, and the per-launch step budget - scored by particle-steps per
 * second of a scaled-down copy of the workload run to completion.  Shapes are
 * cached per device and workload class (grid, surface level, size bucket).
 */
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>
#include <math.h>
#include <chrono>
#include <map>
#include <mutex>
#include <string>
#include <algorithm>
#include <vector>

#include "pool.h"
#include "dsmc_kernels.h"

namespace dsmc {

static double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

/* Runs every admitted run of P to completion; results[g] receives run g's outputs
 * (NULL: none).  done_ms[g] = host time at which run g was seen finished. */
static int pool_drain(Pool *P, dsmc_result *const *results, std::vector<double> *done_ms, double t0)
{
    std::vector<int> fin;
    const int64_t round_limit = (int64_t)1 << 24;
    while (P->inflight > 0 || P->pending > 0) {
        while (P->inflight > 0 && P->pending < 2) {
            int rc = pool_advance(P, 1);
            if (rc) return rc;
        }
        fin.clear();
        int nf = pool_poll(P, &fin);
        if (nf < 0) return nf;
        for (int g : fin) {
            if (done_ms) (*done_ms)[g] = now_ms() - t0;
            if (results && results[g]) {
                int rc = pool_collect(P, g, results[g]);
                if (rc) return rc;
            }
        }
        if (P->rounds > round_limit) {
            fprintf(stderr, "dsmc_hip: batch did not finish within %lld rounds\n", (long long)P->rounds);
            return DSMC_ERR_DEVICE;
        }
    }
    return DSMC_OK;
}

/* ------------------------------------------------------------------ autotuner */
static std::mutex tune_mu;
static std::map<std::string, dsmc_launch_shape> tune_cache;

static std::string class_key(const PoolClass &c, int64_t total, int nruns)
{
    int dev = 0;
    (void)hipGetDevice(&dev);
    char buf[160];
    snprintf(buf, sizeof buf, "d%d m%d a%d g%dx%dx%d L%d n2^%d r2^%d", dev, c.mode, c.adv, c.ftd, c.ffd, c.hd,
             c.levels, (int)floor(log2((double)(total > 1 ? total : 1))),
             (int)floor(log2((double)(nruns > 1 ? nruns : 1))));
    return buf;
}

/* Candidate shapes around `s`, one coordinate at a time. */
template <class Trial>
static int coordinate_descent(dsmc_launch_shape *best, Trial trial, double *best_score)
{
    double bs = -1.0;
    int rc = trial(*best, &bs);
    if (rc) return rc;
    auto try_shape = [&](dsmc_launch_shape c) -> int {
        if (!shape_ok(c)) return DSMC_OK;
        double sc = -1.0;
        int r = trial(c, &sc);
        if (r) return r;
        if (sc > bs * 1.02) { bs = sc; *best = c; }
        return DSMC_OK;
    };
    const int waves_per_cu[] = {8, 12, 16, 20, 24};
    dsmc_launch_shape base = *best;
    for (int w : waves_per_cu) {
        dsmc_launch_shape c = base;
        c.blocks_per_cu = w * 64 / c.block;
        if (c.blocks_per_cu * c.block != w * 64 || w * 64 == base.block * base.blocks_per_cu) continue;
        if ((rc = try_shape(c))) return rc;
    }
    base = *best;
    for (int b : {128, 256}) {
        if (b == base.block) continue;
        dsmc_launch_shape c = base;
        c.block = b;
        c.blocks_per_cu = base.blocks_per_cu * base.block / b;
        if ((rc = try_shape(c))) return rc;
    }
    base = *best;
    for (int r : {16, 32, 40, 48, 64}) {
        if (r == base.relaunch_batch) continue;
        dsmc_launch_shape c = base;
        c.relaunch_batch = r;
        if ((rc = try_shape(c))) return rc;
    }
    base = *best;
    {
        dsmc_launch_shape c = base;
        c.service_inline = 1 - base.service_inline;
        if ((rc = try_shape(c))) return rc;
    }
    base = *best;
    for (int sb : {256, 512, 1024, 2048, 4096}) {
        if (sb == base.step_budget) continue;
        dsmc_launch_shape c = base;
        c.step_budget = sb;
        if ((rc = try_shape(c))) return rc;
    }
    if (best_score) *best_score = bs;
    return DSMC_OK;
}

/* Tunes the pool's shape on a copy of `preps` scaled to `frac` of the particles. */
static int tune_pool(Pool *P, const std::vector<PoolPrep> &preps, double frac, dsmc_launch_shape *out)
{
    std::vector<PoolPrep> sub(preps);
    for (auto &pr : sub) {
        int64_t n = (int64_t)ceil((double)pr.p.nparticles * frac);
        if (n < 256) n = pr.p.nparticles < 256 ? pr.p.nparticles : 256;
        pr.p.nparticles = n;
        pr.E.nparticles = n;
    }
    auto trial = [&](const dsmc_launch_shape &s, double *score) -> int {
        /* the flight kernel is persistent: a grid of more blocks than the CUs
         * hold leaves the extra blocks to start as others end, a shape that
         * won on scaled-down samples but ran the 64-run batch 1-2% slower */
        if (s.blocks_per_cu > pool_occupancy(P, s.block)) { *score = -1.0; return DSMC_OK; }
        int rc = pool_set_shape(P, s);
        if (rc) return rc;
        std::vector<int> gs;
        double t0 = now_ms();
        for (size_t i = 0; i < sub.size(); i++) {
            int g = -1;
            if ((rc = pool_admit(P, sub[i], i, &g))) return rc;
            gs.push_back(g);
        }
        if ((rc = pool_drain(P, nullptr, nullptr, t0))) return rc;
        double ms = now_ms() - t0;
        int64_t steps = 0;
        const unsigned long long *cnt = P->cnt_h + (size_t)((P->consumed - 1) % 2) * P->max_runs * 8;
        for (int g : gs) {
            steps += (int64_t)cnt[(size_t)g * 8 + 6];
            pool_release(P, g);
        }
        *score = (double)steps / ms;
        return DSMC_OK;
    };
    /* the descent starts from the shape that ran the 64-run batch fastest (three
     * 256-lane blocks per CU, services of 40 lanes, 4096-step rounds): it keeps
     * a coordinate's current value unless another beats it by 2% on the sample */
    dsmc_launch_shape best = P->shape;
    best.block = 256;
    best.blocks_per_cu = 3;
    best.relaunch_batch = 40;
    best.step_budget = 4096;
    best.service_inline = 1;
    double score = 0.0;
    int rc = coordinate_descent(&best, trial, &score);
    if (rc) return rc;
    *out = best;
    return pool_set_shape(P, best);
}

}  // namespace dsmc

using namespace dsmc;

/* ------------------------------------------------------------------ staged batch */
struct dsmc_batch {
    Pool P;
    std::vector<PoolPrep> preps;
    std::vector<int> g;                    /* run slot of run i in the last launch */
    std::vector<double> done_ms;
    double wall_ms = 0.0;
    int64_t rounds = 0;
    int launched = 0;
};

static int batch_prepare(int nruns, const dsmc_params *params, const dsmc_surface *surfaces,
                         std::vector<PoolPrep> *preps, PoolClass *cls, int64_t *total)
{
    if (nruns <= 0 || !params || !surfaces) return DSMC_ERR_ARG;
    preps->resize(nruns);
    *total = 0;
    for (int i = 0; i < nruns; i++) {
        PoolClass c;
        int rc = pool_class_of(&params[i], &surfaces[i], &c);
        if (rc) return rc;
        if (i == 0) *cls = c;
        else if (c.ftd != cls->ftd || c.ffd != cls->ffd || c.hd != cls->hd || c.levels != cls->levels)
            return DSMC_ERR_ARG;
        if ((rc = pool_prepare(&params[i], &surfaces[i], &(*preps)[i]))) return rc;
        *total += params[i].nparticles;
    }
    return DSMC_OK;
}

extern "C" dsmc_batch *dsmc_batch_create(int nruns, const dsmc_params *params, const dsmc_surface *surfaces,
                                         int flags, const dsmc_launch_shape *tune, int *status)
{
    dsmc_batch *b = new dsmc_batch();
    PoolClass cls;
    int64_t total = 0;
    const int per_particle = (flags & DSMC_BATCH_PER_PARTICLE) ? 1 : 0;
    const int moments = (flags & DSMC_BATCH_DENSITY_ONLY) ? 0 : 1;
    int rc = (flags & ~(DSMC_BATCH_PER_PARTICLE | DSMC_BATCH_DENSITY_ONLY)) ? DSMC_ERR_ARG : DSMC_OK;
    if (!rc) rc = batch_prepare(nruns, params, surfaces, &b->preps, &cls, &total);
    for (auto &pr : b->preps) pr.moments = moments;
    b->P.late_unbounded = 1;          /* every run is admitted before the first round */
    if (!rc) rc = pool_create(&b->P, cls, nruns, total, per_particle, tune);
    if (!rc && !tune) {
        std::string key = class_key(cls, total, nruns) + (moments ? "" : "/density");
        dsmc_launch_shape s;
        bool hit = false;
        {
            std::lock_guard<std::mutex> lk(tune_mu);
            auto it = tune_cache.find(key);
            if (it != tune_cache.end()) { s = it->second; hit = true; }
        }
        if (hit) rc = pool_set_shape(&b->P, s);
        else {
            /* a copy large enough to fill the device: >= 4M particles or the whole
             * batch (a 12M-particle copy of the 64-run batch took twice as long to
             * tune and chose the same shape) */
            double frac = std::min(1.0, std::max(1.0 / 16.0, 4.0e6 / (double)total));
            rc = tune_pool(&b->P, b->preps, frac, &s);
            if (!rc) {
                std::lock_guard<std::mutex> lk(tune_mu);
                tune_cache[key] = s;
            }
        }
    }
    if (rc) {
        delete b;
        b = nullptr;
    }
    if (status) *status = rc;
    return b;
}

/* Admits every run and runs the batch to completion. */
extern "C" int dsmc_batch_launch(dsmc_batch *b, dsmc_result *const *results)
{
    if (!b) return DSMC_ERR_ARG;
    Pool *P = &b->P;
    for (int g = 0; g < P->max_runs; g++)
        if (P->runs[g].used) pool_release(P, g);
    const int n = (int)b->preps.size();
    b->g.assign(n, -1);
    b->done_ms.assign(n, 0.0);
    const int64_t r0 = P->rounds;
    double t0 = now_ms();
    for (int i = 0; i < n; i++) {
        int rc = pool_admit(P, b->preps[i], (uint64_t)i, &b->g[i]);
        if (rc) return rc;
    }
    std::vector<dsmc_result *> by_slot(P->max_runs, nullptr);
    if (results)
        for (int i = 0; i < n; i++) by_slot[b->g[i]] = results[i];
    std::vector<double> dm(P->max_runs, 0.0);
    int rc = pool_drain(P, by_slot.data(), &dm, t0);
    if (rc) return rc;
    b->wall_ms = now_ms() - t0;
    for (int i = 0; i < n; i++) b->done_ms[i] = dm[b->g[i]];
    b->rounds = P->rounds - r0;
    b->launched = 1;
    return DSMC_OK;
}

extern "C" int dsmc_batch_timing(dsmc_batch *b, double *wall_ms, double *run_done_ms, int64_t *rounds,
                                 dsmc_launch_shape *shape)
{
    if (!b || !b->launched) return DSMC_ERR_ARG;
    if (wall_ms) *wall_ms = b->wall_ms;
    if (run_done_ms) memcpy(run_done_ms, b->done_ms.data(), b->done_ms.size() * sizeof(double));
    if (rounds) *rounds = b->rounds;
    if (shape) *shape = b->P.shape;
    return DSMC_OK;
}

extern "C" void dsmc_batch_destroy(dsmc_batch *b) { delete b; }

extern "C" int dsmc_run_batch(int nruns, const dsmc_params *params, const dsmc_surface *surfaces,
                              dsmc_result *results, const dsmc_launch_shape *tune)
{
    if (nruns <= 0 || !results) return DSMC_ERR_ARG;
    int flags = DSMC_BATCH_DENSITY_ONLY;
    for (int i = 0; i < nruns; i++) {
        if (results[i].end_state || results[i].steps || results[i].fate || results[i].hops)
            flags |= DSMC_BATCH_PER_PARTICLE;
        if (results[i].vmom) flags &= ~DSMC_BATCH_DENSITY_ONLY;
    }
    int rc = 0;
    dsmc_batch *b = dsmc_batch_create(nruns, params, surfaces, flags, tune, &rc);
    if (!b) return rc;
    std::vector<dsmc_result *> rp(nruns);
    for (int i = 0; i < nruns; i++) rp[i] = &results[i];
    rc = dsmc_batch_launch(b, rp.data());
    if (!rc)
        for (int i = 0; i < nruns; i++) results[i].elapsed_ms = b->done_ms[i];
    dsmc_batch_destroy(b);
    return rc;
}

/* Single-run autotune: the batch tuner on a one-run pool of a sample of the run. */
extern "C" int dsmc_autotune(const dsmc_params *p, const dsmc_surface *s, dsmc_launch_shape *out)
{
    if (!p || !s || !out) return DSMC_ERR_ARG;
    dsmc_params q = *p;
    if (q.nparticles > (1 << 19)) q.nparticles = 1 << 19;
    int rc = 0;
    dsmc_batch *b = dsmc_batch_create(1, &q, s, 0, nullptr, &rc);
    if (!b) return rc;
    *out = b->P.shape;
    dsmc_batch_destroy(b);
    return DSMC_OK;
}
