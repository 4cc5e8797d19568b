/*
 * serve.hip - the continuous-batching request loop (dsmc_server_*).
 *
 * Requests are prepared on the host when submitted (environment, exact index
 * tables), queued in arrival order and admitted into one device pool as soon as
 * a contiguous range of particle slots, a grid slot and table slots are free:
 * the pool's particle slots are the server's HBM budget divided by the slot
 * footprint (256 B record + 12 B of queue entries, + 64 B with per-particle
 * outputs), after the grid slots (max_inflight x grid bytes) are set aside.
 * Admitted runs join the particles already in flight at the next replay of the
 * round graph; a finished run's results are copied out on a second stream while
 * the next replay runs, and its slots are free for the next request.
 *
 * Collisional requests run as two pool runs: phase 1 (ballistic, keeping speed
 * histograms in one of the pool's histogram slots) and, once it has finished,
 * phase 2 (collisional) against the fits and n*sigma table built on the device
 * from phase 1 into one of the pool's fit-table slots.  Both slots are taken when
 * phase 1 is admitted; phase 2 goes to the front of the admission queue.
 *
 * Every particle's random stream is keyed by its run's seed and its index in the
 * run, so a request's per-particle results do not depend on what else is in the
 * pool; its grid sums differ only by the order of floating-point atomic
 * additions (integer counts are exact).
 */
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <string.h>
#include <chrono>
#include <deque>
#include <map>
#include <memory>
#include <vector>

#include "pool.h"

using namespace dsmc;

namespace {

double now_ms()
{
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

enum { REQ_QUEUED = 0, REQ_ADMITTED = 1, REQ_DONE = 2 };

struct ServeReq {
    PoolPrep prep;                  /* the run, or phase 1 of a collisional request */
    PoolPrep prep2;                 /* phase 2 of a collisional request */
    int collisional = 0, phase = 1, hist = -1, coll = -1;
    std::vector<int64_t> vcount1;   /* phase-1 deposit counts */
    dsmc_result *r = nullptr;
    double t_submit = 0.0;
    int state = REQ_QUEUED, cancelled = 0, status = 0, g = -1;
};

}  // namespace

struct dsmc_server {
    size_t budget = 0;
    int max_inflight = 0, per_particle = 0, collisional = 0;
    std::unique_ptr<Pool> pool;
    PoolClass cls;
    std::deque<uint64_t> queue;
    std::map<uint64_t, ServeReq> reqs;
    std::vector<uint64_t> ticket_of;   /* run slot -> ticket */
    uint64_t next_ticket = 1;
    int64_t completed = 0, steps_done = 0;   /* requests completed; their particle-steps (both phases) */
};

extern "C" dsmc_server *dsmc_server_create(size_t hbm_budget_bytes, int max_inflight, int flags)
{
    if (hbm_budget_bytes == 0 || max_inflight <= 0 || max_inflight > 4096 ||
        (flags & ~(DSMC_SERVE_PER_PARTICLE | DSMC_SERVE_COLLISIONAL)))
        return nullptr;
    if (ensure_device()) return nullptr;
    dsmc_server *sv = new dsmc_server();
    sv->budget = hbm_budget_bytes;
    sv->max_inflight = max_inflight;
    sv->per_particle = (flags & DSMC_SERVE_PER_PARTICLE) ? 1 : 0;
    sv->collisional = (flags & DSMC_SERVE_COLLISIONAL) ? 1 : 0;
    return sv;
}

static int make_pool(dsmc_server *sv, const PoolClass &c)
{
    const size_t cells = (size_t)c.ftd * c.ffd * c.hd;
    size_t fixed = (size_t)sv->max_inflight * pool_run_bytes(c) + ((size_t)64 << 20);
    if (sv->collisional)            /* histogram and fit-table slots (64 bins, degree 6 at most) */
        fixed += (size_t)POOL_MAX_HIST * cells * 64 * 8 + (size_t)POOL_MAX_COLL * cells * (7 + 4) * 8;
    if (sv->budget <= fixed) return DSMC_ERR_BUDGET;
    int64_t slots = (int64_t)((sv->budget - fixed) / pool_slot_bytes(sv->per_particle));
    if (slots > INT32_MAX) slots = INT32_MAX;
    if (slots <= 0) return DSMC_ERR_BUDGET;
    std::unique_ptr<Pool> P(new Pool());
    int rc = pool_create(P.get(), c, sv->max_inflight, slots, sv->per_particle, nullptr);
    if (rc) return rc;
    sv->pool = std::move(P);
    sv->cls = c;
    sv->ticket_of.assign(sv->max_inflight, 0);
    return DSMC_OK;
}

extern "C" int dsmc_server_set_shape(dsmc_server *sv, const dsmc_launch_shape *shape)
{
    if (!sv || !shape || !sv->pool) return DSMC_ERR_ARG;
    if (sv->pool->pending || sv->pool->inflight) return DSMC_ERR_BUSY;
    return pool_set_shape(sv->pool.get(), *shape);
}

static int submit(dsmc_server *sv, const dsmc_params *p, const dsmc_params *p2, const dsmc_surface *s,
                  dsmc_result *r, uint64_t *ticket);

extern "C" int dsmc_server_submit(dsmc_server *sv, const dsmc_params *p, const dsmc_surface *s, dsmc_result *r,
                                  uint64_t *ticket)
{
    return submit(sv, p, nullptr, s, r, ticket);
}

/* A collisional request: phase 1 with p1, phase 2 with p2 (same grid); r
 * receives the phase-2 results. */
extern "C" int dsmc_server_submit_collisional(dsmc_server *sv, const dsmc_params *p1, const dsmc_params *p2,
                                              const dsmc_surface *s, dsmc_result *r, uint64_t *ticket)
{
    if (!p2) return DSMC_ERR_ARG;
    return submit(sv, p1, p2, s, r, ticket);
}

static int submit(dsmc_server *sv, const dsmc_params *p, const dsmc_params *p2, const dsmc_surface *s,
                  dsmc_result *r, uint64_t *ticket)
{
    if (!sv || !p || !s || !r || !ticket) return DSMC_ERR_ARG;
    if (p2 && !sv->collisional) return DSMC_ERR_ARG;
    if (p2 && (p2->ftd != p->ftd || p2->ffd != p->ffd || p2->hd != p->hd || p2->vhist_bins != p->vhist_bins ||
               p2->vhist_max != p->vhist_max || p2->fit_degree < 1 || p2->fit_degree > 6 || p->vhist_bins <= 0 ||
               p->vhist_bins > 64 || p2->mode != DSMC_MODE_PHYSICAL))
        return DSMC_ERR_ARG;
    PoolClass c;
    int rc = pool_class_of(p, s, &c);
    if (rc) return rc;
    if (!sv->pool) {
        if ((rc = make_pool(sv, c))) return rc;
    } else if (c.ftd != sv->cls.ftd || c.ffd != sv->cls.ffd || c.hd != sv->cls.hd || c.levels != sv->cls.levels) {
        return DSMC_ERR_ARG;
    }
    if (p->nparticles > sv->pool->max_slots || (p2 && p2->nparticles > sv->pool->max_slots)) return DSMC_ERR_BUDGET;
    if (!sv->per_particle && (r->end_state || r->steps || r->fate || r->hops)) return DSMC_ERR_ARG;
    if (p2 && (rc = pool_enable_collisional(sv->pool.get(), p->vhist_bins, p2->fit_degree))) return rc;
    uint64_t t = sv->next_ticket++;
    ServeReq &q = sv->reqs[t];
    if ((rc = pool_prepare(p, s, &q.prep)) || (p2 && (rc = pool_prepare(p2, s, &q.prep2, 1)))) {
        sv->reqs.erase(t);
        return rc;
    }
    /* velocity moments only for requests that ask for them (and never for a
     * collisional request's phase 1, which hands over counts and histograms) */
    q.prep.moments = (!p2 && r->vmom) ? 1 : 0;
    if (p2) {
        q.collisional = 1;
        q.prep2.moments = r->vmom ? 1 : 0;
        q.vcount1.assign((size_t)p->ftd * p->ffd * p->hd, 0);
    }
    q.r = r;
    q.t_submit = now_ms();
    sv->queue.push_back(t);
    *ticket = t;
    return DSMC_OK;
}

static void finish(dsmc_server *sv, uint64_t t, ServeReq &q, int status)
{
    q.state = REQ_DONE;
    q.status = status;
    if (q.r) q.r->elapsed_ms = now_ms() - q.t_submit;
    (void)sv;
}

/* Admits queued requests in arrival order while they fit, keeps two replays of
 * the round graph in flight, and completes the runs the older one finished.
 * Returns the number of requests completed, or a negative status. */
extern "C" int dsmc_server_step(dsmc_server *sv)
{
    if (!sv) return DSMC_ERR_ARG;
    if (!sv->pool) return 0;
    Pool *P = sv->pool.get();
    while (!sv->queue.empty()) {
        uint64_t t = sv->queue.front();
        ServeReq &q = sv->reqs[t];
        int g = -1, rc;
        if (q.collisional && q.phase == 1) {
            /* histogram and fit-table slots for the whole request, taken up front */
            int h = pool_take_slot(P->hist_used);
            int c = h >= 0 ? pool_take_slot(P->coll_used) : -1;
            if (c < 0) { if (h >= 0) P->hist_used[h] = 0; break; }
            rc = pool_admit(P, q.prep, t, &g, h, -1);
            if (rc) { P->hist_used[h] = 0; P->coll_used[c] = 0; }
            else { q.hist = h; q.coll = c; }
        } else if (q.collisional) {
            rc = pool_admit(P, q.prep2, t, &g, -1, q.coll);
        } else {
            rc = pool_admit(P, q.prep, t, &g);
        }
        if (rc == DSMC_ERR_BUSY) break;
        sv->queue.pop_front();
        if (rc) {
            if (q.coll >= 0) { P->coll_used[q.coll] = 0; q.coll = -1; }
            finish(sv, t, q, rc);
            continue;
        }
        q.state = REQ_ADMITTED;
        q.g = g;
        sv->ticket_of[g] = t;
        /* the tables are on the device now; phase 1 keeps the launch CDF (its sum
         * is the emission rate of the number density) */
        PoolPrep &done_prep = q.collisional && q.phase == 2 ? q.prep2 : q.prep;
        std::vector<double> keep_cdf;
        if (q.collisional && q.phase == 1) keep_cdf.swap(done_prep.T.cdf);
        done_prep.T = HostTables();
        done_prep.T.cdf.swap(keep_cdf);
        done_prep.vert = std::vector<double>();
        done_prep.faces = std::vector<int32_t>();
        done_prep.face_T = std::vector<double>();
    }
    while (P->inflight > 0 && P->pending < 2) {
        int rc = pool_advance(P, 1);
        if (rc) return rc;
    }
    if (P->pending == 0) return 0;
    std::vector<int> fin;
    int nf = pool_poll(P, &fin);
    if (nf < 0) return nf;
    int done = 0;
    for (int g : fin) {
        uint64_t t = sv->ticket_of[g];
        ServeReq &q = sv->reqs[t];
        if (q.collisional && q.phase == 1) {
            /* phase 1 done: counts -> number density -> fits on the device -> phase 2 */
            dsmc_result r1;
            memset(&r1, 0, sizeof r1);
            r1.vcount = q.vcount1.data();
            int rc = q.cancelled ? DSMC_ERR_CANCELLED : pool_collect(P, g, &r1);
            if (!rc) sv->steps_done += r1.total_steps;      /* phase-1 particle-steps count too */
            if (!rc) {
                std::vector<double> nd(q.vcount1.size());
                number_density(&q.prep.p, q.prep.T.cdf.back(), q.vcount1.data(), nd.data());
                rc = pool_build_fits(P, q.hist, q.coll, nd.data(), &q.prep2.p);
            }
            P->hist_used[q.hist] = 0;      /* later reuse is ordered after the fit kernel */
            q.hist = -1;
            pool_release(P, g);
            sv->ticket_of[g] = 0;
            if (rc) {
                P->coll_used[q.coll] = 0;
                q.coll = -1;
                finish(sv, t, q, rc);
                done++;
            } else {
                q.phase = 2;
                q.state = REQ_QUEUED;
                sv->queue.push_front(t);
            }
            continue;
        }
        if (q.collisional && q.coll >= 0) { P->coll_used[q.coll] = 0; q.coll = -1; }
        int rc = q.cancelled ? DSMC_ERR_CANCELLED : pool_collect(P, g, q.r);
        if (!q.cancelled && !rc) {
            sv->completed++;
            sv->steps_done += q.r->total_steps;
        }
        pool_release(P, g);
        sv->ticket_of[g] = 0;
        finish(sv, t, q, rc);
        done++;
    }
    return done;
}

extern "C" int dsmc_server_poll(dsmc_server *sv, uint64_t ticket)
{
    if (!sv) return DSMC_ERR_ARG;
    auto it = sv->reqs.find(ticket);
    if (it == sv->reqs.end()) return DSMC_ERR_ARG;
    if (it->second.state != REQ_DONE) return 0;
    return it->second.status ? it->second.status : 1;
}

extern "C" int dsmc_server_wait(dsmc_server *sv, uint64_t ticket)
{
    for (;;) {
        int st = dsmc_server_poll(sv, ticket);
        if (st != 0) return st < 0 ? st : DSMC_OK;
        int rc = dsmc_server_step(sv);
        if (rc < 0) return rc;
    }
}

/* A queued request is dropped at once; an admitted one finishes on the device
 * (its particles hold their slots until then) and its results are discarded. */
extern "C" int dsmc_server_cancel(dsmc_server *sv, uint64_t ticket)
{
    if (!sv) return DSMC_ERR_ARG;
    auto it = sv->reqs.find(ticket);
    if (it == sv->reqs.end()) return DSMC_ERR_ARG;
    ServeReq &q = it->second;
    if (q.state == REQ_DONE) return DSMC_ERR_ARG;
    q.cancelled = 1;
    if (q.state == REQ_QUEUED) {
        for (auto qi = sv->queue.begin(); qi != sv->queue.end(); ++qi)
            if (*qi == ticket) { sv->queue.erase(qi); break; }
        if (q.coll >= 0 && sv->pool) { sv->pool->coll_used[q.coll] = 0; q.coll = -1; }
        finish(sv, ticket, q, DSMC_ERR_CANCELLED);
    }
    return DSMC_OK;
}

/* Forgets a completed request (its ticket becomes invalid). */
extern "C" int dsmc_server_release(dsmc_server *sv, uint64_t ticket)
{
    if (!sv) return DSMC_ERR_ARG;
    auto it = sv->reqs.find(ticket);
    if (it == sv->reqs.end() || it->second.state != REQ_DONE) return DSMC_ERR_ARG;
    sv->reqs.erase(it);
    return DSMC_OK;
}

extern "C" int dsmc_server_get_stats(dsmc_server *sv, dsmc_server_stats *out)
{
    if (!sv || !out) return DSMC_ERR_ARG;
    memset(out, 0, sizeof *out);
    out->completed = sv->completed;
    out->particle_steps = sv->steps_done;
    out->queued = (int64_t)sv->queue.size();
    if (sv->pool) {
        Pool *P = sv->pool.get();
        int64_t running = 0;
        for (int g = 0; g < P->max_runs; g++) running += P->runs[g].used && !P->runs[g].done;
        out->running = running;
        out->particles_in_flight = P->inflight;
        out->slots = P->max_slots;
        out->rounds = P->rounds;
        out->shape = P->shape;
    }
    return DSMC_OK;
}

extern "C" void dsmc_server_destroy(dsmc_server *sv) { delete sv; }
