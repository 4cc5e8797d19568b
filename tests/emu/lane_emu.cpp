/*
 * lane_emu.cpp - TEST-ONLY host build of the product's per-lane state machine
 * (dsmc_suite_amd/csrc/dsmc_lane.h) with plain host accumulation, one particle
 * at a time.  It lets the CPU test suite compare the exact code the gfx950
 * kernels run against the oracle's device-math build without a GPU.  Never
 * linked into the product library.
 */
#include <stdlib.h>
#include <string.h>
#include "../../dsmc_suite_amd/csrc/dsmc_tables.h"

template <int MODE, int ADV>
static void run_one(DsmcLane &L, const DsmcEnv &E, int64_t n, DsmcCounts &cnt)
{
    lane_start<DsmcOpsHost, MODE, ADV>(L, E, n, cnt);
    DsmcCellTabs TB = {E.tq, E.hr, E.cs};
    while (L.phase != LANE_DONE) {
        if (L.phase == LANE_FLIGHT) lane_step<DsmcOpsHost, MODE, ADV>(L, E, TB);
        else lane_service<DsmcOpsHost, MODE, ADV>(L, E, cnt);
    }
}

/* Advanced (collisional) compat run with fits supplied by the caller. */
extern "C" int emu_advanced(const dsmc_params *p, const dsmc_surface *s, const double *copy_atm,
                            int64_t atm_total, const int32_t *count_num, const double *coeffs,
                            const double *pmin, const double *pmax, const int32_t *pvalid, dsmc_result *r)
{
    DsmcEnv E;
    dsmc::HostTables T;
    int rc = dsmc::build_env(p, s, 1, &E, &T);
    if (rc) return rc;
    size_t cells = (size_t)p->hd * p->ftd * p->ffd;
    std::vector<double> den(cells, 0.0), hor((size_t)p->ftd * p->ffd, 0.0), atm;
    if (copy_atm) {
        atm.resize(atm_total);
        for (int64_t i = 0; i < atm_total; i++) atm[i] = copy_atm[i] * p->density_scale;
    }
    E.T = s->T_latlon;
    E.vt_speeds = T.vt_speeds.data(); E.vt_count = T.vt_count.data();
    E.tq = T.tq.data(); E.hr = T.hr.data(); E.cs = T.cs.data();
    E.density = den.data(); E.horizon = hor.data();
    E.copy_atm = copy_atm ? atm.data() : nullptr; E.atm_total = copy_atm ? atm_total : 0;
    E.count_num = count_num; E.coeffs = coeffs; E.pmin = pmin; E.pmax = pmax; E.pvalid = pvalid;
    E.ttime = r->ttime; E.ttime_rows = 9;
    if (E.ttime) memset(E.ttime, 0, sizeof(double) * 9 * p->nparticles);
    E.end_state = r->end_state; E.steps_out = r->steps; E.fate_out = r->fate; E.hops_out = r->hops;
    DsmcCounts cnt;
    memset(&cnt, 0, sizeof cnt);
    for (int64_t n = 0; n < p->nparticles; n++) {
        DsmcLane L;
        memset(&L, 0, sizeof L);
        run_one<0, 1>(L, E, n, cnt);
    }
    if (r->density) memcpy(r->density, den.data(), cells * sizeof(double));
    if (r->horizon) memcpy(r->horizon, hor.data(), hor.size() * sizeof(double));
    r->n_escaped = cnt.esc; r->n_frozen = cnt.frozen; r->n_photolysed = cnt.photo;
    r->n_depleted = cnt.depleted; r->n_capped = cnt.capped; r->n_collisions = cnt.coll;
    r->total_steps = cnt.steps;
    return 0;
}

extern "C" int emu_ballistic(const dsmc_params *p, const dsmc_surface *s, dsmc_result *r)
{
    DsmcEnv E;
    dsmc::HostTables T;
    int rc = dsmc::build_env(p, s, 0, &E, &T);
    if (rc) return rc;
    size_t cells = (size_t)p->hd * p->ftd * p->ffd;
    std::vector<double> den(cells, 0.0), hor((size_t)p->ftd * p->ffd, 0.0);
    std::vector<unsigned long long> vcount(cells, 0);
    std::vector<int32_t> count_nn((size_t)p->ftd * p->hd, 0);
    unsigned long long rec_total = 0;
    E.T = s->T_latlon;
    E.vt_speeds = T.vt_speeds.data(); E.vt_count = T.vt_count.data();
    E.tq = T.tq.data(); E.hr = T.hr.data(); E.cs = T.cs.data();
    E.cdf = T.cdf.data(); E.face_T = s->face_T; E.vert = s->vert; E.faces = s->faces;
    E.enorm = T.enorm.data();
    E.density = den.data(); E.horizon = hor.data();
    E.vmom = r->vmom;
    if (r->vmom) memset(r->vmom, 0, cells * 3 * sizeof(double));
    E.vcount = r->vcount ? vcount.data() : nullptr;
    E.ttime = p->mode == DSMC_MODE_COMPAT ? r->ttime : nullptr;
    E.ttime_rows = 5;
    if (E.ttime) memset(E.ttime, 0, sizeof(double) * 5 * p->nparticles);
    E.end_state = r->end_state; E.steps_out = r->steps; E.fate_out = r->fate; E.hops_out = r->hops;
    E.vrec = r->vrec; E.vrec_td = r->vrec_td; E.vrec_fd = r->vrec_fd; E.vrec_ht = r->vrec_ht;
    E.count_nn = count_nn.data(); E.rec_total = &rec_total;
    DsmcCounts cnt;
    memset(&cnt, 0, sizeof cnt);
    for (int64_t n = 0; n < p->nparticles; n++) {
        DsmcLane L;
        memset(&L, 0, sizeof L);
        if (p->mode == DSMC_MODE_COMPAT) run_one<0, 0>(L, E, n, cnt);
        else run_one<1, 0>(L, E, n, cnt);
    }
    if (r->density) memcpy(r->density, den.data(), cells * sizeof(double));
    if (r->horizon) memcpy(r->horizon, hor.data(), hor.size() * sizeof(double));
    if (r->vcount) for (size_t i = 0; i < cells; i++) r->vcount[i] = (int64_t)vcount[i];
    r->vrec_n = (int64_t)(rec_total < (unsigned long long)p->record_cap_total ? rec_total : p->record_cap_total);
    r->n_escaped = cnt.esc; r->n_frozen = cnt.frozen; r->n_photolysed = cnt.photo;
    r->n_depleted = cnt.depleted; r->n_capped = cnt.capped; r->n_collisions = cnt.coll;
    r->total_steps = cnt.steps;
    return 0;
}
