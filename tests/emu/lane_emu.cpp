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

/* The same particle through the device kernels' record traffic: every launch
 * starts from a lane filled with garbage (as registers are) and loads what the
 * kernel loads.  SPLIT: start -> record; flight (rec_load_flight, at most
 * `budget` steps, rec_store_flight); service (rec_load_all, lane_service,
 * rec_store_all).  INLINE (physical): start -> record; launches of
 * rec_load_phys, steps with the service inline, rec_store_phys on a pause. */
/* The face map, as the library attaches it where a physical surface is uploaded. */
static void attach_face_map(DsmcEnv &E, dsmc::HostTables &T)
{
    if (E.mode != DSMC_MODE_PHYSICAL || E.nfine <= 0) return;
    E.fmap_g = dsmc::face_map_grid(E.nfine);
    dsmc::build_face_map(E, T, E.fmap_g, T.fmap);
    E.fmap = T.fmap.data();
}

static void garbage(DsmcLane &L) { memset(&L, 0xA5, sizeof L); }

template <int MODE, int ADV>
static void run_one_split(const DsmcEnv &E, int64_t n, DsmcCounts &cnt, int64_t budget)
{
    DsmcRec R;
    memset(&R, 0x5A, sizeof R);
    DsmcLane L;
    garbage(L);
    lane_start<DsmcOpsHost, MODE, ADV>(L, E, n, cnt);
    if (L.phase != LANE_FLIGHT) return;
    rec_store_all<MODE>(L, R, 0);
    DsmcCellTabs TB = {E.tq, E.hr, E.cs};
    for (;;) {
        DsmcLane F;
        garbage(F);
        rec_load_flight<MODE, ADV>(F, R);
        const int64_t end = F.steps + budget;
        bool paused = false;
        while (F.phase == LANE_FLIGHT) {
            lane_step<DsmcOpsHost, MODE, ADV>(F, E, TB);
            if (F.phase == LANE_FLIGHT && F.steps >= end) { paused = true; break; }
        }
        if (!paused || MODE == 0) lane_flush<DsmcOpsHost>(F, E);   /* physical pauses keep the visit */
        rec_store_flight<MODE, ADV>(F, R);
        if (paused) continue;
        DsmcLane S;
        garbage(S);
        rec_load_all<MODE>(S, R);
        lane_service<DsmcOpsHost, MODE, ADV>(S, E, cnt);
        if (S.phase == LANE_DONE) return;
        rec_store_all<MODE>(S, R, 0);
    }
}

/* The event-batched physical flight kernel (lean_flight_kernel): lean steps
 * until one needs the cell logic, then lane_tail_phys and a reload of the cell
 * entries; records and the service kernel as the split path. */
static void run_one_lean(const DsmcEnv &E, int64_t n, DsmcCounts &cnt, int64_t budget)
{
    DsmcRec R;
    memset(&R, 0x5A, sizeof R);
    DsmcLane L;
    garbage(L);
    lane_start<DsmcOpsHost, 1, 0>(L, E, n, cnt);
    if (L.phase != LANE_FLIGHT) return;
    rec_store_all<1>(L, R, 0);
    DsmcCellTabs TB = {E.tq, E.hr, E.cs};
    const int64_t cap = E.max_steps > 0 ? E.max_steps : INT64_MAX;
    for (;;) {
        DsmcLane F;
        garbage(F);
        rec_load_flight<1, 0>(F, R);
        const int64_t end = F.steps + budget;
        LeanLane N;
        N.stop = end < cap ? end : cap;
        N.pend = false;
        lean_refresh(N, F, E, TB);
        bool paused = false;
        while (F.phase == LANE_FLIGHT) {
            if (lane_fast_phys(F, E, N)) continue;
            lean_tail<DsmcOpsHost>(F, E, TB, N);
            if (F.phase != LANE_FLIGHT) break;
            if (F.steps >= end) { paused = true; break; }
            lean_refresh(N, F, E, TB);
        }
        if (!paused) lane_flush<DsmcOpsHost>(F, E);   /* pauses keep the visit */
        rec_store_flight<1, 0>(F, R);
        if (paused) continue;
        DsmcLane S;
        garbage(S);
        rec_load_all<1>(S, R);
        lane_service<DsmcOpsHost, 1, 0>(S, E, cnt);
        if (S.phase == LANE_DONE) return;
        rec_store_all<1>(S, R, 0);
    }
}

/* late_inline: a particle's first two flight launches are the event-batched
 * flight kernel's (the service kernel after the first), every later one the
 * flight kernel with the service inline, lean steps and no budget, which takes
 * the particle over from the record the first two left (paused in flight with
 * its open visit, relaunched by the service kernel, or with its hop ended in
 * the second launch, which leaves its landings to the inline launch). */
static void run_one_hybrid(const DsmcEnv &E, int64_t n, DsmcCounts &cnt, int64_t budget)
{
    DsmcRec R;
    memset(&R, 0x5A, sizeof R);
    DsmcLane L;
    garbage(L);
    lane_start<DsmcOpsHost, 1, 0>(L, E, n, cnt);
    if (L.phase != LANE_FLIGHT) return;
    rec_store_all<1>(L, R, 0);
    DsmcCellTabs TB = {E.tq, E.hr, E.cs};
    const int64_t cap = E.max_steps > 0 ? E.max_steps : INT64_MAX;
    for (int round = 0; round < 2; round++) {
        DsmcLane F;
        garbage(F);
        rec_load_flight<1, 0>(F, R);
        const int64_t end = F.steps + budget;
        LeanLane N;
        N.stop = end < cap ? end : cap;
        N.pend = false;
        lean_refresh(N, F, E, TB);
        bool paused = false;
        while (F.phase == LANE_FLIGHT) {
            if (lane_fast_phys(F, E, N)) continue;
            lean_tail<DsmcOpsHost>(F, E, TB, N);
            if (F.phase != LANE_FLIGHT) break;
            if (F.steps >= end) { paused = true; break; }
            lean_refresh(N, F, E, TB);
        }
        if (!paused) lane_flush<DsmcOpsHost>(F, E);
        rec_store_flight<1, 0>(F, R);
        if (paused || round == 1) continue;
        DsmcLane S;
        garbage(S);
        rec_load_all<1>(S, R);
        lane_service<DsmcOpsHost, 1, 0>(S, E, cnt);
        if (S.phase == LANE_DONE) return;
        rec_store_all<1>(S, R, 0);
    }
    DsmcLane F;
    garbage(F);
    rec_load_phys(F, R);
    LeanLane N;
    N.stop = cap;
    N.pend = false;
    lean_refresh(N, F, E, TB);
    for (;;) {
        if (F.phase == LANE_FLIGHT && lane_fast_phys(F, E, N)) continue;
        if (F.phase != LANE_FLIGHT) {      /* a hop that ended in the second launch */
            DsmcLane T = F;
            lane_service<DsmcOpsHost, 1, 0>(T, E, cnt);
            F = T;
            if (F.phase == LANE_DONE) return;
            lean_refresh(N, F, E, TB);
            continue;
        }
        lean_tail<DsmcOpsHost>(F, E, TB, N);
        if (F.phase == LANE_FLIGHT) { lean_refresh(N, F, E, TB); continue; }
        DsmcLane T = F;
        lane_service<DsmcOpsHost, 1, 0>(T, E, cnt);
        F = T;
        if (F.phase == LANE_DONE) return;
        lean_refresh(N, F, E, TB);
    }
}

template <int MODE, int ADV>
static void run_one_inline(const DsmcEnv &E, int64_t n, DsmcCounts &cnt, int64_t budget)
{
    DsmcRec R;
    memset(&R, 0x5A, sizeof R);
    DsmcLane L;
    garbage(L);
    lane_start<DsmcOpsHost, MODE, ADV>(L, E, n, cnt);
    if (L.phase != LANE_FLIGHT) return;
    rec_store_all<MODE>(L, R, 0);
    DsmcCellTabs TB = {E.tq, E.hr, E.cs};
    for (;;) {
        DsmcLane F;
        garbage(F);
        if (MODE == 0) rec_load_all<0>(F, R);
        else rec_load_phys(F, R);
        const int64_t end = F.steps + budget;
        for (;;) {
            bool paused = false;
            lane_step<DsmcOpsHost, MODE, ADV>(F, E, TB);
            if (F.phase == LANE_FLIGHT && F.steps >= end) paused = true;
            if (F.phase == LANE_LANDED || F.phase == LANE_ENDED || F.phase == LANE_COLLIDE) {
                DsmcLane T = F;
                lane_service<DsmcOpsHost, MODE, ADV>(T, E, cnt);
                F = T;
                if (F.phase == LANE_DONE) return;
            }
            if (paused && MODE == 0) {     /* compat: the visit is flushed */
                lane_flush<DsmcOpsHost>(F, E);
                rec_store_all<0>(F, R, 0);
                break;
            }
            if (paused) {                  /* the open visit travels with the record */
                rec_store_phys(F, R);
                break;
            }
        }
    }
}

/* Advanced (collisional) compat run with fits supplied by the caller. */
/* One particle at a time straight through the state machine. */
static int advanced_run(const dsmc_params *p, const dsmc_surface *s, const double *copy_atm, int64_t atm_total,
                        const int32_t *count_num, const double *coeffs, const double *pmin, const double *pmax,
                        const int32_t *pvalid, dsmc_result *r)
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

extern "C" int emu_advanced(const dsmc_params *p, const dsmc_surface *s, const double *copy_atm,
                            int64_t atm_total, const int32_t *count_num, const double *coeffs,
                            const double *pmin, const double *pmax, const int32_t *pvalid, dsmc_result *r)
{
    return advanced_run(p, s, copy_atm, atm_total, count_num, coeffs, pmin, pmax, pvalid, r);
}

/* path 0: one particle at a time straight through the state machine; 1: the
 * split flight / service kernels' record traffic; 2: the physical flight kernel
 * with the service inline; 3: the event-batched physical flight kernel (lean
 * steps) with the service kernel; 4: path 3 for two launches, then the flight
 * kernel with the service inline (late_inline).  `budget`: steps per flight
 * launch (paths 1-3; path 4's first two launches). */
extern "C" int emu_ballistic_path(const dsmc_params *p, const dsmc_surface *s, dsmc_result *r, int path,
                                  int64_t budget)
{
    if (path < 0 || path > 4 || (path && budget < 1) || (path >= 3 && p->mode != DSMC_MODE_PHYSICAL)) return -1;
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
    E.cdf = T.cdf.data(); E.face_T = s->face_T; E.face_sig = T.face_sig.data(); E.vert = s->vert; E.faces = s->faces;
    E.enorm = T.enorm.data();
    E.fnorm = T.fnorm.data();
    attach_face_map(E, T);
    E.density = den.data(); E.horizon = hor.data();
    E.vmom = r->vmom;
    if (r->vmom) memset(r->vmom, 0, cells * 3 * sizeof(double));
    /* physical runs: density = visit count, as the library computes it */
    E.density_from_count = p->mode == DSMC_MODE_PHYSICAL;
    E.vcount = r->vcount || E.density_from_count ? vcount.data() : nullptr;
    E.ttime = p->mode == DSMC_MODE_COMPAT ? r->ttime : nullptr;
    E.ttime_rows = 5;
    if (E.ttime) memset(E.ttime, 0, sizeof(double) * 5 * p->nparticles);
    E.end_state = r->end_state; E.steps_out = r->steps; E.fate_out = r->fate; E.hops_out = r->hops;
    E.vrec = r->vrec; E.vrec_td = r->vrec_td; E.vrec_fd = r->vrec_fd; E.vrec_ht = r->vrec_ht;
    E.count_nn = count_nn.data(); E.rec_total = &rec_total;
    std::vector<unsigned long long> vh;
    if (r->vhist && p->mode == DSMC_MODE_PHYSICAL) {
        vh.assign(cells * p->vhist_bins, 0);
        E.vhist = vh.data();
        E.vhist_bins = p->vhist_bins;
        E.vhist_inv = (double)p->vhist_bins / p->vhist_max;
    }
    DsmcCounts cnt;
    memset(&cnt, 0, sizeof cnt);
    for (int64_t n = 0; n < p->nparticles; n++) {
        DsmcLane L;
        memset(&L, 0, sizeof L);
        if (path == 1 && p->mode == DSMC_MODE_COMPAT) run_one_split<0, 0>(E, n, cnt, budget);
        else if (path == 1) run_one_split<1, 0>(E, n, cnt, budget);
        else if (path == 2 && p->mode == DSMC_MODE_COMPAT) run_one_inline<0, 0>(E, n, cnt, budget);
        else if (path == 2) run_one_inline<1, 0>(E, n, cnt, budget);
        else if (path == 3) run_one_lean(E, n, cnt, budget);
        else if (path == 4) run_one_hybrid(E, n, cnt, budget);
        else if (p->mode == DSMC_MODE_COMPAT) run_one<0, 0>(L, E, n, cnt);
        else run_one<1, 0>(L, E, n, cnt);
    }
    if (r->density && E.density_from_count) for (size_t i = 0; i < cells; i++) r->density[i] = (double)vcount[i];
    else if (r->density) memcpy(r->density, den.data(), cells * sizeof(double));
    if (r->horizon) memcpy(r->horizon, hor.data(), hor.size() * sizeof(double));
    if (r->vcount) for (size_t i = 0; i < cells; i++) r->vcount[i] = (int64_t)vcount[i];
    if (E.vhist) for (size_t i = 0; i < vh.size(); i++) r->vhist[i] = (int64_t)vh[i];
    r->vrec_n = (int64_t)(rec_total < (unsigned long long)p->record_cap_total ? rec_total : p->record_cap_total);
    r->n_escaped = cnt.esc; r->n_frozen = cnt.frozen; r->n_photolysed = cnt.photo;
    r->n_depleted = cnt.depleted; r->n_capped = cnt.capped; r->n_collisions = cnt.coll;
    r->total_steps = cnt.steps;
    return 0;
}

extern "C" int emu_ballistic(const dsmc_params *p, const dsmc_surface *s, dsmc_result *r)
{
    return emu_ballistic_path(p, s, r, 0, 0);
}

/* Physical collisional phase 2 on the lane state machine with the caller's
 * fits and n*sigma table: path 0 straight, 2 the flight kernel with the service
 * inline (budget steps per launch). */
extern "C" int emu_physical_coll(const dsmc_params *p, const dsmc_surface *s, const double *nsig,
                                 const double *coeffs, const double *pmin, const double *pmax, dsmc_result *r,
                                 int path, int64_t budget)
{
    if (p->mode != DSMC_MODE_PHYSICAL || (path != 0 && path != 2) || (path == 2 && budget < 1)) return -1;
    DsmcEnv E;
    dsmc::HostTables T;
    int rc = dsmc::build_env(p, s, 1, &E, &T);
    if (rc) return rc;
    size_t cells = (size_t)p->hd * p->ftd * p->ffd;
    std::vector<double> den(cells, 0.0), hor((size_t)p->ftd * p->ffd, 0.0);
    std::vector<unsigned long long> vcount(cells, 0);
    E.tq = T.tq.data(); E.hr = T.hr.data(); E.cs = T.cs.data();
    E.cdf = T.cdf.data(); E.face_T = s->face_T; E.face_sig = T.face_sig.data(); E.vert = s->vert; E.faces = s->faces;
    E.enorm = T.enorm.data(); E.fnorm = T.fnorm.data();
    attach_face_map(E, T);
    E.density = den.data(); E.horizon = hor.data();
    E.vmom = r->vmom;
    if (r->vmom) memset(r->vmom, 0, cells * 3 * sizeof(double));
    E.vcount = vcount.data();
    E.density_from_count = 1;
    E.end_state = r->end_state; E.steps_out = r->steps; E.fate_out = r->fate; E.hops_out = r->hops;
    E.nsig = nsig; E.coeffs = coeffs; E.pmin = pmin; E.pmax = pmax; E.fit_deg = p->fit_degree;
    DsmcCounts cnt;
    memset(&cnt, 0, sizeof cnt);
    for (int64_t n = 0; n < p->nparticles; n++) {
        DsmcLane L;
        memset(&L, 0, sizeof L);
        if (path == 2) run_one_inline<1, 1>(E, n, cnt, budget);
        else run_one<1, 1>(L, E, n, cnt);
    }
    if (r->density) for (size_t i = 0; i < cells; i++) r->density[i] = (double)vcount[i];
    if (r->horizon) memcpy(r->horizon, hor.data(), hor.size() * sizeof(double));
    if (r->vcount) for (size_t i = 0; i < cells; i++) r->vcount[i] = (int64_t)vcount[i];
    r->n_escaped = cnt.esc; r->n_frozen = cnt.frozen; r->n_photolysed = cnt.photo;
    r->n_depleted = cnt.depleted; r->n_capped = cnt.capped; r->n_collisions = cnt.coll;
    r->total_steps = cnt.steps;
    return 0;
}

/* Face location of n directions: out_near = locate_face_near (hint, then the
 * face map, then the search), out_full = the hierarchical search alone. */
extern "C" int emu_locate(const dsmc_params *p, const dsmc_surface *s, int64_t n, const double *dirs,
                          const int32_t *hints, int32_t *out_near, int32_t *out_full)
{
    if (p->mode != DSMC_MODE_PHYSICAL) return -1;
    DsmcEnv E;
    dsmc::HostTables T;
    int rc = dsmc::build_env(p, s, 0, &E, &T);
    if (rc) return rc;
    E.enorm = T.enorm.data();
    E.fnorm = T.fnorm.data();
    attach_face_map(E, T);
    for (int64_t i = 0; i < n; i++) {
        const double *d = dirs + 3 * i;
        const double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        out_near[i] = locate_face_near(E, hints ? hints[i] : -1, d[0], d[1], d[2], len);
        out_full[i] = locate_face(E, d[0], d[1], d[2]);
    }
    return 0;
}

/* Face location of n directions as the pool kernels make it: no guess, the
 * first four candidates' normals read from the face-map cell (fmap_n). */
extern "C" int emu_locate_pool(const dsmc_params *p, const dsmc_surface *s, int64_t n, const double *dirs,
                               int32_t *out)
{
    if (p->mode != DSMC_MODE_PHYSICAL) return -1;
    DsmcEnv E;
    dsmc::HostTables T;
    int rc = dsmc::build_env(p, s, 0, &E, &T);
    if (rc) return rc;
    E.enorm = T.enorm.data();
    E.fnorm = T.fnorm.data();
    attach_face_map(E, T);
    std::vector<double> fmn;
    dsmc::build_face_map_normals(T.fmap, T.fnorm, E.fmap_g, fmn);
    E.fmap_n = fmn.data();
    for (int64_t i = 0; i < n; i++) {
        const double *d = dirs + 3 * i;
        const double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        out[i] = locate_face_near<true>(E, -1, d[0], d[1], d[2], len);
    }
    return 0;
}

/* Which step of locate_face_near resolves each of n directions (no hint):
 * counts[0] the first four face-map candidates, [1] the next four, [2] the
 * hierarchical search (no candidate holds the direction firmly). */
extern "C" int emu_locate_paths(const dsmc_params *p, const dsmc_surface *s, int64_t n, const double *dirs,
                                int64_t *counts)
{
    if (p->mode != DSMC_MODE_PHYSICAL) return -1;
    DsmcEnv E;
    dsmc::HostTables T;
    int rc = dsmc::build_env(p, s, 0, &E, &T);
    if (rc) return rc;
    E.fnorm = T.fnorm.data();
    attach_face_map(E, T);
    counts[0] = counts[1] = counts[2] = 0;
    for (int64_t i = 0; i < n; i++) {
        const double *d = dirs + 3 * i;
        const double eps = 1e-12 * sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        const int cell = fmap_cell(E.fmap_g, d[0], d[1], d[2]);
        int found = -1;
        for (int j = 0; cell >= 0 && j < DSMC_FMAP_K && E.fmap[(size_t)cell * DSMC_FMAP_K + j] >= 0; j++)
            if (face_holds(E, E.fmap[(size_t)cell * DSMC_FMAP_K + j], d[0], d[1], d[2], eps)) { found = j; break; }
        counts[found < 0 ? 2 : (found < 4 ? 0 : 1)]++;
    }
    return 0;
}
