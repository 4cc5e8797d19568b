/* pool.h - multi-run pools (internal): many independent physical runs share one
 * set of device particle slots, queues and persistent kernels.  Used by the
 * batch API (all runs admitted at once) and by the serving loop (runs admitted
 * while others fly, as slots and memory free up). */
#ifndef DSMC_POOL_H
#define DSMC_POOL_H

#include <hip/hip_runtime.h>
#include <map>
#include <vector>
#include "dsmc_host.h"

namespace dsmc {

/* Runs of one pool share these. */
struct PoolClass {
    int mode, adv, ftd, ffd, hd, levels;
};

struct PoolRun {
    int used = 0, done = 0, cancelled = 0;
    int kind = 0;              /* 0: ballistic (optionally keeping histograms), 1: collisional */
    int hist = -1, coll = -1;  /* histogram slot (phase 1) / fit-table slot (phase 2) */
    int moments = 0;           /* the caller takes velocity moments (every run deposits cell records) */
    dsmc_params p;
    DsmcEnv E;
    int64_t slot0 = 0, n = 0;
    int hr_tab = -1, geo = -1;
    uint64_t tag = 0;          /* caller's id (serving ticket) */
    int64_t first_seq = 0;     /* first replay enqueued after admission: earlier
                                  counter snapshots describe the slot's last run */
};

struct PoolGeo {
    uint64_t key = 0;
    int ref = 0, valid = 0;
};

struct Pool {
    PoolClass cls;
    int max_runs = 0, per_particle = 0;
    int64_t max_slots = 0;
    size_t cells = 0, nfine = 0, nvert = 0, nface_all = 0;
    hipStream_t st = nullptr, cst = nullptr;   /* rounds / result copies */
    hipEvent_t ring_ev[2] = {nullptr, nullptr}; /* end of a replay + its counter copy */
    hipEvent_t last_ev = nullptr;              /* replay whose counters were read last */
    int pending = 0;                           /* replays enqueued, counters not read */
    int64_t issued = 0, consumed = 0;
    int cus = 256;

    DevBuf rec, flyq0, flyq1, svcq, qctl;
    DevBuf cflyq0, cflyq1, cqctl;              /* queues of collisional particles */
    /* physical collisional requests: phase-1 speed histograms and phase-2 fit
     * tables, POOL_MAX_HIST / POOL_MAX_COLL slots (allocated when enabled) */
    int hist_bins = 0, fit_deg = 0;
    DevBuf hist;                               /* [max_hist][cells*hist_bins] */
    DevBuf coll_cf, coll_lo, coll_hi, coll_nsig, coll_valid, coll_cnt;
    std::vector<int> hist_used, coll_used;
    DevBuf envs_d, steps_d, tq, cs, hr;
    DevBuf geo_vert, geo_faces, geo_enorm, geo_fnorm, geo_fmap, geo_fmapn;   /* [max_geo] slots */
    int fmap_g = 0;                            /* face-map resolution of the class */
    DevBuf cdf, face_T, face_sig;              /* [max_runs][nfine] */
    DevBuf vcount, cellacc, horizon, counters; /* [max_runs] grid slots (cellacc: 32-byte cell records) */
    DevBuf end_state, steps_o, fate, hops;     /* [max_slots], optional */

    DsmcEnv *envs_h = nullptr;                 /* pinned host copies */
    DsmcStepEnv *steps_h = nullptr;
    double *stage_h = nullptr;                 /* [max_runs][3][nfine] */
    unsigned long long *cnt_h = nullptr;       /* [max_runs*8] */
    unsigned int *ctl_h = nullptr;             /* [16] */
    void *out_h = nullptr;                     /* result staging */
    size_t out_bytes = 0;

    std::vector<PoolRun> runs;
    std::map<int64_t, int64_t> free_slots;     /* start -> length */
    std::vector<std::vector<double>> hr_tabs;
    std::vector<int> hr_ref;
    std::vector<PoolGeo> geos;

    dsmc_launch_shape shape;
    int graph_rounds = 8, use_graph = 1;
    /* batches (every run admitted before the first round): the rounds after the
     * first of each replay run without a step budget, so the particles paused by
     * the first round fly to their ends in one launch instead of one launch per
     * budget, each with its own tail of a few long particles */
    int late_unbounded = 0;
    hipGraphExec_t exec = nullptr;
    int64_t rounds = 0;
    int64_t inflight = 0;                      /* particles of runs not done */
    ~Pool();
};

enum { POOL_MAX_HR_TABS = 16, POOL_MAX_GEO = 4, POOL_MAX_HIST = 8, POOL_MAX_COLL = 8 };
/* largest run table the pool kernels keep in LDS (else they read it from device memory) */
#define POOL_STEPS_LDS_BYTES 16384

/* Footprint in device memory of one particle slot and of one run's grid. */
size_t pool_slot_bytes(int per_particle);
size_t pool_run_bytes(const PoolClass &c);

/* A run prepared on the host (environment and tables), ready to admit. */
struct PoolPrep {
    int kind = 0;              /* 1: collisional phase 2 */
    dsmc_params p;
    DsmcEnv E;
    HostTables T;
    std::vector<double> face_T;
    std::vector<double> vert;
    std::vector<int32_t> faces;
    uint64_t geo_key = 0;
    int moments = 1;           /* velocity moments per cell (3 more atomics per cell visit) */
};

int pool_class_of(const dsmc_params *p, const dsmc_surface *s, PoolClass *c);
int pool_prepare(const dsmc_params *p, const dsmc_surface *s, PoolPrep *out, int collisional = 0);
/* Histogram and fit-table slots for collisional requests (bins <= 64, degree <= 6). */
int pool_enable_collisional(Pool *P, int bins, int degree);
int pool_take_slot(std::vector<int> &used);    /* -1 if none free */
int pool_create(Pool *P, const PoolClass &c, int max_runs, int64_t max_slots, int per_particle,
                const dsmc_launch_shape *shape);
/* Admits a run: DSMC_OK (*g = its run slot), DSMC_ERR_BUSY if slots, grid slots
 * or table slots are not free now, or another status. */
/* hist: histogram slot for a ballistic run that keeps them (or -1); coll: fit-
 * table slot of a collisional run (required for kind 1). */
int pool_admit(Pool *P, const PoolPrep &pr, uint64_t tag, int *g, int hist = -1, int coll = -1);
/* Builds fit table `coll` from histogram slot `hist` and the number density
 * ndens (host array, [cells]) on the pool's stream; parameters of the phase-2 run. */
int pool_build_fits(Pool *P, int hist, int coll, const double *ndens, const dsmc_params *p2);
/* Enqueues `chunks` graph replays of the pool's rounds (at most 2 unread). */
int pool_advance(Pool *P, int chunks);
/* Waits for the oldest unread replay, reads the run counters it left, marks
 * finished runs.  Returns the number of runs newly finished (or a status). */
int pool_poll(Pool *P, std::vector<int> *finished);
/* Copies run g's results (the outputs `r` asks for) and its counters. */
int pool_collect(Pool *P, int g, dsmc_result *r);
/* Frees a finished run's slots and grid slot (a run still in flight cannot be
 * released: its particles hold their slots until they finish). */
int pool_release(Pool *P, int g);
int pool_set_shape(Pool *P, const dsmc_launch_shape &s);
/* Blocks of `block` threads of the pool's flight kernel one CU holds at once. */
int pool_occupancy(Pool *P, int block);

}  // namespace dsmc

#endif
