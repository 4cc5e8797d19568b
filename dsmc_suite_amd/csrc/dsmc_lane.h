/*
 * dsmc_lane.h - the per-particle state machine of the DSMC path, written once
 * for the gfx950 kernels (one particle per lane) and for a test-only host build.
 *
 * A particle alternates between FLIGHT (one Stormer-Verlet step per call of
 * lane_step) and the hop service (landing bookkeeping + relaunch, lane_service),
 * so the kernels can run the cheap, uniform flight step for a whole wavefront and
 * batch the divergent relaunch work.  Semantics follow the reference drivers:
 *   compat ballistic  src/dsmc/ballistic.c:102-281
 *   compat advanced   src/dsmc/advanced.c:241-450 (collision branch :369-398)
 *   physical          the benchmark exosphere model (icosphere surface, flux-
 *                     Maxwellian emission, freeze-out, Monte-Carlo photolysis)
 * and every floating-point expression is evaluated in the reference's order with
 * no contraction (the translation units are built with -ffp-contract=off), so a
 * particle's trajectory is bit-identical to the oracle's device-math build.
 *
 * Per-step cell indices avoid transcendentals:
 *   td  latitude: q = z/Ri compared against a host-built table tq[k] = smallest
 *       q with ceil((asin(q)+pi/2)/pi*ftd)-1 >= k (libm asin, grid.c:28-39);
 *   ht  radius: Ri compared against hr[k] = smallest Ri whose height index is
 *       >= k (uniform: ballistic.c:211, exponential: grid.c:91-99 with libm log);
 *   fd  longitude: kept while the point is farther than 1e-12 (relative) from
 *       both sector boundaries (a sign test against the boundary directions),
 *       otherwise recomputed as ceil(atan2(x,y)/2pi*ffd)-1 with dm_atan2.
 */
#ifndef DSMC_LANE_H
#define DSMC_LANE_H

#include <stdint.h>
#include <math.h>
#include "dsmc_math.h"
#include "philox.h"

#if defined(__HIPCC__)
#define DSMC_HD __host__ __device__ inline
#else
#define DSMC_HD inline
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* FLIGHT: in a hop; LANDED: hop ended on the surface (or the per-hop limit);
 * ENDED: the particle's fate was decided during a step and its outputs are
 * still to be written by the service; DONE: finished; COLLIDE: a physical
 * collisional particle reached its collision depth (service scatters it). */
enum { LANE_IDLE = 0, LANE_FLIGHT = 1, LANE_LANDED = 2, LANE_DONE = 3, LANE_ENDED = 4, LANE_COLLIDE = 5 };
enum { DSMC_VIRTUAL_RECORDS = 50 };   /* physical fits: P(vmin) = 50, as the reference's 50 records */
enum { DSMC_VT_NT = 23, DSMC_VT_NV = 1000 };
#define DSMC_FD_MARGIN 1e-12
enum { DSMC_FMAP_K = 8 };             /* candidate faces per face-map cell */

/* Everything a lane reads or writes; device pointers on the GPU. */
struct DsmcEnv {
    /* run scalars */
    int mode, ftd, ffd, hd, exp_grid, advanced;
    int64_t nparticles, max_hop_steps, max_steps;
    uint64_t seed;
    uint32_t stream;
    double GM, R, h, c1, c2, decay, dt_rot, ve, max_height, w_min, t_freeze, alpha, period;
    double mass, k_B, lifetime, sun_x, sun_y, omega, unih, alpha2;
    /* tables */
    const double *T;          /* [ftd*ffd] surface temperature (compat) */
    const double *vt_speeds;  /* [23*1000] velocity table (velocity.c:6-47) */
    const int32_t *vt_count;  /* [23] */
    const double *tq;         /* [ftd+1] latitude thresholds on z/Ri */
    const double *hr;         /* [hd+2] radial thresholds */
    const double *cs;         /* [(ffd+1)*2] (cos, sin) of sector boundary k*2pi/ffd */
    /* physical surface */
    int levels, nfine;
    const double *cdf;        /* [nfine] inclusive prefix sum of launch weights */
    const double *face_T;     /* [nfine] */
    const double *face_sig;   /* [nfine] sqrt(k_B face_T / mass): the emission's thermal speed at each face */
    const double *vert;       /* [nvert*3] */
    const int32_t *faces;     /* all levels, [sum*3] */
    const double *enorm;      /* all levels, [sum*9]: cross(a,b), cross(b,c), cross(c,a) */
    const double *fnorm;      /* finest level, [nfine*9]: the same normals, unit length */
    const int32_t *fmap;      /* [6*G*G*DSMC_FMAP_K] candidate finest faces per cube-map cell, -1 = none */
    int fmap_g;               /* G; fmap == nullptr: no map */
    /* [6*G*G][4][9]: fnorm's rows of each cell's first four candidates (a -1
     * candidate: the first one's), so the first test needs no face index;
     * nullptr: the candidates' rows are read from fnorm */
    const double *fmap_n;
    int level_off[12];
    /* advanced inputs */
    const double *copy_atm;   /* scaled phase-1 density, [atm_total] */
    int64_t atm_total;
    double sigma;
    int fit_deg;
    const int32_t *count_num; /* [ftd*ffd*hd] cells [td][fd][ht] */
    const double *coeffs;     /* [cells*(deg+1)] */
    const double *pmin, *pmax;
    const int32_t *pvalid;
    /* physical collisional: n*sigma per cell [1/m], 0 where no partner fit */
    const double *nsig;
    /* physical phase 1: steps per speed bin of every cell visit, [cell][vhist_bins] */
    unsigned long long *vhist;
    int vhist_bins;
    double vhist_inv;         /* vhist_bins / vhist_max */
    /* outputs */
    int density_from_count;   /* physical mode with counts: W == 1, density = vcount */
    double *density, *horizon, *vmom;
    unsigned long long *vcount;
    /* physical runs with moments: every deposit goes to a 32-byte record per
     * cell (sum vx, sum vy, sum vz, count), [cells*4], folded into vmom and
     * vcount when the run ends; nullptr: deposits go to vmom / vcount */
    double *cellacc;
    double *ttime, *end_state;
    int64_t *steps_out;
    int32_t *fate_out, *hops_out;
    int ttime_rows;
    double *vrec;
    int32_t *vrec_td, *vrec_fd, *vrec_ht;
    int32_t *count_nn;        /* [ftd*hd] record counter per (td, ht) */
    unsigned long long *rec_total;
    int rec_cap_cell;
    int64_t rec_cap_total;
    /* fate/step counters of this run: escaped, frozen, photolysed, depleted,
     * capped, collisions, particle-steps (device kernels add here) */
    unsigned long long *counters;
};

/* What one physical flight step reads, held per lane in registers when the lanes
 * of a wavefront belong to different runs (multi-run pools).  Field names match
 * DsmcEnv so the step is one template for both.  ftd/ffd/hd are the pool's. */
struct DsmcStepEnv {
    double h, GM, c1, c2, R, max_height, sun_x, sun_y;
    int64_t max_steps;
    int ftd, ffd, hd, density_from_count;
    double *density, *vmom;
    unsigned long long *vcount;
    double *cellacc;          /* moments runs: per-cell records (as DsmcEnv::cellacc) */
    int hr_off;               /* this run's radial thresholds within the pool's table */
    int pad_;
    const double *nsig;       /* collisional runs: n*sigma per cell */
    const DsmcEnv *env;       /* the run's full environment (device memory) */
};

struct DsmcLane {
    double sx, sy, sz, vx, vy, vz, ox, oy, oz;
    double Ri, W, time, flight_time, thi, fi;
    double acc_w, acx, acy, acz;
    double real_col_p, col_p, rest_time, col_lo;
    double lsx, lsy, lsz;            /* advanced: position before a frozen-scan rotation */
    int64_t steps, ii, n_day, n_photo, pid;
    uint64_t k, blk;
    uint32_t rb0, rb1, rb2, rb3;
    int phase, fate, hops, num_col, td, fd, ht, acc_cell, acc_n, rec_key, rec_full, rotated;
    int face;                        /* physical: face of the last launch */
    double ncur;                     /* physical collisional: n*sigma of cell ncell */
    int ncell;
};

/* A particle between kernels: one 256-byte record per slot in device memory.
 * Physical records use the compat-only slots for physical state: ii = launch
 * face; rotated, thi, fi, lsx, lsy, lsz = the open deposit visit (cell, weight
 * sum, step count, velocity sums), kept across launches so that a launch
 * boundary never splits a visit; real_col_p, col_p = collision depth and its
 * threshold (collisional runs). */
struct DsmcRec {
    double sx, sy, sz, vx, vy, vz, ox, oy, oz;
    double Ri, W, time, flight_time, thi, fi, rest_time, real_col_p, col_p, col_lo;
    double lsx, lsy, lsz;
    int64_t steps, n_day, n_photo, pid;
    uint64_t k;
    int32_t ii, phase, fate, hops, num_col, td, fd, ht, rotated, run;
};

/* Per-step lookup tables (the kernels keep a copy in LDS). */
struct DsmcCellTabs {
    const double *tq, *hr, *cs;
};

/* Speed-histogram bin of one cell visit (n deposits with velocity sums vx, vy,
 * vz): the speed of the mean velocity, vhist_bins bins up to vhist_max, the
 * last bin open-ended.  NULL when the run keeps no histograms. */
DSMC_HD unsigned long long *vhist_bin(const DsmcEnv &E, int cell, double vx, double vy, double vz, int n)
{
    if (!E.vhist || n <= 0) return nullptr;
    double b = sqrt(vx * vx + vy * vy + vz * vz) / (double)n * E.vhist_inv;
    int bin = b < (double)(E.vhist_bins - 1) ? (int)b : E.vhist_bins - 1;   /* NaN -> last bin */
    return E.vhist + (size_t)cell * E.vhist_bins + bin;
}
DSMC_HD unsigned long long *vhist_bin(const DsmcStepEnv &S, int cell, double vx, double vy, double vz, int n)
{
    return S.env ? vhist_bin(*S.env, cell, vx, vy, vz, n) : nullptr;
}

/* Atomic/plain accumulation, supplied by the caller (GPU atomics or host adds). */
struct DsmcOpsHost {
    template <class EnvT>
    static inline void deposit(const EnvT &E, int cell, double w, double vx, double vy, double vz, int n)
    {
        unsigned long long *hb = vhist_bin(E, cell, vx, vy, vz, n);
        if (hb) *hb += (unsigned long long)n;
        if (!E.density_from_count) E.density[cell] += w;
        if (E.vmom) {
            E.vmom[(size_t)cell * 3 + 0] += vx;
            E.vmom[(size_t)cell * 3 + 1] += vy;
            E.vmom[(size_t)cell * 3 + 2] += vz;
        }
        if (E.vcount) E.vcount[cell] += (unsigned long long)n;
    }
    static inline void add_f64(double *p, double v) { *p += v; }
    static inline void add_u64(unsigned long long *p, unsigned long long v) { *p += v; }
    static inline int32_t claim_i32(int32_t *p) { return (*p)++; }
    static inline unsigned long long claim_u64(unsigned long long *p) { return (*p)++; }
};

/* ------------------------------------------------------------------ RNG */
/* One Philox block; out of line on the device so that the several draws of an
 * emission share one copy of the rounds (and its registers). */
#if defined(__HIP_DEVICE_COMPILE__)
__attribute__((noinline))
#endif
DSMC_HD dsmc_u32x4 lane_block(uint64_t seed, uint64_t pid, uint32_t stream, uint64_t b)
{
    return dsmc_rng_block(seed, pid, stream, b);
}

/* INL: the Philox rounds inline (a caller that is itself out of line and draws
 * several blocks in a row: no call per block); same words either way. */
template <bool INL = false>
DSMC_HD uint32_t lane_word(DsmcLane &L, const DsmcEnv &E)
{
    uint64_t b = L.k >> 2;
    if (b != L.blk) {
        dsmc_u32x4 r = INL ? dsmc_rng_block(E.seed, (uint64_t)L.pid, E.stream, b)
                           : lane_block(E.seed, (uint64_t)L.pid, E.stream, b);
        L.rb0 = r.w[0]; L.rb1 = r.w[1]; L.rb2 = r.w[2]; L.rb3 = r.w[3];
        L.blk = b;
    }
    uint32_t j = (uint32_t)(L.k & 3);
    L.k++;
    return j == 0 ? L.rb0 : j == 1 ? L.rb1 : j == 2 ? L.rb2 : L.rb3;
}
DSMC_HD int lane_draw(DsmcLane &L, const DsmcEnv &E) { return (int)(lane_word(L, E) >> 1); }
template <bool INL = false>
DSMC_HD double lane_u01(DsmcLane &L, const DsmcEnv &E)
{
    uint32_t hi = lane_word<INL>(L, E);
    uint32_t lo = lane_word<INL>(L, E);
    return dsmc_u01_from_words(hi, lo);
}
DSMC_HD double lane_urand(DsmcLane &L, const DsmcEnv &E)   /* (double)rand_r / RAND_MAX */
{
    return (double)lane_draw(L, E) / 2147483647;
}

/* ------------------------------------------------------------------ cells */
/* grid.c:28-39 */
DSMC_HD void sph_index(double thi, double fi, int ftd, int ffd, int *td, int *fd)
{
    *td = (int)ceil((thi + M_PI / 2.0) / M_PI * ftd) - 1;
    *fd = (int)ceil(fi / (2.0 * M_PI) * ffd) - 1;
    if (*td < 0) *td = 0;
    if (*td >= ftd) *td = ftd - 1;
    if (*fd < 0) *fd = 0;
    if (*fd >= ffd) *fd = ffd - 1;
}

/* ballistic.c:29-33 with the device-math elementary functions */
DSMC_HD void pos_angles(double x, double y, double z, double Ri, double *thi, double *fi)
{
    *thi = dm_asin(z / Ri);
    *fi = dm_atan2(x, y);
    if (*fi < 0) *fi += 2.0 * M_PI;
}

/* Out of line on the device: it runs only within 1e-12 of a sector boundary, and
 * inlined it would set the register footprint of the whole flight kernel. */
#if defined(__HIP_DEVICE_COMPILE__)
__attribute__((noinline))
#endif
DSMC_HD int fd_exact(double x, double y, int ffd)
{
    double fi = dm_atan2(x, y);
    if (fi < 0) fi += 2.0 * M_PI;
    int fd = (int)ceil(fi / (2.0 * M_PI) * ffd) - 1;
    if (fd < 0) fd = 0;
    if (fd >= ffd) fd = ffd - 1;
    return fd;
}

/* True if (x, y) is farther than eps from both boundaries of sector k and inside it. */
DSMC_HD bool fd_inside(double x, double y, int k, double eps, const double *cs)
{
    double dlo = x * cs[2 * k] - y * cs[2 * k + 1];
    double dhi = x * cs[2 * k + 2] - y * cs[2 * k + 3];
    return dlo > eps && dhi < -eps;
}

/* Longitude sector of (x, y): the current sector or a neighbour when the point is
 * safely inside it (a sign test against the boundary directions), otherwise the
 * exact ceil(atan2(x, y) / 2pi * ffd) - 1. */
DSMC_HD int fd_update(double x, double y, int fd, int ffd, const double *cs)
{
    double eps = DSMC_FD_MARGIN * (fabs(x) + fabs(y));
    if (fd_inside(x, y, fd, eps, cs)) return fd;
    int up = fd + 1 == ffd ? 0 : fd + 1, dn = fd == 0 ? ffd - 1 : fd - 1;
    if (fd_inside(x, y, up, eps, cs)) return up;
    if (fd_inside(x, y, dn, eps, cs)) return dn;
    return fd_exact(x, y, ffd);
}

/* Latitude index from (z, Ri): cell td, or a neighbour, while z/Ri is farther
 * than 1e-12 (relative) from both of its thresholds (no division: such a margin
 * leaves the walk on q = z/Ri no other answer), else walked on q = z/Ri. */
DSMC_HD int td_update(double z, double Ri, int td, int ftd, const double *tq)
{
    double eps = 1e-12 * Ri;
    const double lo = tq[td] * Ri, hi = tq[td + 1] * Ri;
    if (z - lo > eps && hi - z > eps) return td;
    if (td > 0 && z - tq[td - 1] * Ri > eps && lo - z > eps) return td - 1;
    if (td < ftd - 1 && z - hi > eps && tq[td + 2] * Ri - z > eps) return td + 1;
    double q = z / Ri;
    while (td > 0 && q < tq[td]) td--;
    while (td < ftd - 1 && q >= tq[td + 1]) td++;
    return td;
}

/* Index k with tab[k] <= v < tab[k+1], starting from a nearby guess. */
DSMC_HD int table_walk(const double *tab, int k, int kmax, double v)
{
    while (k > 0 && v < tab[k]) k--;
    while (k < kmax && v >= tab[k + 1]) k++;
    return k;
}

/* The current cell's table entries of a physical step, read before the push so
 * that the table reads are in flight while it computes: latitude thresholds
 * tq[td], tq[td+1], sector directions cs[2fd..2fd+3], radial thresholds
 * hr[ht], hr[ht+1] (ht <= hd; the radial table has hd + 2 entries). */
struct CellPre {
    double tlo, thi, c0, s0, c1, s1, rlo, rhi;
};

DSMC_HD CellPre cell_pre(int td, int fd, int ht, const double *tq, const double *cs, const double *hr)
{
    CellPre P;
    P.tlo = tq[td]; P.thi = tq[td + 1];
    P.c0 = cs[2 * fd]; P.s0 = cs[2 * fd + 1]; P.c1 = cs[2 * fd + 2]; P.s1 = cs[2 * fd + 3];
    P.rlo = hr[ht]; P.rhi = hr[ht + 1];
    return P;
}

/* td_update, fd_update and table_walk with their first test on preloaded
 * entries: each returns the unchanged index exactly when the original's first
 * test (or both loop conditions) would, else the original runs. */
DSMC_HD int td_update_pre(double z, double Ri, int td, int ftd, const double *tq, const CellPre &P)
{
    double eps = 1e-12 * Ri;
    if (z - P.tlo * Ri > eps && P.thi * Ri - z > eps) return td;
    return td_update(z, Ri, td, ftd, tq);
}

DSMC_HD int fd_update_pre(double x, double y, int fd, int ffd, const double *cs, const CellPre &P)
{
    double eps = DSMC_FD_MARGIN * (fabs(x) + fabs(y));
    if (x * P.c0 - y * P.s0 > eps && x * P.c1 - y * P.s1 < -eps) return fd;
    return fd_update(x, y, fd, ffd, cs);
}

/* table_walk keeps k iff (k == 0 or !(v < tab[k])) and (k == kmax or
 * !(v >= tab[k+1])); a NaN v keeps k both ways. */
DSMC_HD int table_walk_pre(const double *tab, int k, int kmax, double v, const CellPre &P)
{
    if ((k == 0 || v >= P.rlo) && (k == kmax || v < P.rhi)) return k;
    return table_walk(tab, k, kmax, v);
}

/* ------------------------------------------------------------------ deposit */
template <class Ops, class EnvT>
DSMC_HD void lane_flush(DsmcLane &L, const EnvT &E)
{
    if (L.acc_cell >= 0 && L.acc_n > 0) Ops::deposit(E, L.acc_cell, L.acc_w, L.acx, L.acy, L.acz, L.acc_n);
    L.acc_cell = -1;
}

/* A physical hop opens its first deposit visit at the launch cell with nothing
 * in it: a first step that stays in that cell adds to it what a visit opened by
 * that step would hold, one that leaves closes it with nothing deposited
 * (lane_flush skips an empty visit), and the lean steps take the hop's first
 * step too. */
template <class EnvT>
DSMC_HD void lane_open_visit(DsmcLane &L, const EnvT &E)
{
    if (L.ht < E.hd) {
        L.acc_cell = (L.ht * E.ftd + L.td) * E.ffd + L.fd;
        L.acc_w = 0.0; L.acx = 0.0; L.acy = 0.0; L.acz = 0.0; L.acc_n = 0;
    }
}

/* WEIGHT: the visit's weight sum is kept (compat runs); physical particles all
 * weigh 1 and their density is the visit count, so they skip it. */
template <class Ops, bool WEIGHT = true, class EnvT>
DSMC_HD void lane_deposit(DsmcLane &L, const EnvT &E, int cell)
{
    if (cell != L.acc_cell) {
        lane_flush<Ops>(L, E);
        L.acc_cell = cell;
        L.acc_w = 0.0; L.acx = 0.0; L.acy = 0.0; L.acz = 0.0; L.acc_n = 0;
    }
    if (WEIGHT) L.acc_w += L.W;
    L.acx += L.vx; L.acy += L.vy; L.acz += L.vz;
    L.acc_n++;
}

/* ballistic.c:215-232: capped velocity records, first come first served per
 * (td, ht) cell (cap ignores fd) and globally. */
template <class Ops>
DSMC_HD void lane_record(DsmcLane &L, const DsmcEnv &E)
{
    int key = L.td * E.hd + L.ht;
    if (key != L.rec_key) { L.rec_key = key; L.rec_full = 0; }
    if (L.rec_full) return;
    int32_t slot = Ops::claim_i32(&E.count_nn[key]);
    if (slot >= E.rec_cap_cell) { L.rec_full = 1; return; }
    unsigned long long g = Ops::claim_u64(E.rec_total);
    if ((int64_t)g >= E.rec_cap_total) { L.rec_full = 1; return; }
    E.vrec[g * 3 + 0] = L.vx;
    E.vrec[g * 3 + 1] = L.vy;
    E.vrec[g * 3 + 2] = L.vz;
    E.vrec_td[g] = L.td;
    E.vrec_fd[g] = L.fd;
    E.vrec_ht[g] = L.ht;
}

/* ------------------------------------------------------------------ end of a particle */
struct DsmcCounts {
    unsigned long long esc, frozen, photo, depleted, capped, coll, steps, done;
};

template <int MODE = 0>
DSMC_HD void lane_finish(DsmcLane &L, const DsmcEnv &E, DsmcCounts &cnt)
{
    int64_t n = L.pid;
    if (MODE == 0 && E.ttime) {
        int64_t N = E.nparticles;
        if (L.fate != 1 /* escaped: rows 1+ stay 0 */) {
            E.ttime[1 * N + n] = L.thi;
            E.ttime[2 * N + n] = L.time;
            E.ttime[3 * N + n] = L.W;
            E.ttime[4 * N + n] = L.flight_time;
            if (E.ttime_rows >= 9) {
                E.ttime[7 * N + n] = L.rest_time;
                E.ttime[8 * N + n] = (double)L.num_col;
            }
        }
    }
    if (E.fate_out) E.fate_out[n] = L.fate;
    if (E.steps_out) E.steps_out[n] = L.steps;
    if (E.hops_out) E.hops_out[n] = L.hops;
    if (E.end_state) {
        double *e = E.end_state + (size_t)n * 6;
        if (L.steps == 0) { e[0] = e[1] = e[2] = e[3] = e[4] = e[5] = 0.0; }
        else if (MODE == 0 && L.rotated) { e[0] = L.lsx; e[1] = L.lsy; e[2] = L.lsz; e[3] = L.vx; e[4] = L.vy; e[5] = L.vz; }
        else { e[0] = L.sx; e[1] = L.sy; e[2] = L.sz; e[3] = L.vx; e[4] = L.vy; e[5] = L.vz; }
    }
    cnt.steps += (unsigned long long)L.steps;
    cnt.coll += (unsigned long long)L.num_col;
    switch (L.fate) {
    case 1: cnt.esc++; break;
    case 2: cnt.frozen++; break;
    case 3: cnt.photo++; break;
    case 4: cnt.depleted++; break;
    case 5: cnt.capped++; break;
    default: break;
    }
    cnt.done++;
    L.phase = LANE_DONE;
}

/* Moves counts gathered in `cnt` to the run's device counters and clears them. */
template <class Ops>
DSMC_HD void counts_flush(DsmcCounts &cnt, const DsmcEnv &E)
{
    if (!E.counters) return;
    unsigned long long v[8] = {cnt.esc, cnt.frozen, cnt.photo, cnt.depleted, cnt.capped, cnt.coll, cnt.steps,
                               cnt.done};
    for (int i = 0; i < 8; i++)
        if (v[i]) Ops::add_u64(&E.counters[i], v[i]);
    cnt.esc = cnt.frozen = cnt.photo = cnt.depleted = cnt.capped = cnt.coll = cnt.steps = cnt.done = 0;
}

/* ------------------------------------------------------------------ records */
/* The open deposit visit of a physical particle, kept in its record. */
DSMC_HD void rec_load_visit(DsmcLane &L, const DsmcRec &R)
{
    L.acc_cell = R.rotated;
    L.acc_w = R.thi; L.acc_n = (int)R.fi;
    L.acx = R.lsx; L.acy = R.lsy; L.acz = R.lsz;
}
DSMC_HD void rec_store_visit(const DsmcLane &L, DsmcRec &R)
{
    R.rotated = L.acc_cell;
    R.thi = L.acc_w; R.fi = (double)L.acc_n;
    R.lsx = L.acx; R.lsy = L.acy; R.lsz = L.acz;
}

/* Only what a flight step reads or writes travels with the flight kernel. */
template <int MODE, int ADV>
DSMC_HD void rec_load_flight(DsmcLane &L, const DsmcRec &R)
{
    L.sx = R.sx; L.sy = R.sy; L.sz = R.sz;
    L.vx = R.vx; L.vy = R.vy; L.vz = R.vz;
    L.ox = R.ox; L.oy = R.oy; L.oz = R.oz;
    L.Ri = R.Ri; L.W = R.W; L.steps = R.steps;
    L.td = R.td; L.fd = R.fd; L.ht = R.ht;
    L.phase = R.phase; L.fate = R.fate;
    if (MODE == 0) { L.time = R.time; L.flight_time = R.flight_time; L.ii = R.ii; }
    else { L.n_day = R.n_day; L.n_photo = R.n_photo; }
    if (ADV) {
        L.real_col_p = R.real_col_p; L.col_p = R.col_p; L.col_lo = R.col_lo;
        L.num_col = R.num_col; L.k = R.k; L.pid = R.pid; L.blk = ~0ull;
    }
    L.acc_cell = -1; L.rec_key = -1; L.rec_full = 0; L.ncell = -1;
    if (MODE != 0) rec_load_visit(L, R);
}

template <int MODE, int ADV>
DSMC_HD void rec_store_flight(const DsmcLane &L, DsmcRec &R)
{
    R.sx = L.sx; R.sy = L.sy; R.sz = L.sz;
    R.vx = L.vx; R.vy = L.vy; R.vz = L.vz;
    R.ox = L.ox; R.oy = L.oy; R.oz = L.oz;
    R.Ri = L.Ri; R.W = L.W; R.steps = L.steps;
    R.td = L.td; R.fd = L.fd; R.ht = L.ht;
    R.phase = L.phase; R.fate = L.fate;
    if (MODE == 0) { R.time = L.time; R.flight_time = L.flight_time; R.ii = (int32_t)L.ii; }
    else { R.n_day = L.n_day; rec_store_visit(L, R); }
    if (ADV) { R.real_col_p = L.real_col_p; R.col_p = L.col_p; R.num_col = L.num_col; R.k = L.k; }
}

/* The whole particle (start and service kernels).  The record's ii slot holds
 * the hop step counter in compat mode and the launch face in physical mode. */
template <int MODE>
DSMC_HD void rec_load_all(DsmcLane &L, const DsmcRec &R)
{
    L.sx = R.sx; L.sy = R.sy; L.sz = R.sz; L.vx = R.vx; L.vy = R.vy; L.vz = R.vz;
    L.ox = R.ox; L.oy = R.oy; L.oz = R.oz;
    L.Ri = R.Ri; L.W = R.W; L.time = R.time; L.flight_time = R.flight_time; L.thi = R.thi; L.fi = R.fi;
    L.rest_time = R.rest_time; L.real_col_p = R.real_col_p; L.col_p = R.col_p; L.col_lo = R.col_lo;
    L.lsx = R.lsx; L.lsy = R.lsy; L.lsz = R.lsz;
    L.steps = R.steps; L.n_day = R.n_day; L.n_photo = R.n_photo; L.pid = R.pid; L.k = R.k;
    if (MODE == 0) { L.ii = R.ii; L.face = -1; }
    else { L.ii = 0; L.face = R.ii; }
    L.phase = R.phase; L.fate = R.fate; L.hops = R.hops; L.num_col = R.num_col;
    L.td = R.td; L.fd = R.fd; L.ht = R.ht; L.rotated = R.rotated;
    L.blk = ~0ull; L.acc_cell = -1; L.rec_key = -1; L.rec_full = 0; L.ncell = -1;
    if (MODE != 0) { L.rotated = 0; rec_load_visit(L, R); }
}

template <int MODE>
DSMC_HD void rec_store_all(const DsmcLane &L, DsmcRec &R, int32_t run)
{
    R.sx = L.sx; R.sy = L.sy; R.sz = L.sz; R.vx = L.vx; R.vy = L.vy; R.vz = L.vz;
    R.ox = L.ox; R.oy = L.oy; R.oz = L.oz;
    R.Ri = L.Ri; R.W = L.W; R.time = L.time; R.flight_time = L.flight_time; R.thi = L.thi; R.fi = L.fi;
    R.rest_time = L.rest_time; R.real_col_p = L.real_col_p; R.col_p = L.col_p; R.col_lo = L.col_lo;
    R.lsx = L.lsx; R.lsy = L.lsy; R.lsz = L.lsz;
    R.steps = L.steps; R.n_day = L.n_day; R.n_photo = L.n_photo; R.pid = L.pid; R.k = L.k;
    R.ii = MODE == 0 ? (int32_t)L.ii : L.face;
    R.phase = L.phase; R.fate = L.fate; R.hops = L.hops; R.num_col = L.num_col;
    R.td = L.td; R.fd = L.fd; R.ht = L.ht; R.rotated = L.rotated; R.run = run;
    if (MODE != 0) rec_store_visit(L, R);
}

/* What a physical particle carries between launches of the flight kernel with
 * the landing service inline (records written by rec_store_all load the same). */
DSMC_HD void rec_load_phys(DsmcLane &L, const DsmcRec &R)
{
    L.sx = R.sx; L.sy = R.sy; L.sz = R.sz; L.vx = R.vx; L.vy = R.vy; L.vz = R.vz;
    L.ox = R.ox; L.oy = R.oy; L.oz = R.oz;
    L.Ri = R.Ri; L.W = R.W;
    L.steps = R.steps; L.n_day = R.n_day; L.n_photo = R.n_photo; L.pid = R.pid; L.k = R.k;
    L.phase = R.phase; L.fate = R.fate; L.hops = R.hops;
    L.td = R.td; L.fd = R.fd; L.ht = R.ht;
    L.face = R.ii;                   /* physical records keep the launch face in ii */
    L.num_col = R.num_col; L.rotated = 0;
    L.real_col_p = R.real_col_p; L.col_p = R.col_p;
    L.blk = ~0ull; L.rec_key = -1; L.rec_full = 0; L.ncell = -1;
    rec_load_visit(L, R);
}

DSMC_HD void rec_store_phys(const DsmcLane &L, DsmcRec &R)
{
    R.sx = L.sx; R.sy = L.sy; R.sz = L.sz; R.vx = L.vx; R.vy = L.vy; R.vz = L.vz;
    R.ox = L.ox; R.oy = L.oy; R.oz = L.oz;
    R.Ri = L.Ri; R.W = L.W;
    R.steps = L.steps; R.n_day = L.n_day; R.k = L.k;
    R.phase = L.phase; R.fate = L.fate; R.hops = L.hops;
    R.td = L.td; R.fd = L.fd; R.ht = L.ht;
    R.ii = L.face;
    R.num_col = L.num_col; R.real_col_p = L.real_col_p; R.col_p = L.col_p;
    rec_store_visit(L, R);
}

/* ------------------------------------------------------------------ hop start */
/* Starts a hop from the surface.  The cell indices are walked from the cell the
 * lane was last in (`near`: landing) or from a fixed guess (launch); the walks
 * and the sector test are exact, so both give the same cell. */
DSMC_HD void lane_begin_flight(DsmcLane &L, const DsmcEnv &E, bool near = false)
{
    double Ri = sqrt(L.sx * L.sx + L.sy * L.sy + L.sz * L.sz);
    L.ox = E.GM * L.sx / (Ri * Ri * Ri);
    L.oy = E.GM * L.sy / (Ri * Ri * Ri);
    L.oz = E.GM * L.sz / (Ri * Ri * Ri);
    L.Ri = Ri;
    L.ii = 0;
    /* current cell guesses for the incremental walks */
    double q = L.sz / Ri;
    if (near) {
        L.td = table_walk(E.tq, L.td, E.ftd - 1, q);
        L.ht = table_walk(E.hr, L.ht, E.hd, Ri);
        L.fd = fd_update(L.sx, L.sy, L.fd, E.ffd, E.cs);
    } else {
        L.td = table_walk(E.tq, E.ftd / 2, E.ftd - 1, q);
        L.ht = table_walk(E.hr, 0, E.hd, Ri);
        L.fd = fd_exact(L.sx, L.sy, E.ffd);
    }
    L.phase = LANE_FLIGHT;
}

/* ------------------------------------------------------------------ velocity table use */
/* velocity.c:49-68: the table row of a local temperature (rows 21-22 are never
 * chosen: every T >= 220 K takes row 20). */
DSMC_HD int vt_row(double T)
{
    int idx = T >= 220.0 ? 20 : (int)ceil(T / 10.0) - 2;
    if (idx < 0) idx = 0;
    if (idx >= DSMC_VT_NT) idx = DSMC_VT_NT - 1;
    return idx;
}

/* velocity.c:70-73: outgoing speed with sticking coefficient alpha. */
DSMC_HD double sticking_speed(double vr, double v_old_sq, double alpha)
{
    return sqrt(alpha * vr * vr + (1.0 - alpha) * v_old_sq);
}

/* collision.c:15-41 as the reference evaluates it: v <- (v + v2) + 0.5 (v - v2). */
DSMC_HD void collide_compat(double &vx, double &vy, double &vz, double v2x, double v2y, double v2z)
{
    const double dvx = vx - v2x, dvy = vy - v2y, dvz = vz - v2z;
    const double cvx = vx + v2x, cvy = vy + v2y, cvz = vz + v2z;
    vx = cvx + 0.5 * dvx;
    vy = cvy + 0.5 * dvy;
    vz = cvz + 0.5 * dvz;
}

/* ------------------------------------------------------------------ compat launch */
/* Body of `while (W >= 0.01)` up to the flight loop: ballistic.c:124-186 and
 * advanced.c:271-338. */
template <class Ops, int ADV>
DSMC_HD void lane_launch_compat(DsmcLane &L, const DsmcEnv &E, DsmcCounts &cnt)
{
    const int ftd = E.ftd, ffd = E.ffd;
    if (!(L.W >= E.w_min)) { L.fate = 4; lane_finish(L, E, cnt); return; }
    double KF = 2.0 * M_PI * E.R * fabs(dm_cos(L.thi)) / E.period;
    int td, fd;
    sph_index(L.thi, L.fi, ftd, ffd, &td, &fd);
    double local_T = E.T[td * ffd + fd];
    if (local_T < E.t_freeze) {
        int found = 0;
        if (!ADV) {                                  /* ballistic.c:132-153 */
            for (int add = 0; add < ffd; add++) {
                L.time += E.dt_rot;
                int cfd = (fd + add + 1) % ffd;
                if (E.T[td * ffd + cfd] >= E.t_freeze) {
                    fd = cfd;
                    L.fi = (fd + 0.5) / ffd * 2.0 * M_PI;
                    found = 1;
                    break;
                }
            }
        } else {                                     /* advanced.c:278-309 */
            Ops::add_f64(&E.horizon[td * ffd + fd], L.W);
            for (int add = 1; add <= ffd; add++) {
                L.time += E.dt_rot;
                L.rest_time += E.dt_rot;
                if (E.T[td * ffd + (add % ffd)] >= E.t_freeze) {
                    double rot_fi = (double)(add - fd) / ffd * 2.0 * M_PI;
                    double sn, cs;
                    dm_sincos(rot_fi, &sn, &cs);
                    if (!L.rotated) { L.lsx = L.sx; L.lsy = L.sy; L.lsz = L.sz; L.rotated = 1; }
                    double nx = L.sx * cs - L.sy * sn;
                    double ny = L.sx * sn + L.sy * cs;
                    L.sx = nx; L.sy = ny;
                    fd = add % ffd;
                    L.fi = (fd + 0.5) / ffd * 2.0 * M_PI;
                    found = 1;
                    break;
                }
            }
        }
        if (!found) {
            Ops::add_f64(&E.horizon[td * ffd + fd], L.W);
            L.fate = 2;
            lane_finish(L, E, cnt);
            return;
        }
        local_T = E.T[td * ffd + fd];
    }
    const int idx = vt_row(local_T);                                      /* velocity.c:49-68 */
    int r = lane_draw(L, E) % E.vt_count[idx];
    double vr = E.vt_speeds[idx * DSMC_VT_NV + r];
    double v_old_sq = L.vx * L.vx + L.vy * L.vy + L.vz * L.vz;
    double vout = sticking_speed(vr, v_old_sq, E.alpha);
    double flyd = lane_urand(L, E) * M_PI + L.thi - M_PI / 2.0;
    double xi = lane_urand(L, E) * M_PI + L.fi - M_PI / 2.0;
    double sfl, cfl, sxi, cxi, sr, cr;
    dm_sincos(flyd, &sfl, &cfl);
    dm_sincos(xi, &sxi, &cxi);
    dm_sincos(L.fi + M_PI / 2.0, &sr, &cr);
    double nvx = cfl * sxi * vout;
    double nvy = cfl * cxi * vout;
    double nvz = sfl * vout;
    nvx += KF * sr;
    nvy += KF * cr;
    L.hops++;
    if (sqrt(nvx * nvx + nvy * nvy + nvz * nvz) >= E.ve) {   /* ballistic.c:173 */
        L.fate = 1;
        lane_finish(L, E, cnt);
        return;
    }
    L.vx = nvx; L.vy = nvy; L.vz = nvz;
    L.rotated = 0;
    lane_begin_flight(L, E);
}

/* ------------------------------------------------------------------ physical emission */
struct EmitDraws {
    double u1, u2, u3, u4;
};

/* The four uniforms of an emission, words k .. k+7 of the particle's stream (the
 * words lane_u01 would return four times in a row), from one out-of-line call
 * that computes their two or three blocks together.  Drawn word by word, every
 * draw was a call site of its own whose cached block the caller kept (spilled,
 * in the register-heavy flight kernels) between draws. */
#if defined(__HIP_DEVICE_COMPILE__)
__attribute__((noinline))
#endif
DSMC_HD EmitDraws emit_draws(uint64_t seed, uint64_t pid, uint32_t stream, uint64_t k)
{
    const uint64_t b0 = k >> 2;
    const uint32_t off = (uint32_t)(k & 3);
    const dsmc_u32x4 a = dsmc_rng_block(seed, pid, stream, b0);
    const dsmc_u32x4 b = dsmc_rng_block(seed, pid, stream, b0 + 1);
    const dsmc_u32x4 c = off ? dsmc_rng_block(seed, pid, stream, b0 + 2) : b;
    const uint32_t s[12] = {a.w[0], a.w[1], a.w[2], a.w[3], b.w[0], b.w[1], b.w[2], b.w[3],
                            c.w[0], c.w[1], c.w[2], c.w[3]};
    uint32_t w[8];
#pragma unroll
    for (int i = 0; i < 8; i++) w[i] = off == 0 ? s[i] : off == 1 ? s[i + 1] : off == 2 ? s[i + 2] : s[i + 3];
    EmitDraws d;
    d.u1 = dsmc_u01_from_words(w[0], w[1]);
    d.u2 = dsmc_u01_from_words(w[2], w[3]);
    d.u3 = dsmc_u01_from_words(w[4], w[5]);
    d.u4 = dsmc_u01_from_words(w[6], w[7]);
    return d;
}

/* sig: sqrt(k_B T / m) at the launch face (DsmcEnv::face_sig). */
DSMC_HD void lane_emit(DsmcLane &L, const DsmcEnv &E, double sig)
{
    const EmitDraws D = emit_draws(E.seed, (uint64_t)L.pid, E.stream, L.k);
    L.k += 8;
    L.blk = ~0ull;                     /* the cached block is no longer the current one */
    double u1 = D.u1;
    double u2 = D.u2;
    double w = -dm_log(u1 * u2);
    double speed = sig * sqrt(2.0 * w);
    double u3 = D.u3;
    double cos_t = sqrt(u3), sin_t = sqrt(1.0 - u3);
    double phi = 2.0 * M_PI * D.u4;
    double sph, cph;
    dm_sincos(phi, &sph, &cph);
    double Ri = sqrt(L.sx * L.sx + L.sy * L.sy + L.sz * L.sz);
    double nx = L.sx / Ri, ny = L.sy / Ri, nz = L.sz / Ri;
    double hxy = sqrt(nx * nx + ny * ny);
    double e1x, e1y, e1z;
    if (hxy > 1e-12) { e1x = -ny / hxy; e1y = nx / hxy; e1z = 0.0; }
    else { e1x = 1.0; e1y = 0.0; e1z = 0.0; }
    double e2x = ny * e1z - nz * e1y;
    double e2y = nz * e1x - nx * e1z;
    double e2z = nx * e1y - ny * e1x;
    double a = sin_t * cph, b = sin_t * sph;
    L.vx = speed * (a * e1x + b * e2x + cos_t * nx) + E.omega * L.sy;
    L.vy = speed * (a * e1y + b * e2y + cos_t * ny) - E.omega * L.sx;
    L.vz = speed * (a * e1z + b * e2z + cos_t * nz);
}

DSMC_HD double face_margin(const DsmcEnv &E, int f, double dx, double dy, double dz)
{
    const double *n = E.enorm + (size_t)f * 9;
    double m0 = dx * n[0] + dy * n[1] + dz * n[2];
    double m1 = dx * n[3] + dy * n[4] + dz * n[5];
    double m2 = dx * n[6] + dy * n[7] + dz * n[8];
    double m = m0 < m1 ? m0 : m1;
    return m < m2 ? m : m2;
}

/* Finest icosphere face containing direction d (largest minimum edge margin at
 * every level of the subdivision hierarchy, lowest index on ties).  The margins
 * of a group of faces are computed before any of them is compared, so the
 * group's normals load together (one memory round trip per group: 5 for the 20
 * top faces, 1 per level below); the comparisons run in index order as a
 * one-at-a-time scan would make them. */
DSMC_HD int locate_face(const DsmcEnv &E, double dx, double dy, double dz)
{
    int best = 0;
    double bm = face_margin(E, 0, dx, dy, dz);
#pragma unroll 1
    for (int f0 = 1; f0 < 20; f0 += 4) {
        double m[4];
#pragma unroll
        for (int j = 0; j < 4; j++) m[j] = f0 + j < 20 ? face_margin(E, f0 + j, dx, dy, dz) : 0.0;
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (f0 + j < 20 && m[j] > bm) { bm = m[j]; best = f0 + j; }
    }
    /* level l starts at face 20 (4^l - 1) / 3 of the all-levels arrays */
    int off = 0, nl = 20;
#pragma unroll 1
    for (int l = 1; l <= E.levels; l++) {
        off += nl;
        nl *= 4;
        const int base = 4 * best;
        const double m0 = face_margin(E, off + base, dx, dy, dz), m1 = face_margin(E, off + base + 1, dx, dy, dz);
        const double m2 = face_margin(E, off + base + 2, dx, dy, dz), m3 = face_margin(E, off + base + 3, dx, dy, dz);
        int pick = base;
        double pm = m0;
        if (m1 > pm) { pm = m1; pick = base + 1; }
        if (m2 > pm) { pm = m2; pick = base + 2; }
        if (m3 > pm) { pm = m3; pick = base + 3; }
        best = pick;
    }
    return best;
}

/* Direction d at least eps inside the face of unit edge normals u[0..8]. */
DSMC_HD bool normals_hold(const double *u, double dx, double dy, double dz, double eps)
{
    const bool a = dx * u[0] + dy * u[1] + dz * u[2] >= eps;
    const bool b = dx * u[3] + dy * u[4] + dz * u[5] >= eps;
    const bool c = dx * u[6] + dy * u[7] + dz * u[8] >= eps;
    return a & b & c;
}

/* Direction d at least eps inside finest face f (all three unit edge normals). */
DSMC_HD bool face_holds(const DsmcEnv &E, int f, double dx, double dy, double dz, double eps)
{
    const double *u = E.fnorm + (size_t)f * 9;
    const bool a = dx * u[0] + dy * u[1] + dz * u[2] >= eps;
    const bool b = dx * u[3] + dy * u[4] + dz * u[5] >= eps;
    const bool c = dx * u[6] + dy * u[7] + dz * u[8] >= eps;
    return a & b & c;
}

/* Cube-map cell of direction d: the cube face its largest component points
 * through (x+, x-, y+, y-, z+, z-; ties to the earlier axis), then a G x G grid
 * over that face in gnomonic coordinates (the other two components over the
 * largest: x-faces (y, z), y-faces (z, x), z-faces (x, y)), index
 * (face * G + row) * G + column.  Float arithmetic: the cell only chooses which
 * faces are tested, each exactly.  -1 for a zero or non-finite direction. */
DSMC_HD int fmap_cell(int G, double dx, double dy, double dz)
{
    const float x = (float)dx, y = (float)dy, z = (float)dz;
    const float ax = fabsf(x), ay = fabsf(y), az = fabsf(z);
    int f;
    float u, v, m;
    if (ax >= ay && ax >= az) { f = x < 0.0f; m = ax; u = y; v = z; }
    else if (ay >= az) { f = 2 + (y < 0.0f); m = ay; u = z; v = x; }
    else { f = 4 + (z < 0.0f); m = az; u = x; v = y; }
    if (!(m > 0.0f) || !(m <= 3.0e38f)) return -1;
    const float half = 0.5f * (float)G, r = half / m;
    const float a = fminf(fmaxf(u * r + half, 0.0f), (float)G - 1.0f);
    const float b = fminf(fmaxf(v * r + half, 0.0f), (float)G - 1.0f);
    return (f * G + (int)b) * G + (int)a;
}

/* locate_face with a guess: a direction at least 1e-12 |d| inside a finest face
 * (unit edge normals) is inside every ancestor of it by at least as much, and
 * outside every other face of each level, so the hierarchical search would
 * return that face.  Tested in turn: `hint`, then the candidates the face map
 * holds for d's cube-map cell (four at a time, their normals loaded together);
 * when none holds d that firmly, or |d| is zero or not finite, the search runs. */
struct alignas(16) FmapQuad { int32_t f[4]; };

/* FMN: the caller's environment may hold fmap_n (pools); without it the code
 * for the stored normals is not compiled in (the single-run kernels, whose
 * register budget it would tighten). */
template <bool FMN = false>
DSMC_HD int locate_face_near(const DsmcEnv &E, int hint, double dx, double dy, double dz, double dlen)
{
    if (!E.fnorm || !(dlen > 0.0 && dlen <= 1.0e308)) return locate_face(E, dx, dy, dz);
    const double eps = 1e-12 * dlen;
    if (hint >= 0 && hint < E.nfine && face_holds(E, hint, dx, dy, dz, eps)) return hint;
    const int cell = E.fmap ? fmap_cell(E.fmap_g, dx, dy, dz) : -1;
    if (cell >= 0) {
        const FmapQuad *q = (const FmapQuad *)(E.fmap + (size_t)cell * DSMC_FMAP_K);
        int g0 = 0;
        if (FMN && E.fmap_n) {
            /* the first four candidates' normals stored with the cell: loaded
             * together with the candidates, one memory round trip */
            const FmapQuad c = q[0];
            const double *nn = E.fmap_n + (size_t)cell * 36;
            bool in[4];
#pragma unroll
            for (int j = 0; j < 4; j++) in[j] = normals_hold(nn + 9 * j, dx, dy, dz, eps);
            if (c.f[0] < 0) return locate_face(E, dx, dy, dz);
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (in[j] && c.f[j] >= 0) return c.f[j];
            g0 = 1;
        }
#pragma unroll 1
        for (int g = g0; g < DSMC_FMAP_K / 4; g++) {
            const FmapQuad c = q[g];
            if (c.f[0] < 0) break;
            bool in[4];
#pragma unroll
            for (int j = 0; j < 4; j++) in[j] = face_holds(E, c.f[j] < 0 ? c.f[0] : c.f[j], dx, dy, dz, eps);
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (in[j] && c.f[j] >= 0) return c.f[j];
        }
    }
    return locate_face(E, dx, dy, dz);
}

/* A fit polynomial held in registers for the root searches below (degree at
 * most 6, checked where runs are staged): the coefficients, highest power
 * first, right-aligned in DSMC_FIT_MAX_DEG + 1 slots behind leading zeros.
 * Horner's rule over the zeros keeps v = +0 until the first real coefficient
 * (0 * x + 0 = +0 for every finite x; a NaN or infinite x gives NaN either way),
 * so fit_eval returns exactly what Horner's rule over the deg + 1 coefficients
 * alone gives. */
#define DSMC_FIT_MAX_DEG 6
struct FitPoly {
    double c[DSMC_FIT_MAX_DEG + 1];
};

DSMC_HD FitPoly fit_load(const double *c, int deg)
{
    FitPoly f;
#pragma unroll
    for (int i = 0; i <= DSMC_FIT_MAX_DEG; i++) {
        const int j = i - (DSMC_FIT_MAX_DEG - deg);
        f.c[i] = j >= 0 ? c[j] : 0.0;
    }
    return f;
}

DSMC_HD double fit_eval(const FitPoly &f, double x)
{
    double v = 0.0;
#pragma unroll
    for (int i = 0; i <= DSMC_FIT_MAX_DEG; i++) v = v * x + f.c[i];
    return v;
}

/* Physical partner speed: the fit P(x) of the reverse cumulative speed
 * histogram in normalised speed x = (v - min_v) / (max_v - min_v), P(0) = 50;
 * the speed with P(x) = u (50 - P(1)) + P(1), by bisection (the sign at the
 * left end tracked, an exact zero ends it) to a 0.01 m/s bracket. */
DSMC_HD double sample_speed_phys(const double *cg, int deg, double min_v, double max_v, double u)
{
    const FitPoly c = fit_load(cg, deg);
    const double pv = fit_eval(c, 1.0);
    const double span = max_v - min_v;
    const double ranking = u * ((double)DSMC_VIRTUAL_RECORDS - pv) + pv;
    double x0 = 0.0, x1 = 1.0;
    double f0 = fit_eval(c, 0.0) - ranking;
    for (int iter = 0; iter < 100; iter++) {
        double xm = (x0 + x1) * 0.5;
        double fm = fit_eval(c, xm) - ranking;
        if (fm == 0.0) { x0 = x1 = xm; break; }
        if ((fm < 0.0) == (f0 < 0.0)) { x0 = xm; f0 = fm; }
        else x1 = xm;
        if ((x1 - x0) * span < 0.01) break;
    }
    return min_v + (x0 + x1) * 0.5 * span;
}

/* Elastic hard-sphere collision with a partner of the cell's speed distribution
 * moving in an isotropic direction; equal masses, isotropic scattering in the
 * centre-of-mass frame.  Then a new collision depth threshold.  INL: its twelve
 * random words from inline Philox rounds (for the out-of-line collision path). */
template <bool INL = false>
DSMC_HD void lane_collide(DsmcLane &L, const DsmcEnv &E)
{
    if (L.ht >= E.hd) { L.real_col_p = 0.0; L.phase = LANE_FLIGHT; return; }   /* cannot happen: depth grows in-grid */
    const int c = (L.ht * E.ftd + L.td) * E.ffd + L.fd;
    const int deg = E.fit_deg;
    double u = lane_u01<INL>(L, E);
    double vv2 = sample_speed_phys(E.coeffs + (size_t)c * (deg + 1), deg, E.pmin[c], E.pmax[c], u);
    double ct = 2.0 * lane_u01<INL>(L, E) - 1.0, st = sqrt(1.0 - ct * ct);
    double sp, cp;
    dm_sincos(2.0 * M_PI * lane_u01<INL>(L, E), &sp, &cp);
    double v2x = vv2 * st * cp, v2y = vv2 * st * sp, v2z = vv2 * ct;
    double gx = L.vx - v2x, gy = L.vy - v2y, gz = L.vz - v2z;
    double g = sqrt(gx * gx + gy * gy + gz * gz);
    double ct2 = 2.0 * lane_u01<INL>(L, E) - 1.0, st2 = sqrt(1.0 - ct2 * ct2);
    double sp2, cp2;
    dm_sincos(2.0 * M_PI * lane_u01<INL>(L, E), &sp2, &cp2);
    L.vx = 0.5 * (L.vx + v2x) + 0.5 * g * st2 * cp2;
    L.vy = 0.5 * (L.vy + v2y) + 0.5 * g * st2 * sp2;
    L.vz = 0.5 * (L.vz + v2z) + 0.5 * g * ct2;
    L.num_col++;
    L.real_col_p = 0.0;
    L.col_p = -dm_log(lane_u01<INL>(L, E));
    L.phase = LANE_FLIGHT;
}

/* ------------------------------------------------------------------ particle start */
template <class Ops, int MODE, int ADV>
DSMC_HD void lane_start(DsmcLane &L, const DsmcEnv &E, int64_t pid, DsmcCounts &cnt)
{
    L.pid = pid;
    L.k = 0; L.blk = ~0ull;
    L.W = 1.0; L.time = 0.0; L.flight_time = 0.0; L.rest_time = 0.0;
    L.steps = 0; L.hops = 0; L.num_col = 0; L.n_day = 0;
    L.fate = 0; L.rotated = 0;
    L.acc_cell = -1; L.rec_key = -1; L.rec_full = 0;
    L.vx = 0.0; L.vy = 0.0; L.vz = 0.0;
    L.real_col_p = 0.0; L.col_p = 0.0; L.col_lo = 0.0;
    if (MODE == 0) {
        if (ADV) {
            L.col_p = lane_urand(L, E);               /* advanced.c:246 */
            /* 1 - exp(-x) > col_p needs x > -log(1 - col_p); below this bound
             * (with a wide relative margin) the exact test cannot pass */
            double omc = 1.0 - L.col_p + 1e-15;
            L.col_lo = omc >= 1.0 ? 0.0 : -dm_log(omc) * (1.0 - 1e-9);
        }
        L.thi = lane_urand(L, E) * M_PI - M_PI / 2.0;
        L.fi = lane_urand(L, E) * 2.0 * M_PI;
        double st, ct, sf, cf;
        dm_sincos(L.thi, &st, &ct);
        dm_sincos(L.fi, &sf, &cf);
        L.sx = E.R * ct * sf;
        L.sy = E.R * ct * cf;
        L.sz = E.R * st;
        if (E.ttime) {
            E.ttime[0 * E.nparticles + pid] = L.thi;
            if (E.ttime_rows >= 9) {
                E.ttime[5 * E.nparticles + pid] = L.fi;
                int td, fd;
                sph_index(L.thi, L.fi, E.ftd, E.ffd, &td, &fd);
                E.ttime[6 * E.nparticles + pid] = E.T[td * E.ffd + fd];
            }
        }
        lane_launch_compat<Ops, ADV>(L, E, cnt);
        return;
    }
    /* physical: face by inverse CDF of the launch weights */
    const int nf = E.nfine;
    double x = lane_u01(L, E) * E.cdf[nf - 1];
    int lo = 0, hi = nf - 1;
    while (lo < hi) { int mid = (lo + hi) / 2; if (E.cdf[mid] > x) hi = mid; else lo = mid + 1; }
    int face = lo;
    const int32_t *fv = E.faces + 3 * (size_t)(E.level_off[E.levels] + face);
    const double *a = E.vert + 3 * (size_t)fv[0];
    const double *b = E.vert + 3 * (size_t)fv[1];
    const double *e = E.vert + 3 * (size_t)fv[2];
    double u1 = lane_u01(L, E);
    double u2 = lane_u01(L, E);
    double r1 = sqrt(u1);
    double wa = 1.0 - r1, wb = r1 * (1.0 - u2), wc = r1 * u2;
    double px = wa * a[0] + wb * b[0] + wc * e[0];
    double py = wa * a[1] + wb * b[1] + wc * e[1];
    double pz = wa * a[2] + wb * b[2] + wc * e[2];
    double sc = E.R / sqrt(px * px + py * py + pz * pz);
    L.sx = px * sc; L.sy = py * sc; L.sz = pz * sc;
    double Ex = -dm_log(lane_u01(L, E));
    double nph = ceil(Ex * E.lifetime / E.h);
    L.n_photo = nph > 9.0e18 ? (int64_t)9000000000000000000LL : (int64_t)nph;
    L.face = face;
    lane_emit(L, E, E.face_sig[face]);
    if (ADV) {                       /* collision depth threshold ~ Exp(1) */
        L.real_col_p = 0.0;
        L.col_p = -dm_log(lane_u01(L, E));
    }
    L.ncell = -1;
    L.hops = 1;
    lane_begin_flight(L, E);
    lane_open_visit(L, E);
}

/* ------------------------------------------------------------------ service */
/* After a flight ends on the surface (or the per-hop step limit). */
template <class Ops, int MODE, int ADV, bool FMN = false>
DSMC_HD void lane_service(DsmcLane &L, const DsmcEnv &E, DsmcCounts &cnt)
{
    lane_flush<Ops>(L, E);
    if (L.phase == LANE_ENDED) { lane_finish<MODE>(L, E, cnt); return; }
    if (MODE == 0) {                         /* ballistic.c:239-256 / advanced.c:405-420 */
        double Ri = sqrt(L.sx * L.sx + L.sy * L.sy + L.sz * L.sz);
        pos_angles(L.sx, L.sy, L.sz, Ri, &L.thi, &L.fi);
        int td, fd;
        sph_index(L.thi, L.fi, E.ftd, E.ffd, &td, &fd);
        if (td <= 1 || td >= E.ftd - 2) {
            Ops::add_f64(&E.horizon[td * E.ffd + fd], 0.05 * L.W);
            L.W *= 0.95;
        }
        lane_launch_compat<Ops, ADV>(L, E, cnt);
        return;
    }
    if (ADV && L.phase == LANE_COLLIDE) { lane_collide(L, E); return; }
    int64_t cap = E.max_steps > 0 ? E.max_steps : INT64_MAX;
    if (L.steps >= cap) { L.fate = 5; lane_finish<MODE>(L, E, cnt); return; }
    const double Rl = sqrt(L.sx * L.sx + L.sy * L.sy + L.sz * L.sz);
    /* no guess: the hop's launch face holds the landing point in ~15% of the
     * landings of the 64-run batch and the face map's first four candidates in
     * ~84%, so testing the launch face first costs most landings a dependent
     * round trip to memory (the 64-run batch: 2% faster without it) */
    int f = locate_face_near<FMN>(E, -1, L.sx, L.sy, L.sz, Rl);
    double Tf = E.face_T[f];
    if (Tf < E.t_freeze) {
        /* the horizon cell of sph_index(pos_angles(...)) by the exact walks from
         * the lane's last cell (the threshold tables reproduce those functions'
         * indices; no arcsine or arctangent unless a point lies within 1e-12 of
         * a sector boundary) */
        const int td = td_update(L.sz, Rl, L.td, E.ftd, E.tq);
        const int fd = fd_update(L.sx, L.sy, L.fd, E.ffd, E.cs);
        Ops::add_f64(&E.horizon[td * E.ffd + fd], L.W);
        L.fate = 2;
        lane_finish<MODE>(L, E, cnt);
        return;
    }
    double sc = E.R / Rl;
    L.sx *= sc; L.sy *= sc; L.sz *= sc;
    L.face = f;
    lane_emit(L, E, E.face_sig[f]);
    L.hops++;
    lane_begin_flight(L, E, true);
    lane_open_visit(L, E);
}

/* ------------------------------------------------------------------ collision (advanced) */

/* collision.c:43-101 */
DSMC_HD double sample_speed(const double *cg, int deg, double min_v, double max_v, double u, int nn)
{
    const FitPoly c = fit_load(cg, deg);
    double val_at_max = fit_eval(c, max_v);
    double ranking = u * ((double)nn - val_at_max) + val_at_max;
    double v0 = min_v, v1 = max_v;
    for (int iter = 0; iter < 100; iter++) {
        double vm = (v0 + v1) * 0.5;
        double fm = fit_eval(c, vm) - ranking;
        double f1 = fit_eval(c, v1) - ranking;
        if (fm * f1 < 0) v0 = vm;
        else v1 = vm;
        if (fabs(v1 - v0) < 0.01) break;
    }
    return (v0 + v1) * 0.5;
}

/* integrator.c:4-55: one Stormer-Verlet step of position s and velocity v under
 * the attraction GM / r^2, with the previous step's acceleration o (replaced by
 * this step's); Ri = |s|, c1 = 1.5 h^2, c2 = h^2 / 6.  DS: *ds = |s_new - s|.
 * Returns |s_new|.  The reference's operation order, no contraction; the
 * division and square root are the IEEE ones (dm_div_pos, dm_sqrt_pos: GM,
 * r^3 and |s|^2 of a body's exosphere lie far inside their range). */
template <bool DS>
DSMC_HD double verlet_push(double &sx, double &sy, double &sz, double &vx, double &vy, double &vz, double &ox,
                           double &oy, double &oz, double Ri, double GM, double h, double c1, double c2, double *ds)
{
    const double g = dm_div_pos(GM, Ri * Ri * Ri);
    /* the previous acceleration's term first, so the new acceleration can take
     * its registers: the same products and sums, each rounded as before */
    const double tx = ox * c2, ty = oy * c2, tz = oz * c2;
    ox = g * sx; oy = g * sy; oz = g * sz;
    const double nx = sx + vx * h - ox * c1 - tx;
    const double ny = sy + vy * h - oy * c1 - ty;
    const double nz = sz + vz * h - oz * c1 - tz;
    if (DS) {
        const double dx = nx - sx, dy = ny - sy, dz = nz - sz;
        *ds = sqrt(dx * dx + dy * dy + dz * dz);
    }
    vx -= ox * h;
    vy -= oy * h;
    vz -= oz * h;
    sx = nx; sy = ny; sz = nz;
    return dm_sqrt_pos(nx * nx + ny * ny + nz * nz);
}

template <class Ops, int ADV>
DSMC_HD void lane_step_compat(DsmcLane &L, const DsmcEnv &E, const DsmcCellTabs &TB, double nx, double ny,
                              double nz, double Rn, double ds);

/* ------------------------------------------------------------------ one flight step */
/* A physical step after its push (the landing, escape, photolysis, cell, deposit
 * and cap logic of lane_step), with the current cell's table entries CP read
 * before the push.  Returns with L.phase = FLIGHT, LANDED, ENDED or COLLIDE. */
template <class Ops, int ADV, class EnvT>
DSMC_HD void lane_tail_phys(DsmcLane &L, const EnvT &E, const DsmcCellTabs &TB, const CellPre &CP, double Rn,
                            double ds)
{
    const double nx = L.sx, ny = L.sy, nz = L.sz;
    if (Rn <= E.R) { L.phase = LANE_LANDED; return; }
    if (Rn >= E.max_height) { lane_flush<Ops>(L, E); L.fate = 1; L.phase = LANE_ENDED; return; }
    if (nx * E.sun_x + ny * E.sun_y > 0.0) {
        L.n_day++;
        if (L.n_day >= L.n_photo) { lane_flush<Ops>(L, E); L.fate = 3; L.phase = LANE_ENDED; return; }
    }
    L.td = td_update_pre(nz, Rn, L.td, E.ftd, TB.tq, CP);
    L.fd = fd_update_pre(nx, ny, L.fd, E.ffd, TB.cs, CP);
    L.ht = table_walk_pre(TB.hr, L.ht, E.hd, Rn, CP);
    if (L.ht < E.hd) {
        const int cell = (L.ht * E.ftd + L.td) * E.ffd + L.fd;
        if constexpr (ADV != 0) {
            if (cell != L.ncell) { L.ncur = E.nsig[cell]; L.ncell = cell; }
        }
        lane_deposit<Ops, false>(L, E, cell);
    } else if constexpr (ADV != 0) {
        L.ncur = 0.0;
        L.ncell = -1;
    }
    if constexpr (ADV != 0) L.real_col_p += L.ncur * ds;     /* collision depth */
    int64_t cap = E.max_steps > 0 ? E.max_steps : INT64_MAX;
    if (L.steps >= cap) { lane_flush<Ops>(L, E); L.fate = 5; L.phase = LANE_ENDED; }
    else if (ADV != 0 && L.real_col_p >= L.col_p) L.phase = LANE_COLLIDE;
}

/* integrator.c:4-55 followed by ballistic.c:195-236 / advanced.c:347-402, or the
 * physical-mode step.  Returns with L.phase = FLIGHT, LANDED or DONE. */
template <class Ops, int MODE, int ADV, class EnvT>
DSMC_HD void lane_step(DsmcLane &L, const EnvT &E, const DsmcCellTabs &TB)
{
    const double h = E.h;
    CellPre CP;
    if constexpr (MODE != 0) CP = cell_pre(L.td, L.fd, L.ht, TB.tq, TB.cs, TB.hr);
    if (MODE == 0) { L.flight_time += h; L.time += h; }
    L.steps++;
    double ds = 0.0;
    const double Rn = verlet_push<ADV != 0>(L.sx, L.sy, L.sz, L.vx, L.vy, L.vz, L.ox, L.oy, L.oz, L.Ri, E.GM, h,
                                            E.c1, E.c2, &ds);
    L.Ri = Rn;
    if constexpr (MODE != 0) {
        lane_tail_phys<Ops, ADV>(L, E, TB, CP, Rn, ds);
    } else {
        const double nx = L.sx, ny = L.sy, nz = L.sz;
        if (Rn <= E.R) { L.phase = LANE_LANDED; return; }
        if (Rn >= E.max_height) { lane_flush<Ops>(L, E); L.fate = 1; L.phase = LANE_ENDED; return; }
        lane_step_compat<Ops, ADV>(L, E, TB, nx, ny, nz, Rn, ds);
    }
}

/* ------------------------------------------------------------------ lean physical step */
/* What a lane of the event-batched flight loops holds besides DsmcLane: its
 * cell's lean bounds (in registers, so a step reads no table), the step at which
 * the lane must stop (the step cap or the launch's budget, whichever comes
 * first), the sunlit steps taken since the last refresh and whether a step may
 * accumulate into the open visit without the cell logic: the cell is above the
 * grid (no deposit), or it is the open visit's cell.
 *
 * Lean bounds: each of the cell tests of a full step (td_update_pre,
 * fd_update_pre, table_walk_pre: inside by a margin of 1e-12 relative, landed,
 * escaped) reduced to one product and one comparison, by bounds moved inward:
 * latitude thresholds by LEAN_TQ_SHIFT (z > tl Rn implies z - tlo Rn > 1e-12 Rn
 * with every rounding counted: the shift is 4e-12 against a margin of 1e-12
 * and roundings of ~1e-16); sector boundary directions rotated inward by an
 * angle of sine LEAN_CS_SHIFT (x c0' - y s0' > 0 gives rho sin(psi - beta0) >
 * ~4e-12 rho > 1e-12 (|x| + |y|)); the radial range intersected with (R,
 * max_height), exactly (Rn > R is Rn >= the next double above R).  A step that
 * passes them passes the full step's tests: only the converse can fail, and a
 * lane that fails them parks and runs the full step.  The full step's first
 * tests (the _pre functions) take the same moved bounds: inside them a point is
 * inside the cell, and otherwise those functions fall back to the exact walks
 * on the tables, so the cell they return is the one the original bounds give
 * (lane_tail_phys tests landing and escape before the radial walk). */
#define LEAN_TQ_SHIFT 4e-12
#define LEAN_CS_SHIFT 4e-12

struct LeanLane {
    CellPre B;      /* the lean bounds, in CellPre's layout */
    int64_t stop;
    int32_t sun_steps;  /* sunlit lean steps since the refresh, not yet in L.n_day */
    int32_t sun_lim;    /* L.n_photo - L.n_day - 1 at the refresh, at most INT32_MAX */
    bool easy;      /* the next steps may take lane_fast_phys's shortcut */
    bool pend;      /* a pushed step waits for lean_tail */
};

/* What lane_fast_phys reads of a run's constants (a pool lane holds these in
 * registers and reads its run's full step constants only for the cell logic). */
struct LeanEnv {
    double GM, h, c1, c2, sun_x, sun_y;
    int hd;
};

template <class EnvT>
DSMC_HD LeanEnv lean_env(const EnvT &E)
{
    LeanEnv F;
    F.GM = E.GM; F.h = E.h; F.c1 = E.c1; F.c2 = E.c2; F.sun_x = E.sun_x; F.sun_y = E.sun_y;
    F.hd = E.hd;
    return F;
}

/* The lean bounds of cell (td, fd, ht): a function of the cell alone. */
template <class EnvT>
DSMC_HD CellPre lean_bounds(int td, int fd, int ht, const EnvT &E, const DsmcCellTabs &TB)
{
    const CellPre P = cell_pre(td, fd, ht, TB.tq, TB.cs, TB.hr);
    CellPre B;
    B.tlo = P.tlo + LEAN_TQ_SHIFT;
    B.thi = P.thi - LEAN_TQ_SHIFT;
    B.c0 = P.c0 - P.s0 * LEAN_CS_SHIFT;
    B.s0 = P.s0 + P.c0 * LEAN_CS_SHIFT;
    B.c1 = P.c1 + P.s1 * LEAN_CS_SHIFT;
    B.s1 = P.s1 - P.c1 * LEAN_CS_SHIFT;
    const double r_up = nextafter(E.R, INFINITY);          /* Rn > R  <=>  Rn >= r_up */
    B.rlo = P.rlo > r_up ? P.rlo : r_up;
    B.rhi = P.rhi < E.max_height ? P.rhi : E.max_height;
    return B;
}

/* The shortcut flag and the sunlit limit after the cell logic ran or a hop began
 * (the bounds: lean_bounds). */
template <class EnvT>
DSMC_HD void lean_refresh_flags(LeanLane &N, const DsmcLane &L, const EnvT &E)
{
    N.easy = L.ht >= E.hd || L.acc_cell == (L.ht * E.ftd + L.td) * E.ffd + L.fd;
    const int64_t lim = L.n_photo - L.n_day - 1;
    N.sun_lim = lim < (int64_t)INT32_MAX ? (int32_t)lim : INT32_MAX;
    N.sun_steps = 0;
}

/* Refreshes the lean bounds, the shortcut flag and the sunlit limit after the
 * cell logic ran. */
template <class EnvT>
DSMC_HD void lean_refresh(LeanLane &N, const DsmcLane &L, const EnvT &E, const DsmcCellTabs &TB)
{
    N.B = lean_bounds(L.td, L.fd, L.ht, E, TB);
    N.easy = L.ht >= E.hd || L.acc_cell == (L.ht * E.ftd + L.td) * E.ffd + L.fd;
    const int64_t lim = L.n_photo - L.n_day - 1;
    N.sun_lim = lim < (int64_t)INT32_MAX ? (int32_t)lim : INT32_MAX;
    N.sun_steps = 0;
}

/* Adds the lean steps' sunlit count to the lane (before the cell logic reads it). */
DSMC_HD void lean_sync(LeanLane &N, DsmcLane &L)
{
    L.n_day += N.sun_steps;
    N.sun_steps = 0;
}

/* One physical step (ballistic, density and moments by visit) whose outcome is,
 * in most steps, decided by a few comparisons: the push, then - when the point
 * passes the lean bounds (firmly inside the current cell, neither landed nor
 * escaped), is not photolysed on this step, is short of the stop step, and the
 * open visit is this cell's - the sunlit count (held in N.sun_steps until
 * lean_sync) and the visit's sums, exactly as lane_tail_phys would apply them,
 * and true.  Otherwise only the push is applied and it returns false: lean_tail
 * completes the step. */
template <class EnvT>
DSMC_HD bool lane_fast_phys(DsmcLane &L, const EnvT &E, LeanLane &N)
{
    L.steps++;
    const double Rn = verlet_push<false>(L.sx, L.sy, L.sz, L.vx, L.vy, L.vz, L.ox, L.oy, L.oz, L.Ri, E.GM, E.h,
                                         E.c1, E.c2, nullptr);
    L.Ri = Rn;
    const double x = L.sx, y = L.sy, z = L.sz;
    /* fl(a + b) > 0 exactly when a > -b, and fl(a - b) > 0 exactly when a > b
     * (a rounded sum or difference keeps the sign of the exact one): the same
     * decisions as lane_tail_phys's sun test and the sector tests' differences */
    const bool sun = x * E.sun_x > -(y * E.sun_y);
    const CellPre &B = N.B;
    const bool stay = (z > B.tlo * Rn) & (z < B.thi * Rn) & (x * B.c0 > y * B.s0) & (x * B.c1 < y * B.s1) &
                      (Rn >= B.rlo) & (Rn < B.rhi);
    /* photolysed on this step: the sunlit count L.n_day + N.sun_steps + 1 reaches
     * n_photo (with sun_lim clamped, a lane parks before sun_steps overflows) */
    const bool ok = stay & (L.steps < N.stop) & N.easy & !(sun & (N.sun_steps >= N.sun_lim));
    if (ok) {
        N.sun_steps += sun ? 1 : 0;
        if (L.ht < E.hd) {
            L.acx += L.vx; L.acy += L.vy; L.acz += L.vz;
            L.acc_n++;
        }
    }
    return ok;
}

/* The rest of a parked lane's step: the full step logic (lane_tail_phys),
 * its first cell tests on the lean bounds. */
template <class Ops, class EnvT>
DSMC_HD void lean_tail(DsmcLane &L, const EnvT &E, const DsmcCellTabs &TB, LeanLane &N)
{
    lean_sync(N, L);
    lane_tail_phys<Ops, 0>(L, E, TB, N.B, L.Ri, 0.0);
}

/* The rest of a compat step (ballistic.c:195-236 / advanced.c:347-402). */
template <class Ops, int ADV>
DSMC_HD void lane_step_compat(DsmcLane &L, const DsmcEnv &E, const DsmcCellTabs &TB, double nx, double ny,
                              double nz, double Rn, double ds)
{
    L.td = td_update(nz, Rn, L.td, E.ftd, TB.tq);
    L.fd = fd_update(nx, ny, L.fd, E.ffd, TB.cs);
    if (L.fd < E.ffd / 2) L.W *= E.decay;
    int64_t ii = L.ii++;
    if (ii + 1 >= E.max_hop_steps) L.phase = LANE_LANDED;   /* loop bound reached */
    if (ii == 0) return;
    L.ht = table_walk(TB.hr, L.ht, E.hd, Rn);
    if (L.ht >= E.hd) return;
    if (ADV) {
        if (E.copy_atm && L.ht < E.hd / 3) {           /* advanced.c:369-398 */
            int64_t atm_idx = (int64_t)L.ht * E.ftd * E.ffd + L.td * E.ffd + L.fd;
            double atm = atm_idx < E.atm_total ? E.copy_atm[atm_idx] : 0.0;
            L.real_col_p += ds * atm * E.sigma;
            if (L.real_col_p >= L.col_lo) {
                int pidx = L.td * E.ffd * E.hd + L.fd * E.hd + L.ht;
                if (L.col_p < 1.0 - dm_exp(-L.real_col_p) && E.pvalid[pidx]) {
                    L.real_col_p = 0.0;
                    L.num_col++;
                    double u = lane_urand(L, E);
                    double vv2 = E.pvalid[pidx]
                        ? sample_speed(E.coeffs + (size_t)pidx * (E.fit_deg + 1), E.fit_deg,
                                       E.pmin[pidx], E.pmax[pidx], u, E.count_num[pidx])
                        : 0.0;
                    double thi2 = M_PI * (double)lane_draw(L, E) / 2147483647 - M_PI / 2.0;
                    double fi2 = 2.0 * M_PI * (double)lane_draw(L, E) / 2147483647;
                    double st2, ct2, sf2, cf2;
                    dm_sincos(thi2, &st2, &ct2);
                    dm_sincos(fi2, &sf2, &cf2);
                    collide_compat(L.vx, L.vy, L.vz, vv2 * ct2 * cf2, vv2 * ct2 * sf2, vv2 * st2);
                    return;
                }
            }
        }
    } else if (E.rec_cap_cell > 0 && E.vrec) {
        lane_record<Ops>(L, E);
    }
    lane_deposit<Ops>(L, E, (L.ht * E.ftd + L.td) * E.ffd + L.fd);
}

#endif /* DSMC_LANE_H */
