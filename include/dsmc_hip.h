/*
 * dsmc_hip.h - C ABI of the MI355X-native test-particle DSMC path.
 *
 * Two groups of entry points:
 *
 *  1. Drop-in replacements for dsmc_suite's drivers and file I/O.  Same names,
 *     same signatures, same file side effects (`<prefix>_den.dat`, ...):
 *       run_ballistic  replaces src/dsmc/ballistic.c:35-36 (declared extern at
 *                      src/main.c:18-19, called at src/main.c:147,194)
 *       run_advanced   replaces src/dsmc/advanced.c:119-121 (declared extern at
 *                      src/main.c:20-23, called at src/main.c:150,202)
 *       read_ / write_*_dat  replace include/fileio.h:8-44 (src/common/fileio.c)
 *     `nthreads` keeps its meaning of host threads; the device is implicit
 *     (HIP device 0 unless DSMC_DEVICE is set in the environment).
 *
 *  2. The parameterised in-memory API (dsmc_run_*), the multi-run batch API and
 *     the serving API.  They return an int status (0 = ok, <0 = dsmc_status).
 *
 * All pointers are plain host pointers owned by the caller unless stated.
 * No torch or HIP types appear in any signature.
 */
#ifndef DSMC_HIP_H
#define DSMC_HIP_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
enum dsmc_status {
    DSMC_OK = 0,
    DSMC_ERR_ARG = -1,        /* invalid argument (sizes, NULL where required) */
    DSMC_ERR_IO = -2,         /* file missing, short or malformed */
    DSMC_ERR_DEVICE = -3,     /* no usable HIP device / kernel image */
    DSMC_ERR_NOMEM = -4,      /* device or host allocation failed */
    DSMC_ERR_BUDGET = -5,     /* request does not fit the serving HBM budget */
    DSMC_ERR_CANCELLED = -6,  /* serving request cancelled before completion */
    DSMC_ERR_BUSY = -7        /* serving queue full / result not ready */
};

/* ------------------------------------------------------------------ modes */
enum dsmc_mode {
    /* Reproduces dsmc_suite's run_ballistic / run_advanced state machines,
     * quirks included (uniform-angle launch, velocity table, launch-cell frozen
     * scan, polar deposit, fixed dayside fd < ffd/2, W-decay photolysis). */
    DSMC_MODE_COMPAT = 0,
    /* The exosphere model of the benchmark configurations: icosphere surface,
     * sublimation-flux-weighted launch, flux-Maxwellian (Lambertian) emission at
     * the local face temperature, freeze-out on impact below t_freeze,
     * Monte-Carlo photolysis on the sunlit half-space, per-particle step cap. */
    DSMC_MODE_PHYSICAL = 1
};

enum dsmc_fate {
    DSMC_FATE_NONE = 0,
    DSMC_FATE_ESCAPED = 1,     /* reached max_height, or launch speed >= v_esc (compat) */
    DSMC_FATE_FROZEN = 2,      /* frozen on a cold surface */
    DSMC_FATE_PHOTOLYSED = 3,  /* photodissociated (physical mode) */
    DSMC_FATE_DEPLETED = 4,    /* weight fell below w_min (compat: photolysis + polar loss) */
    DSMC_FATE_CAPPED = 5       /* per-particle step cap reached (physical mode) */
};

enum dsmc_rng_stream {
    DSMC_STREAM_BALLISTIC = 1,
    DSMC_STREAM_ADVANCED = 2
};

/* ------------------------------------------------------------------ params */
/* One run's scalars.  Derived quantities are computed from these exactly the way
 * the reference computes them (GM = G*M, v_esc = sqrt(2*G*M/R), molecule mass =
 * mmass/mole, dt_rot = period/ffd), so compat runs reproduce its arithmetic. */
typedef struct dsmc_params {
    int mode;                 /* enum dsmc_mode */
    int stream;               /* enum dsmc_rng_stream (Philox counter word 3) */
    uint64_t seed;            /* Philox key */
    int64_t nparticles;       /* test particles in this run */

    /* body and species (include/constants.h:5-35 for the Ceres/H2O preset) */
    double G, M, radius;
    double k_B, mmass, mole;
    double lifetime;          /* photolysis lifetime tau [s] */
    double period;            /* rotation period P [s] */
    double alpha;             /* sticking coefficient (compat velocity_sticking) */
    double v_esc;             /* escape speed used by the compat launch test */

    /* integration */
    double dt;                /* Verlet step h [s] */
    int64_t max_hop_steps;    /* per-hop step limit (reference: 1e7) */
    int64_t max_steps;        /* per-particle step cap (physical), 0 = none */

    /* thresholds */
    double t_freeze;          /* compat: launch-cell T < t_freeze freezes; physical: impact T */
    double w_min;             /* particle ends when W < w_min (reference 0.01) */
    double subsolar_lon;      /* physical: longitude of the subsolar point [rad] */

    /* grid */
    int ftd, ffd, hd;         /* latitude bins, longitude bins, radial shells */
    int exp_grid;             /* 0: uniform shells of unih, 1: exponential (alpha2) */
    double unih, alpha2, max_height;

    /* velocity records (compat ballistic v_record, reference ballistic.c:77-82) */
    int record_cap_cell;      /* per (td, ht) cap, reference 200; 0 disables */
    int64_t record_cap_total; /* global cap, reference 100*ftd*hd */

    /* collisions (advanced) */
    double cross_section;     /* sigma [m^2] */
    int fit_degree;           /* polynomial CDF degree (reference POLY_DEGREE = 3) */
    int max_per_cell;         /* speeds kept per cell for the fit (reference 50) */
    int min_fit_samples;      /* fit only cells with nn > this (reference 8) */
    double density_scale;     /* multiplies the phase-1 density read as copy_atm */

    /* physical-mode emission */
    double sublimation_scale; /* multiplies the emission rate (collisional runs: x10) */
    double production_rate;   /* collisional runs: emission rate [molecules/s] behind the
                                 phase-1 number density; 0 = sum of face weights x R^2 */

    /* physical collisional runs: phase-1 speed histograms per cell (one sample
     * per cell visit: the speed of the visit's mean velocity) */
    int vhist_bins;           /* bins per cell, <= 64 */
    double vhist_max;         /* upper edge of the last regular bin [m/s]; the last bin is open */
} dsmc_params;

/* Presets.  COMPAT_BALLISTIC / COMPAT_ADVANCED reproduce the reference CLI
 * (ballistic.c:49-67, advanced.c:131-152; ftd/ffd come from the T file).
 * CERES_PHYSICAL is the benchmark's Ceres/H2O ballistic configuration. */
enum dsmc_preset {
    DSMC_PRESET_COMPAT_BALLISTIC = 0,
    DSMC_PRESET_COMPAT_ADVANCED = 1,
    DSMC_PRESET_CERES_PHYSICAL = 2,
    DSMC_PRESET_CERES_PHYSICAL_COLL = 3
};
int dsmc_params_preset(dsmc_params *p, int preset, int ftd, int ffd);

/* ------------------------------------------------------------------ surface */
/* Compat runs read a lat/lon temperature field [ftd][ffd].  Physical runs use an
 * icosphere: faces of all subdivision levels are stored level by level (20, 80,
 * 320, ... faces); the children of face f of level l are faces 4f..4f+3 of level
 * l+1 (the subdivision order of scripts/vesta_raytrace_gif.py:35-102). */
typedef struct dsmc_surface {
    int ftd, ffd;
    const double *T_latlon;   /* [ftd*ffd], compat */

    int levels;               /* subdivision levels L; finest level has 20*4^L faces */
    int nvert;
    const double *vert;       /* [nvert*3], unit vectors */
    const int32_t *faces;     /* [sum_l 20*4^l][3] vertex indices, all levels */
    const double *face_T;     /* [20*4^L] temperature of each finest face [K] */
    const double *face_w;     /* [20*4^L] launch weight of each finest face */
} dsmc_surface;

/* Builds an icosphere with `levels` subdivisions into caller buffers sized by
 * dsmc_icosphere_sizes, with the synthetic temperature
 * T = t_night + t_amp * max(cos(sza), 0)^0.25 and Antoine/Hertz-Knudsen
 * sublimation weights (face area x flux).  Returns 0 or DSMC_ERR_ARG. */
int dsmc_icosphere_sizes(int levels, int *nvert, int *nface_all, int *nface_fine);
int dsmc_icosphere_build(int levels, double subsolar_lon, double t_night, double t_amp,
                         double molecule_mass, double k_B,
                         double *vert, int32_t *faces, double *face_T, double *face_w);
/* Rasterises the face temperatures to a lat/lon grid (cell centre -> face). */
int dsmc_icosphere_rasterise(const dsmc_surface *s, int ftd, int ffd, double *T_latlon);

/* ------------------------------------------------------------------ results */
/* Every pointer may be NULL (that output is then not produced). */
typedef struct dsmc_result {
    double *density;          /* [hd*ftd*ffd], [ht][td][fd], sum of W per step */
    double *horizon;          /* [ftd*ffd] frozen / polar deposits */
    double *vmom;             /* [hd*ftd*ffd*3] sum of velocity per deposit (physical) */
    int64_t *vcount;          /* [hd*ftd*ffd] deposits per cell (physical) */
    double *ttime;            /* [rows*n], rows 5 (ballistic) or 9 (advanced); compat runs only: a physical run zeroes 5 rows */
    double *end_state;        /* [n*6] final position and velocity */
    int64_t *steps;           /* [n] Verlet steps taken by each particle */
    int32_t *fate;            /* [n] enum dsmc_fate */
    int32_t *hops;            /* [n] launches per particle */
    int64_t *vhist;           /* [hd*ftd*ffd*vhist_bins] physical: steps per speed bin of cell visits */
    /* compat ballistic velocity records, capacity record_cap_total */
    double *vrec;             /* [cap*3] */
    int32_t *vrec_td, *vrec_fd, *vrec_ht;
    int64_t vrec_n;           /* out: records written */
    /* out: counters */
    int64_t n_escaped, n_frozen, n_photolysed, n_depleted, n_capped;
    int64_t n_collisions, total_steps;
    double elapsed_ms;        /* out: device time of the run */
} dsmc_result;

/* Phase-2 inputs (advanced): the phase-1 density read as copy_atm and the
 * velocity records.  Both optional, as in advanced.c:160-209. */
typedef struct dsmc_phase2_input {
    const double *copy_atm;   /* [atm_hd*atm_ftd*atm_ffd] or NULL (no collisions) */
    int atm_ftd, atm_ffd, atm_hd;
    const double *vrec;       /* [nrec*3] or NULL (no fits) */
    const int32_t *vrec_td, *vrec_fd, *vrec_ht;
    int64_t nrec;
    /* physical collisional runs instead: the phase-1 number density [m^-3] and
     * speed histograms of every cell, both in the [ht][td][fd] cell order */
    const double *ndens;
    const int64_t *vhist;
} dsmc_phase2_input;

/* Single runs on the current device.  Inputs stay in host memory; the call
 * uploads them, runs, and copies results back. */
int dsmc_run_ballistic(const dsmc_params *p, const dsmc_surface *s, dsmc_result *r);
int dsmc_run_advanced(const dsmc_params *p, const dsmc_surface *s,
                      const dsmc_phase2_input *in, dsmc_result *r);

/* dsmc_run_ballistic without allocating anything: no hipMalloc, hipFree or
 * hipHostMalloc, no stream or event is created, buffers are cleared by a kernel
 * (no memset).  Every device buffer of the run (particle records and queues,
 * tables, grids, per-particle outputs) is taken from one 96 MiB block of device
 * memory that the library holds on each device from load until the process
 * ends, and the work runs on the null stream; the host tables reuse the
 * previous call's storage.  Same inputs, outputs and results as
 * dsmc_run_ballistic.  No counterpart in the reference: it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid allocation).  A run that does not fit the block is rejected with
 * DSMC_ERR_BUDGET and a message (at the 50 x 36 x 72 grid with velocity
 * moments: about 270k particles with every per-particle output, 335k without).
 * Every call on a device uses that device's one block, which keeps the last
 * call's contents: calls overlapping on one GPU would race, so the library
 * runs the calls of a process one at a time. */
int dsmc_run_ballistic_noalloc(const dsmc_params *p, const dsmc_surface *s, dsmc_result *r);

/* Launch shape of the persistent particle kernel. */
typedef struct dsmc_launch_shape {
    int block;                /* threads per block of the flight kernel (64, 128, 192 or 256) */
    int blocks_per_cu;        /* persistent flight blocks per CU */
    int relaunch_batch;       /* split kernels: idle lanes a wavefront gathers before pulling
                                 new particles; service inline: lanes whose hop ended that a
                                 wavefront gathers before running the landing service */
    int step_budget;          /* steps a particle takes per flight launch before yielding (batches:
                                 the first round only; their later rounds run without a budget,
                                 every run being admitted before the first) */
    int service_inline;       /* 1 = landings serviced inside the flight kernel (physical and
                                 compat ballistic runs), 0 = by a separate service kernel per
                                 round (physical advanced runs are always serviced inline,
                                 compat advanced runs always by the separate kernel) */
    int late_step_budget;     /* single runs with the separate service kernel: the step budget of
                                 every round after the first two; 0 = step_budget (pools ignore it) */
    int park_batch;           /* physical ballistic flights: lanes whose step needs the cell logic
                                 wait until this many of a wavefront do (1-64); 0 = the default */
    int late_inline;          /* single physical ballistic runs with the separate service kernel:
                                 k = 1-64: after the first k rounds, the rest of the run in one launch
                                 of the flight kernel with the landing service inline and no step
                                 budget; 0 = rounds of the split kernels to the end (pools ignore it) */
} dsmc_launch_shape;

/* Staged runs: inputs uploaded once and kept resident in device memory, the run
 * relaunched any number of times (outputs are cleared before each launch).
 * `outputs` only says which outputs to allocate (non-NULL pointers). */
typedef struct dsmc_run dsmc_run;
dsmc_run *dsmc_run_create(const dsmc_params *p, const dsmc_surface *s, const dsmc_result *outputs,
                          int *status);
/* Async on the run's stream; shape NULL = the default for the run's mode, body
 * and size. */
int dsmc_run_launch(dsmc_run *run, const dsmc_launch_shape *shape);
int dsmc_run_sync(dsmc_run *run);
double dsmc_run_kernel_ms(dsmc_run *run);        /* HIP-event time of the last launch */
/* Device time spent in the flight kernel and number of rounds of the last launch. */
int dsmc_run_timing(dsmc_run *run, double *flight_ms, int *rounds);
/* The event-batched flight kernel's share of the last launch (physical ballistic
 * runs with the separate service kernel): its device time, its launches and the
 * particle-steps it took, counted on the device (0 when the run took none). */
int dsmc_run_lean_timing(dsmc_run *run, double *lean_ms, int *launches, int64_t *steps);
int dsmc_run_collect(dsmc_run *run, dsmc_result *r);
void dsmc_run_destroy(dsmc_run *run);

/* Diagnostics on the device: independent Verlet steps (integrator.c:4-55) of
 * n states [pos, vel, oa] x 9, optional trajectory [n][nsteps][pos, vel, status];
 * and the device elementary functions (0 sin, 1 cos, 2 asin, 3 atan2(x, y),
 * 4 exp, 5 log, 6 sqrt, 7 x / y, 8 x / y and 9 sqrt(x) for positive operands in
 * [2^-500, 2^500] as the push evaluates them). */
int dsmc_verlet_steps(double *state, int n, int nsteps, double GM, double R, double max_height,
                      double h, double *traj);
int dsmc_elementary(int fn, const double *x, const double *y, double *out, int n);

/* Physical collisional runs.
 *
 * Number density of a phase-1 run: the steps spent in each cell (vcount) x dt x
 * the body's emission rate (production_rate, or the sum of face weights x R^2
 * when it is 0; times sublimation_scale; molecules/s) / (nparticles x cell
 * volume); cells [ht][td][fd], spherical-shell sectors between the grid's
 * radial, latitude and longitude edges. */
int dsmc_number_density(const dsmc_params *p1, const dsmc_surface *s, const int64_t *vcount, double *ndens);
/* Partner-speed fits from phase-1 histograms (device): per cell with more than
 * min_fit_samples steps, the reverse cumulative histogram P = 50 (steps at or
 * above the bin) / steps at the occupied bins, fitted with a degree fit_degree
 * polynomial in the normalised speed (polyfit.c's normal equations).  coeffs
 * [cell*(degree+1)], highest power first; min_v/max_v the occupied range. */
int dsmc_build_hist_fits(const dsmc_params *p, const int64_t *vhist, double *coeffs, double *min_v,
                         double *max_v, int32_t *valid);
/* Phase 1 (p1: physical ballistic, histograms kept) then phase 2 (p2:
 * collisional, dt and grid of p2; p2 must match p1's grid) with the phase-1
 * density and fits handed over on the device.  r1 may be NULL. */
int dsmc_run_collisional(const dsmc_params *p1, const dsmc_params *p2, const dsmc_surface *s, dsmc_result *r1,
                         dsmc_result *r2);

/* Per-cell polynomial CDF fits (advanced.c:32-117 + polyfit.c:10-101) computed on
 * the device from velocity records: outputs [ftd*ffd*hd] cells in the
 * reference's [td][fd][ht] order, coeffs[cell*(degree+1)], highest power first. */
int dsmc_build_fits(const dsmc_params *p, const dsmc_phase2_input *in,
                    int32_t *count_num, double *coeffs, double *min_v, double *max_v,
                    int32_t *valid);

/* ------------------------------------------------------------------ batch */
/* Many independent physical runs (each its own body, species, surface
 * temperatures, seed and particle count; one grid size and icosphere level for
 * all) in one pool of device particle slots: their particles share persistent
 * flight / service kernels whose rounds are replayed from one captured hipGraph
 * until every run has finished.  `tune` selects the launch shape; NULL = the
 * autotuned shape (tuned on a scaled copy of the batch of at least 4M particles
 * or 1/16 of it, cached per device and workload class).  results[i].elapsed_ms
 * = host time from the start of the batch until run i was seen finished. */
int dsmc_run_batch(int nruns, const dsmc_params *params, const dsmc_surface *surfaces,
                   dsmc_result *results, const dsmc_launch_shape *tune);

/* Staged form: inputs prepared and device memory allocated once (and the shape
 * tuned), then launched any number of times.  flags: DSMC_BATCH_PER_PARTICLE
 * keeps per-particle outputs (end_state, steps, fate, hops; 64 B per particle);
 * DSMC_BATCH_DENSITY_ONLY skips the velocity moments (one atomic per cell visit
 * instead of four; results' vmom, if given, are zero).  dsmc_run_batch is
 * density-only when no result asks for vmom. */
typedef struct dsmc_batch dsmc_batch;
enum dsmc_batch_flags {
    DSMC_BATCH_PER_PARTICLE = 1,
    DSMC_BATCH_DENSITY_ONLY = 2
};
dsmc_batch *dsmc_batch_create(int nruns, const dsmc_params *params, const dsmc_surface *surfaces,
                              int flags, const dsmc_launch_shape *tune, int *status);
/* Runs every run to completion; results[i] (or results == NULL) receive run i. */
int dsmc_batch_launch(dsmc_batch *b, dsmc_result *const *results);
/* Host wall time of the last launch, each run's finish time, rounds, shape. */
int dsmc_batch_timing(dsmc_batch *b, double *wall_ms, double *run_done_ms, int64_t *rounds,
                      dsmc_launch_shape *shape);
void dsmc_batch_destroy(dsmc_batch *b);

/* Launch shape for runs like *p (the batch tuner on a one-run pool of at most
 * 2^19 particles of it). */
int dsmc_autotune(const dsmc_params *p, const dsmc_surface *s, dsmc_launch_shape *out);

/* ------------------------------------------------------------------ serving */
/* A continuous-batching server.  Requests (physical runs of one grid size and
 * icosphere level) are prepared on the host at submission, queued in arrival
 * order and admitted into one device pool as soon as particle slots (the HBM
 * budget less max_inflight grid slots, at 276 B per particle, 340 B with
 * per-particle outputs), a grid slot and table slots are free; they join the
 * particles already in flight and complete independently.  A request's
 * per-particle results do not depend on what else runs with it; its grid sums
 * differ only by the order of floating-point atomic additions. */
typedef struct dsmc_server dsmc_server;
typedef struct dsmc_server_stats {
    int64_t completed, particle_steps;     /* requests completed, their particle-steps */
    int64_t queued, running, particles_in_flight, slots, rounds;
    dsmc_launch_shape shape;
} dsmc_server_stats;
enum dsmc_serve_flags {
    DSMC_SERVE_PER_PARTICLE = 1,   /* keep per-particle outputs (64 B per particle slot) */
    DSMC_SERVE_COLLISIONAL = 2     /* accept collisional requests (histogram and fit slots) */
};
dsmc_server *dsmc_server_create(size_t hbm_budget_bytes, int max_inflight, int flags);
int dsmc_server_set_shape(dsmc_server *sv, const dsmc_launch_shape *shape);   /* while idle */
/* r must stay valid until the request completes; r->elapsed_ms receives the time
 * from submission to completion.  Velocity moments are computed only for
 * requests whose r->vmom is set. */
int dsmc_server_submit(dsmc_server *sv, const dsmc_params *p, const dsmc_surface *s,
                       dsmc_result *r, uint64_t *ticket);
/* A physical collisional request: phase 1 (p1, speed histograms kept), then
 * phase 2 (p2, same grid) with the phase-1 density and fits built on the
 * device; r receives the phase-2 results. */
int dsmc_server_submit_collisional(dsmc_server *sv, const dsmc_params *p1, const dsmc_params *p2,
                                   const dsmc_surface *s, dsmc_result *r, uint64_t *ticket);
/* Admits what fits, keeps the device busy, completes finished requests; waits
 * for at most one replay of the round graph.  Returns requests completed. */
int dsmc_server_step(dsmc_server *sv);
int dsmc_server_poll(dsmc_server *sv, uint64_t ticket);   /* 1 done, 0 pending, <0 error */
int dsmc_server_wait(dsmc_server *sv, uint64_t ticket);
int dsmc_server_cancel(dsmc_server *sv, uint64_t ticket);
int dsmc_server_release(dsmc_server *sv, uint64_t ticket);  /* forget a completed request */
int dsmc_server_get_stats(dsmc_server *sv, dsmc_server_stats *out);
void dsmc_server_destroy(dsmc_server *sv);

/* ------------------------------------------------------------------ drop-in */
void run_ballistic(int totaln, const char *temp_file, const char *output_prefix,
                   int nthreads);
void run_advanced(int totaln, const char *temp_file, const char *prev_den_file,
                  const char *prev_v_record_file, const char *output_prefix,
                  int nthreads);

/* The CLI's `pipeline` command (src/main.c:153-204) without its ray-tracing
 * step: the thermal model on 36 x 72 cells written to <prefix>_temperature.dat,
 * then run_ballistic into <prefix>_phase1_*.dat and run_advanced on those into
 * <prefix>_phase2_*.dat.  The same files and messages as those three steps;
 * each phase takes the previous one's outputs from memory, passed through the
 * same text conversion as the files (so it sees the values the files hold)
 * instead of reading the files back.  Returns 0 or a dsmc_status. */
int dsmc_pipeline(int totaln, const char *output_prefix, int nthreads);

/* include/fileio.h:8-44, same semantics and printf formats */
double *read_density_dat(const char *filename, int *ftd, int *ffd, int *hd, int has_hd);
int write_density_dat(const char *filename, const double *den, int ftd, int ffd, int hd);
double *read_temperature_dat(const char *filename, int *ftd, int *ffd);
double *read_temperature_time_dat(const char *filename, int *ftd, int *ffd);
int write_ttime_dat(const char *filename, const double *ttime, int nrows, int totaln);
int write_v_record_dat(const char *filename, const double *v_record, const int *td_list,
                       const int *fd_list, const int *ht_list, int nrecords);
int read_v_record_dat(const char *filename, double **v_record, int **td_list,
                      int **fd_list, int **ht_list, int *nrecords);
double *read_flux_dat(const char *filename, int *nface, int *tt);
int write_flux_dat(const char *filename, const double *F, int nface, int tt);
int write_horizon_dat(const char *filename, const double *horizon, int ftd, int ffd);
/* Host threads of the functions above (they format and parse chunks of a file
 * concurrently; the bytes written and values read do not depend on it): n > 0
 * sets it, 0 restores the default (the machine's cores, at most 16).  Returns
 * the previous setting.  The drop-in drivers use their own nthreads.  No
 * counterpart in the reference, whose file functions are serial. */
int dsmc_io_threads(int n);

/* ------------------------------------------------------------------ thermal model */
/* src/thermal/thermal_model.c: per surface cell (td, fd) a 1-D heat equation in
 * depth (Crank-Nicolson, insulating bottom), the surface temperature found by
 * bisection on the energy balance (absorbed flux + conduction - emission -
 * sublimation) before each substep; on the device, one lane per cell.
 *
 * dsmc_thermal_params: field for field the reference's ThermalParams
 * (include/thermal.h), so its callers link against this library unchanged. */
typedef struct {
    double density;       /* regolith density [kg/m^3] */
    double conductivity;  /* [W/m/K] */
    double Cp;            /* specific heat [J/kg/K] */
    double albedo;
    double distance_AU;
    double depth;         /* [m] */
    int n_substep;        /* substeps per rotation step */
    double z_angle;       /* obliquity [rad] (not read by the solve) */
    double p_mat[5];      /* Antoine vapour pressure coefficients */
    double LH_mat[5];     /* latent heat coefficients */
    double tran_unit_p;   /* Torr to Pa */
    double tran_unit_LH;  /* cal to J */
} dsmc_thermal_params;

/* The reference's SphericalGrid (include/types.h:56-64), same layout. */
typedef struct {
    int ftd, ffd, hd;
    double *density, *temperature, *volume, *horizon;
} dsmc_sph_grid;

/* grid.c:7-17: a grid with every array zeroed (density and volume [hd*ftd*ffd],
 * temperature and horizon [ftd*ffd]); NULL for a negative size or when out of
 * memory.  The reference CLI allocates the thermal model's grid with it. */
dsmc_sph_grid *grid_create(int ftd, int ffd, int hd);
/* grid.c:19-26: frees the grid and its arrays; NULL is ignored. */
void grid_free(dsmc_sph_grid *g);

/* thermal_model.c:23-49 */
void thermal_params_default(dsmc_thermal_params *tp);
/* thermal_model.c:105-283: fills grid->temperature [ftd*ffd]; flux [ftd*ffd]
 * absorbed solar flux or NULL for the built-in cos(lat) sin(lon) day side.
 * Prints the reference's three progress lines; nthreads is unused (device). */
void thermal_solve(dsmc_sph_grid *grid, const dsmc_thermal_params *tp, const double *flux, int nthreads);
/* The same with a status: T_surface [ftd*ffd]; 0 or a dsmc_status. */
int dsmc_thermal_solve(int ftd, int ffd, const dsmc_thermal_params *tp, const double *flux, double *T_surface);

/* ------------------------------------------------------------------ MC shadow flux */
/* src/raytracer/shadow_maker.c: for every face nf and time step t, the sun at
 * longitude (t + 0.5) / tt * 2 pi + pi and elevation -za; F[nf*tt + t] = (1 -
 * shadowed / Nmonte) * max(0, sun . normal), with Nmonte points drawn on the
 * face (glibc rand_r seeded t*nface + nf + 42, as the reference) and each
 * point shadowed when the ray from it (offset 1e-3 along the normal) toward
 * the sun hits any face.  Normals as mesh_compute_normals (outward = away from
 * the origin).  The reference tests rays with Embree in single precision; here
 * the test is double-precision Moller-Trumbore (closed triangles, t > 0) through
 * a bounding-volume hierarchy, one lane per (face, time step).
 * vert [npoint*3]; faces [nface*3], 0-based; F [nface*tt].  0 or a dsmc_status. */
int dsmc_shadow_flux(int npoint, const double *vert, int nface, const int32_t *faces, double za_deg, int tt,
                     int nmonte, double *F);
/* glibc rand_r's sequence from `seed` (n values) as the shadow kernel computes
 * it; no counterpart in the reference (it checks that restatement). */
int dsmc_rand_r_sequence(unsigned int seed, int n, int *out);
/* The same n values, each reached by jumping 3 i generator steps from `seed`
 * (how the kernel's lanes find their sample points). */
int dsmc_rand_r_at(unsigned int seed, int n, int *out);
/* shadow_maker.c:242-271: reads the vertex and (1-based) face files, writes
 * output_file with write_flux_dat; nthreads is unused (device). */
void run_raytracer(const char *mesh_vertex_file, const char *mesh_face_file, int npoint, int nface,
                   double za_deg, int tt, int Nmonte, const char *output_file, int nthreads);

/* ------------------------------------------------------------------ misc */
const char *dsmc_version(void);
int dsmc_device_info(char *buf, size_t len);   /* "gfx950, 256 CUs, ..." */

#ifdef __cplusplus
}
#endif
#endif /* DSMC_HIP_H */
