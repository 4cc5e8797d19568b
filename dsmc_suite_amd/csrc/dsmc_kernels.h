/* dsmc_kernels.h - host-callable launchers of kernels.hip (internal). */
#ifndef DSMC_KERNELS_H
#define DSMC_KERNELS_H

#include <hip/hip_runtime.h>
#include "dsmc_lane.h"

/* Slot queues of a set of runs.  ctl (unsigned int):
 *   [0] fly_count[0] [1] fly_head[0] [2] fly_count[1] [3] fly_head[1]
 *   [4] svc_count    [8] rounds executed
 *   [12..13] particle-steps taken by the event-batched flight kernel (uint64) */
struct DsmcQueues {
    int32_t *fly[2];
    int32_t *svc;
    unsigned int *ctl;
};

/* A pool of runs sharing one set of particle slots and queues (multi-run
 * launches).  Run g's particles carry rec.run = g; envs[g] / steps[g] describe
 * it.  Every run of a pool has the pool's mode and grid dimensions. */
struct DsmcPoolDev {
    DsmcRec *rec;
    const DsmcEnv *envs;       /* [max_runs] full environments (start/service) */
    const DsmcStepEnv *steps;  /* [max_runs] what a flight step reads */
    const double *tq, *cs;     /* latitude thresholds, sector directions (pool dims) */
    const double *hr;          /* radial threshold tables, [nhr] doubles in all */
    int ftd, ffd, hd, nhr;
    int mode, adv;
    int steps_lds;             /* > 0: the pool kernels copy steps[0, steps_lds) into LDS */
};

extern "C" {
/* Starts particles [0, n) of one run in slots slot0.. (records carry run_id). */
hipError_t dsmc_launch_start(const DsmcEnv *E, int32_t run_id, DsmcRec *rec, int64_t slot0, int64_t n,
                             DsmcQueues Q, int p, int grid, hipStream_t st);
/* flight kernel of round p (one run) */
hipError_t dsmc_launch_flight(const DsmcEnv *E, DsmcRec *rec, DsmcQueues Q, int p, int grid, int block,
                              int refill_batch, int64_t step_budget, int park_batch, hipStream_t st);
/* in place of the service kernel of round p: its service queue appended to the
 * next fly queue (for a launch that services landings itself), then the reset */
hipError_t dsmc_launch_handover(DsmcQueues Q, int p, int grid, hipStream_t st);
/* service kernel of round p (one run), then the queue reset */
extern "C" int dsmc_flight_occupancy(const DsmcEnv *E, int inline_svc, int block);
extern "C" int dsmc_pool_occupancy(const DsmcPoolDev *P, int block);
hipError_t dsmc_launch_service(const DsmcEnv *E, DsmcRec *rec, DsmcQueues Q, int p, int grid, hipStream_t st);
/* the same two for a pool of runs */
hipError_t dsmc_launch_pool_flight(const DsmcPoolDev *P, DsmcQueues Q, int p, int grid, int block,
                                   int refill_batch, int64_t step_budget, hipStream_t st);
hipError_t dsmc_launch_pool_service(const DsmcPoolDev *P, DsmcQueues Q, int p, int grid, hipStream_t st);
/* physical runs: flight with the landing service inline, then the queue reset;
 * one run (E on the host, Ed its copy in device memory, P == NULL) or a pool (P) */
hipError_t dsmc_launch_phys(const DsmcEnv *E, const DsmcEnv *Ed, const DsmcPoolDev *P, DsmcRec *rec, DsmcQueues Q,
                            int p, int grid, int block, int svc_batch, int64_t step_budget, hipStream_t st);
/* physical collisional tables: partner-speed fits from speed histograms, and
 * n*sigma (ndens may alias nsig) */
hipError_t dsmc_launch_hist_fits(const unsigned long long *vhist, int ncell, int bins, double dv, int degree,
                                 int min_samples, double *coeffs, double *pmin, double *pmax, int32_t *valid,
                                 int32_t *count_num, hipStream_t st);
hipError_t dsmc_launch_nsig(const double *ndens, const int32_t *valid, double sigma, int ncell, double *nsig,
                            hipStream_t st);
/* a physical run's per-cell records into vmom and vcount (no-op without records) */
hipError_t dsmc_launch_fold(const DsmcEnv *E, int64_t cells, hipStream_t st);
/* n zeros of 8 bytes (graph-safe clear) */
/* counts of a pool run's 32-byte cell records into its vcount */
hipError_t dsmc_launch_fold_counts(const double *acc, unsigned long long *vcount, int64_t cells, hipStream_t st);
hipError_t dsmc_launch_fill0(void *p, size_t n8, hipStream_t st);
hipError_t dsmc_launch_verlet(double *state, int n, int nsteps, double GM, double R,
                              double max_height, double h, double *traj, hipStream_t s);
hipError_t dsmc_launch_math(int fn, const double *x, const double *y, double *out, int n,
                            hipStream_t s);
}

#endif
