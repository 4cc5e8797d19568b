/* dsmc_kernels.h - host-callable launchers of kernels.hip (internal). */
#ifndef DSMC_KERNELS_H
#define DSMC_KERNELS_H

#include <hip/hip_runtime.h>
#include "dsmc_lane.h"

/* Slot queues of a set of runs.  ctl (unsigned int):
 *   [0] fly_count[0] [1] fly_head[0] [2] fly_count[1] [3] fly_head[1]
 *   [4] svc_count    [8] rounds executed */
struct DsmcQueues {
    int32_t *fly[2];
    int32_t *svc;
    unsigned int *ctl;
};

extern "C" {
hipError_t dsmc_launch_start(const DsmcEnv *E0, const DsmcEnv *envs, const int64_t *run_base, int nruns,
                             DsmcRec *rec, int64_t slot0, int64_t nslots, DsmcQueues Q, int p, int grid,
                             hipStream_t st);
/* flight kernel of round p */
hipError_t dsmc_launch_flight(const DsmcEnv *E0, const DsmcEnv *envs, DsmcRec *rec, DsmcQueues Q, int p,
                              int grid, int refill_batch, int64_t step_budget, hipStream_t st);
/* service kernel of round p, then the queue reset */
hipError_t dsmc_launch_service(const DsmcEnv *E0, const DsmcEnv *envs, DsmcRec *rec, DsmcQueues Q, int p,
                               int grid, hipStream_t st);
hipError_t dsmc_launch_verlet(double *state, int n, int nsteps, double GM, double R,
                              double max_height, double h, double *traj, hipStream_t s);
hipError_t dsmc_launch_math(int fn, const double *x, const double *y, double *out, int n,
                            hipStream_t s);
}

#endif
