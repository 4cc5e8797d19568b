/* dsmc_kernels.h - host-callable launchers of kernels.hip (internal). */
#ifndef DSMC_KERNELS_H
#define DSMC_KERNELS_H

#include <hip/hip_runtime.h>
#include "dsmc_lane.h"

extern "C" {
hipError_t dsmc_launch_particles(const DsmcEnv *E, unsigned long long *next_pid,
                                 unsigned long long *counters, int grid, int block,
                                 int relaunch_batch, hipStream_t stream);
hipError_t dsmc_launch_verlet(double *state, int n, int nsteps, double GM, double R,
                              double max_height, double h, double *traj, hipStream_t s);
hipError_t dsmc_launch_math(int fn, const double *x, const double *y, double *out, int n,
                            hipStream_t s);
}

#endif
