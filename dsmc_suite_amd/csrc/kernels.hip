/*
 * kernels.hip - gfx950 kernels of the DSMC path.
 *
 * dsmc_particles_kernel is persistent: each 64-lane wavefront owns 64 particles
 * at a time and pulls new particle ids from a global counter (one atomic per
 * wave per refill).  The flight step (Stormer-Verlet + cell tracking + deposit,
 * FP64 VALU only) runs for every flying lane; lanes that land wait, masked, until
 * enough of the wave has landed, and then the divergent relaunch work (angles,
 * frozen scan, velocity sampling, trigonometry) runs once for all of them.  The
 * per-lane logic is dsmc_lane.h; this file only schedules it.
 *
 * Deposits are accumulated in registers while a particle stays in one cell and
 * merged with one global FP64 atomic per cell change (the grid is L2-resident).
 */
#include <hip/hip_runtime.h>
#include "dsmc_lane.h"
#include "dsmc_kernels.h"

struct DsmcOpsDev {
    __device__ static inline void add_f64(double *p, double v) { atomicAdd(p, v); }
    __device__ static inline void add_u64(unsigned long long *p, unsigned long long v) { atomicAdd(p, v); }
    __device__ static inline int32_t claim_i32(int32_t *p) { return atomicAdd(p, 1); }
    __device__ static inline unsigned long long claim_u64(unsigned long long *p) { return atomicAdd(p, 1ull); }
};

__device__ static inline uint64_t lanemask_lt()
{
    uint32_t lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

template <int MODE, int ADV>
__global__ void __launch_bounds__(256)
dsmc_particles_kernel(DsmcEnv E, unsigned long long *next_pid, unsigned long long *counters,
                      int relaunch_batch)
{
    const uint32_t lane = __lane_id();
    const uint64_t ltmask = lanemask_lt();
    DsmcLane L;
    L.phase = LANE_IDLE;
    DsmcCounts cnt = {0, 0, 0, 0, 0, 0, 0};
    bool exhausted = false;
    const int64_t N = E.nparticles;

    for (;;) {
        /* refill lanes without a particle (wave-aggregated fetch) */
        while (!exhausted) {
            uint64_t need = __ballot(L.phase == LANE_DONE || L.phase == LANE_IDLE);
            if (need == 0) break;
            int nneed = __popcll(need);
            int leader = __ffsll((long long)need) - 1;
            unsigned long long base = 0;
            if ((int)lane == leader) base = atomicAdd(next_pid, (unsigned long long)nneed);
            base = __shfl(base, leader);
            if ((int64_t)(base + nneed) >= N) exhausted = true;
            if (need & (1ull << lane)) {
                int64_t pid = (int64_t)base + __popcll(need & ltmask);
                if (pid < N) lane_start<DsmcOpsDev, MODE, ADV>(L, E, pid, cnt);
                else L.phase = LANE_IDLE;
            }
        }
        uint64_t flying = __ballot(L.phase == LANE_FLIGHT);
        uint64_t landed = __ballot(L.phase == LANE_LANDED);
        if (flying == 0 && landed == 0) {
            if (exhausted) break;
            continue;
        }
        /* flight: step until enough lanes wait for service */
        int active = __popcll(__ballot(L.phase != LANE_IDLE && L.phase != LANE_DONE));
        int target = active / 2 < relaunch_batch ? active / 2 : relaunch_batch;
        if (target < 1) target = 1;
        while (flying != 0) {
            uint64_t waiting = landed | (exhausted ? 0ull : __ballot(L.phase == LANE_DONE));
            if (__popcll(waiting) >= target) break;
            if (L.phase == LANE_FLIGHT) lane_step<DsmcOpsDev, MODE, ADV>(L, E, cnt);
            flying = __ballot(L.phase == LANE_FLIGHT);
            landed = __ballot(L.phase == LANE_LANDED);
        }
        /* service: landing bookkeeping + relaunch for all waiting lanes */
        if (L.phase == LANE_LANDED) lane_service<DsmcOpsDev, MODE, ADV>(L, E, cnt);
    }
    /* fate and step counters: one atomic per counter per lane with nonzero value */
    if (cnt.esc) atomicAdd(&counters[0], cnt.esc);
    if (cnt.frozen) atomicAdd(&counters[1], cnt.frozen);
    if (cnt.photo) atomicAdd(&counters[2], cnt.photo);
    if (cnt.depleted) atomicAdd(&counters[3], cnt.depleted);
    if (cnt.capped) atomicAdd(&counters[4], cnt.capped);
    if (cnt.coll) atomicAdd(&counters[5], cnt.coll);
    if (cnt.steps) atomicAdd(&counters[6], cnt.steps);
}

/* Plain FP64 Verlet steps of independent states (integrator.c:4-55), for the
 * integrator parity test: state[i*9..] = pos, vel, oa; traj[(i*nsteps+k)*7]. */
__global__ void dsmc_verlet_kernel(double *state, int n, int nsteps, double GM, double R,
                                   double max_height, double h, double *traj)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double *s = state + (size_t)i * 9;
    double x = s[0], y = s[1], z = s[2], vx = s[3], vy = s[4], vz = s[5], ox = s[6], oy = s[7], oz = s[8];
    const double h2 = h * h, c1 = 1.5 * h2, c2 = h2 / 6.0;
    for (int k = 0; k < nsteps; k++) {
        double Ri = sqrt(x * x + y * y + z * z);
        double g = GM / (Ri * Ri * Ri);
        double ax = g * x, ay = g * y, az = g * z;
        double nx = x + vx * h - ax * c1 - ox * c2;
        double ny = y + vy * h - ay * c1 - oy * c2;
        double nz = z + vz * h - az * c1 - oz * c2;
        vx -= ax * h; vy -= ay * h; vz -= az * h;
        ox = ax; oy = ay; oz = az;
        x = nx; y = ny; z = nz;
        double Rn = sqrt(x * x + y * y + z * z);
        double st = Rn <= R ? 1.0 : (Rn >= max_height ? 2.0 : 0.0);
        if (traj) {
            double *t = traj + ((size_t)i * nsteps + k) * 7;
            t[0] = x; t[1] = y; t[2] = z; t[3] = vx; t[4] = vy; t[5] = vz; t[6] = st;
        }
    }
    s[0] = x; s[1] = y; s[2] = z; s[3] = vx; s[4] = vy; s[5] = vz; s[6] = ox; s[7] = oy; s[8] = oz;
}

/* Elementary functions of dsmc_math.h on the device (bitwise host/device check):
 * fn 0 sin, 1 cos, 2 asin, 3 atan2(x,y), 4 exp, 5 log, 6 sqrt, 7 x/y. */
__global__ void dsmc_math_kernel(int fn, const double *x, const double *y, double *out, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a = x[i], b = y ? y[i] : 0.0, r;
    switch (fn) {
    case 0: r = dm_sin(a); break;
    case 1: r = dm_cos(a); break;
    case 2: r = dm_asin(a); break;
    case 3: r = dm_atan2(a, b); break;
    case 4: r = dm_exp(a); break;
    case 5: r = dm_log(a); break;
    case 6: r = sqrt(a); break;
    default: r = a / b; break;
    }
    out[i] = r;
}

/* ------------------------------------------------------------------ launchers */
extern "C" hipError_t dsmc_launch_particles(const DsmcEnv *E, unsigned long long *next_pid,
                                            unsigned long long *counters, int grid, int block,
                                            int relaunch_batch, hipStream_t stream)
{
    const size_t lds = 0;
    if (E->mode == 0 && !E->advanced)
        hipLaunchKernelGGL((dsmc_particles_kernel<0, 0>), dim3(grid), dim3(block), lds, stream, *E,
                           next_pid, counters, relaunch_batch);
    else if (E->mode == 0)
        hipLaunchKernelGGL((dsmc_particles_kernel<0, 1>), dim3(grid), dim3(block), lds, stream, *E,
                           next_pid, counters, relaunch_batch);
    else
        hipLaunchKernelGGL((dsmc_particles_kernel<1, 0>), dim3(grid), dim3(block), lds, stream, *E,
                           next_pid, counters, relaunch_batch);
    return hipGetLastError();
}

extern "C" hipError_t dsmc_launch_verlet(double *state, int n, int nsteps, double GM, double R,
                                         double max_height, double h, double *traj, hipStream_t s)
{
    hipLaunchKernelGGL(dsmc_verlet_kernel, dim3((n + 255) / 256), dim3(256), 0, s, state, n, nsteps,
                       GM, R, max_height, h, traj);
    return hipGetLastError();
}

extern "C" hipError_t dsmc_launch_math(int fn, const double *x, const double *y, double *out, int n,
                                       hipStream_t s)
{
    hipLaunchKernelGGL(dsmc_math_kernel, dim3((n + 255) / 256), dim3(256), 0, s, fn, x, y, out, n);
    return hipGetLastError();
}
