/* dsmc_host.h - internal host-side types shared by the library's translation units. */
#ifndef DSMC_HOST_H
#define DSMC_HOST_H

#include <hip/hip_runtime.h>
#include "../../include/dsmc_hip.h"
#include "dsmc_tables.h"
#include "dsmc_kernels.h"
#include <vector>

namespace dsmc {

/* A fixed block of device memory handed out front to back, 256-byte aligned;
 * nothing is allocated or freed. */
struct Arena {
    char *base = nullptr;
    size_t cap = 0, used = 0;
    void *take(size_t n)
    {
        const size_t a = (used + 255) & ~(size_t)255;
        const size_t len = (n + 255) & ~(size_t)255;
        if (a > cap || len > cap - a) return nullptr;
        used = a + len;
        return base + a;
    }
};

/* Device allocation: owned (hipMalloc / hipFree), or borrowed from an Arena. */
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    bool borrowed = false;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf();
    void release();
    int alloc(size_t n);
    int upload(const void *src, size_t n, hipStream_t s);
    int zeros(size_t n, hipStream_t s);
    /* the same three from an arena: no allocation, cleared by a fill kernel */
    int take(Arena &a, size_t n);
    int upload(Arena &a, const void *src, size_t n, hipStream_t s);
    int zeros(Arena &a, size_t n, hipStream_t s);
};

/* A non-blocking stream and two timing events for one call, released on every
 * return path (declare it before the call's RunDev, so the buffers go first). */
struct TimedStream {
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    TimedStream() = default;
    TimedStream(const TimedStream &) = delete;
    TimedStream &operator=(const TimedStream &) = delete;
    int create()
    {
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) { st = nullptr; return DSMC_ERR_DEVICE; }
        if (hipEventCreate(&e0) != hipSuccess) { e0 = nullptr; return DSMC_ERR_DEVICE; }
        if (hipEventCreate(&e1) != hipSuccess) { e1 = nullptr; return DSMC_ERR_DEVICE; }
        return DSMC_OK;
    }
    ~TimedStream()
    {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (st) (void)hipStreamDestroy(st);
    }
};

/* Device-side state of one run. */
struct RunDev {
    DsmcEnv E;
    HostTables T;
    DevBuf tq, hr, cs, vt_speeds, vt_count, Tsurf, cdf, enorm, fnorm, fmap, vert, faces, face_T, face_sig;
    DevBuf density, horizon, vmom, vcount, cellacc, ttime, end_state, steps, fate, hops;
    DevBuf vrec, vrec_i, count_nn, ctl;
    /* advanced */
    DevBuf atm, count_num, coeffs, pmin, pmax, pvalid, speeds, nsig, vhist;
    /* particle records and slot queues */
    DevBuf rec, flyq0, flyq1, svcq, qctl, counters, rec_total, env_d;
    unsigned int *host_ctl = nullptr;     /* pinned readback of qctl */
    Arena *arena = nullptr;               /* device buffers from this block instead of hipMalloc */
    unsigned int ctl_local[16];           /* readback of qctl when nothing may be allocated */
    bool ctl_pinned = false;
    int rounds = 0;
    double flight_ms = 0.0;
    /* the event-batched flight kernel's share of the last launch */
    int lean_launches = 0;
    double lean_ms = 0.0;
    int64_t lean_steps = 0;
    int short_hops = 0;                   /* physical: landings every few dozen steps */
    void drop_buffers();                  /* every buffer released: a reused RunDev starts empty */
    ~RunDev();
};

int ensure_device();
int device_props(int *cus);
int stage_run(const dsmc_params *p, const dsmc_surface *s, const dsmc_result *r, int advanced,
              RunDev *rd, hipStream_t st);
int launch_run(RunDev *rd, const dsmc_launch_shape *shape, hipStream_t st, int timed = 0);
int collect_run(RunDev *rd, const dsmc_params *p, dsmc_result *r, hipStream_t st);
dsmc_launch_shape default_shape(const DsmcEnv &E, int short_hops);
void number_density(const dsmc_params *p1, double wsum, const int64_t *vcount, double *ndens);
bool shape_ok(const dsmc_launch_shape &s);

}  // namespace dsmc

#endif
