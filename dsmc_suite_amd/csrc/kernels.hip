/*
 * kernels.hip - gfx950 kernels of the DSMC path (event-based).
 *
 * A particle lives in a 256-byte record (DsmcRec) in device memory between
 * kernels.  One "round" of a run is three launches on one stream:
 *
 *   flight   persistent; each wavefront pulls particle slots from fly queue p
 *            (one atomic per refill of many lanes), loads only the fields a
 *            Stormer-Verlet step touches, steps every lane until its hop ends
 *            on the surface or its fate is decided, flushes its register deposit
 *            accumulator, stores the record and appends the slot to the service
 *            queue.  Lanes that used up a per-launch step budget go to fly queue
 *            1-p instead (time slicing for the serving loop).
 *   service  one lane per queued slot: landing bookkeeping and relaunch (angles,
 *            frozen scan, emission, trigonometry), or the particle's outputs;
 *            relaunched particles are appended to fly queue 1-p.
 *   reset    clears the consumed queues' counters (a kernel, so rounds can be
 *            captured in a hipGraph and replayed).
 *
 * The flight kernel holds no relaunch code, which keeps its register footprint -
 * and so its occupancy - set by the Verlet step alone.  Every particle's
 * arithmetic is the per-lane state machine of dsmc_lane.h.
 *
 * Pools: many runs share the particle slots and the queues (rec.run names the
 * run).  Pool flight lanes load their run's step constants (DsmcStepEnv) into
 * registers when they pick up a particle, so a wavefront can hold particles of
 * different runs; the per-body radial thresholds of every run live in LDS.
 */
#include <hip/hip_runtime.h>
#include "dsmc_lane.h"
#include "dsmc_kernels.h"
#include <string.h>

static_assert(sizeof(DsmcRec) == 256, "particle record must stay 256 bytes");

/* Deposits straight into the device grid (global FP64 atomics). */
/* Atomics on device memory through global (not generic) addresses.  A generic
 * ("flat") memory instruction also counts against the wavefront's LDS wait
 * counter, so an outstanding flat atomic - a deposit to a contended cell takes
 * microseconds - would hold up the next LDS table read of the flight loop; a
 * global one does not.  Every pointer given to these is device memory. */
#define DSMC_GLOBAL __attribute__((address_space(1)))
__device__ static inline void gadd(double *p, double v)
{
    __hip_atomic_fetch_add((DSMC_GLOBAL double *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ static inline void gadd(unsigned long long *p, unsigned long long v)
{
    __hip_atomic_fetch_add((DSMC_GLOBAL unsigned long long *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ static inline int32_t gclaim(int32_t *p)
{
    return __hip_atomic_fetch_add((DSMC_GLOBAL int32_t *)p, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ static inline unsigned long long gclaim(unsigned long long *p)
{
    return __hip_atomic_fetch_add((DSMC_GLOBAL unsigned long long *)p, 1ull, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
}

template <class EnvT>
__device__ static inline void grid_deposit(const EnvT &E, int cell, double w, double vx, double vy,
                                           double vz, int n)
{
    if (E.cellacc) {      /* physical runs with moments: the cell's 32-byte record */
        double *a = E.cellacc + 4 * (size_t)cell;
        gadd(&a[0], vx); gadd(&a[1], vy); gadd(&a[2], vz); gadd(&a[3], (double)n);
        return;
    }
    if (!E.density_from_count) gadd(&E.density[cell], w);
    if (E.vmom) {
        gadd(&E.vmom[(size_t)cell * 3 + 0], vx);
        gadd(&E.vmom[(size_t)cell * 3 + 1], vy);
        gadd(&E.vmom[(size_t)cell * 3 + 2], vz);
    }
    if (E.vcount) gadd(&E.vcount[cell], (unsigned long long)n);
}

/* The speed histogram of a visit (physical runs that keep one). */
template <class EnvT>
__device__ static inline void visit_hist(const EnvT &E, int cell, double vx, double vy, double vz, int n)
{
    unsigned long long *hb = vhist_bin(E, cell, vx, vy, vz, n);
    if (hb) gadd(hb, (unsigned long long)n);
}

struct DsmcOpsDev {
    template <class EnvT>
    __device__ static inline void deposit(const EnvT &E, int cell, double w, double vx, double vy,
                                          double vz, int n)
    {
        visit_hist(E, cell, vx, vy, vz, n);
        grid_deposit(E, cell, w, vx, vy, vz, n);
    }
    __device__ static inline void add_f64(double *p, double v) { gadd(p, v); }
    __device__ static inline void add_u64(unsigned long long *p, unsigned long long v) { gadd(p, v); }
    __device__ static inline int32_t claim_i32(int32_t *p) { return gclaim(p); }
    __device__ static inline unsigned long long claim_u64(unsigned long long *p) { return gclaim(p); }
};

/* LDS-privatised deposit of the flight kernel: a per-block table of DEP_SLOTS grid
 * cells.  A cell claims the slot its hash selects with one LDS compare-and-swap
 * (slots are never evicted during the launch); deposits of claimed cells are LDS
 * atomics, others fall through to the global grid.  At the end of the launch each
 * block merges its claimed slots into the grid with one global atomic per value. */
#define DEP_SLOTS 512
__shared__ int32_t dep_tag[DEP_SLOTS];
__shared__ int32_t dep_n[DEP_SLOTS];
__shared__ double dep_w[DEP_SLOTS], dep_vx[DEP_SLOTS], dep_vy[DEP_SLOTS], dep_vz[DEP_SLOTS];

__device__ static inline int dep_slot(int cell) { return (int)(((uint32_t)cell * 2654435761u) >> 23); }

struct DsmcOpsLds {
    template <class EnvT>
    __device__ static inline void deposit(const EnvT &E, int cell, double w, double vx, double vy,
                                          double vz, int n)
    {
        visit_hist(E, cell, vx, vy, vz, n);
        int s = dep_slot(cell);
        int t = dep_tag[s];
        if (t == -1) {
            t = atomicCAS(&dep_tag[s], -1, cell);
            if (t == -1) t = cell;
        }
        if (t != cell) { grid_deposit(E, cell, w, vx, vy, vz, n); return; }
        if (!E.density_from_count) atomicAdd(&dep_w[s], w);
        if (E.vmom) { atomicAdd(&dep_vx[s], vx); atomicAdd(&dep_vy[s], vy); atomicAdd(&dep_vz[s], vz); }
        if (E.vcount) atomicAdd(&dep_n[s], n);
    }
    __device__ static inline void add_f64(double *p, double v) { gadd(p, v); }
    __device__ static inline void add_u64(unsigned long long *p, unsigned long long v) { gadd(p, v); }
    __device__ static inline int32_t claim_i32(int32_t *p) { return gclaim(p); }
    __device__ static inline unsigned long long claim_u64(unsigned long long *p) { return gclaim(p); }
};

__device__ static inline void dep_init()
{
    for (int i = threadIdx.x; i < DEP_SLOTS; i += blockDim.x) {
        dep_tag[i] = -1; dep_n[i] = 0;
        dep_w[i] = 0.0; dep_vx[i] = 0.0; dep_vy[i] = 0.0; dep_vz[i] = 0.0;
    }
}

__device__ static inline void dep_merge(const DsmcEnv &E)
{
    for (int i = threadIdx.x; i < DEP_SLOTS; i += blockDim.x)
        if (dep_tag[i] >= 0) grid_deposit(E, dep_tag[i], dep_w[i], dep_vx[i], dep_vy[i], dep_vz[i], dep_n[i]);
}

__device__ static inline uint64_t lanemask_lt()
{
    uint32_t lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

/* Wave-aggregated append of `slot` to queue q (counter *cnt) for lanes with pred. */
__device__ static inline void wave_append(bool pred, int32_t slot, int32_t *q, unsigned int *cnt)
{
    uint64_t m = __ballot(pred);
    if (m == 0) return;
    int leader = __ffsll((long long)m) - 1;
    unsigned int base = 0;
    if ((int)__lane_id() == leader) base = atomicAdd(cnt, (unsigned int)__popcll(m));
    base = __shfl(base, leader);
    if (pred) q[base + __popcll(m & lanemask_lt())] = slot;
}

/* Staged deposits of physical runs with velocity moments (weight 1, density =
 * count, every deposit to the cell's 32-byte record).  A device-scope atomic
 * leaves the L2 as one request per lane, and one request carries a 32-byte
 * sector: four lanes adding the four words of one cell's record (sum vx, sum
 * vy, sum vz, count) in one instruction make one request where one lane adding
 * them in turn makes four (on MI355X, 129600 cells, random cells: 2.36e10
 * against 5.8e9 cell deposits/s).  A lane's deposit goes to one of its own
 * STAGE_SLOTS LDS slots (the record's address and the four values); then, in
 * uniform control flow, the wavefront issues every staged deposit, 16 records
 * per atomic instruction.  Runs without moments deposit one count per visit,
 * one request either way: those go straight to the grid. */
#define STAGE_SLOTS 2
__shared__ double stg_val[256 * STAGE_SLOTS * 4];
__shared__ unsigned long long stg_ptr[256 * STAGE_SLOTS];
__shared__ int32_t stg_cnt[256];
__shared__ int32_t stg_idx[4][64 * STAGE_SLOTS];

struct DsmcOpsStage {
    template <class EnvT>
    __device__ static inline void deposit(const EnvT &E, int cell, double w, double vx, double vy, double vz,
                                          int n)
    {
        visit_hist(E, cell, vx, vy, vz, n);
        const int t = threadIdx.x, k = stg_cnt[t];
        if (!E.cellacc || k >= STAGE_SLOTS) { grid_deposit(E, cell, w, vx, vy, vz, n); return; }
        const int e = t * STAGE_SLOTS + k;
        stg_ptr[e] = (unsigned long long)(E.cellacc + 4 * (size_t)cell);
        stg_val[4 * e + 0] = vx; stg_val[4 * e + 1] = vy; stg_val[4 * e + 2] = vz; stg_val[4 * e + 3] = (double)n;
        stg_cnt[t] = k + 1;
    }
    __device__ static inline void add_f64(double *p, double v) { gadd(p, v); }
    __device__ static inline void add_u64(unsigned long long *p, unsigned long long v) { gadd(p, v); }
    __device__ static inline int32_t claim_i32(int32_t *p) { return gclaim(p); }
    __device__ static inline unsigned long long claim_u64(unsigned long long *p) { return gclaim(p); }
};

__device__ static inline void stage_init() { stg_cnt[threadIdx.x] = 0; }

/* Issues the wavefront's staged deposits; every lane of the wavefront calls it.
 * List: slot 0 of every lane that has one, then slot 1.  LAZY (after a batch of
 * cell logic, which stages at most one record per lane): only once some lane's
 * slots are full or 16 records (one full atomic instruction) wait. */
template <bool LAZY = false>
__device__ static inline void stage_issue()
{
    const int t = threadIdx.x, lane = (int)__lane_id();
    const int k = stg_cnt[t];
    const uint64_t m1 = __ballot(k >= 1);
    if (m1 == 0) return;
    const uint64_t m2 = __ballot(k >= 2);
    if (LAZY && m2 == 0 && __popcll(m1) < 16) return;
    const int n1 = __popcll(m1), tot = n1 + __popcll(m2);
    const uint64_t lt = lanemask_lt();
    int32_t *idx = stg_idx[t >> 6];
    if (k >= 1) idx[__popcll(m1 & lt)] = t * STAGE_SLOTS;
    if (k >= 2) idx[n1 + __popcll(m2 & lt)] = t * STAGE_SLOTS + 1;
    stg_cnt[t] = 0;
    /* the other lanes' slots and list entries: LDS operations of one wavefront
     * complete in order; the fence keeps the compiler from moving the reads up */
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const int word = lane & 3;
    for (int g = 0; g < tot; g += 16) {
        const int r = g + (lane >> 2);
        if (r < tot) {
            const int e = idx[r];
            gadd((double *)stg_ptr[e] + word, stg_val[4 * e + word]);
        }
    }
    __builtin_amdgcn_wave_barrier();
}

/* Fate / step counters of finished particles, aggregated over the lanes of the
 * wavefront that share a counter block (one run): one atomic per counter and
 * run instead of one per lane.  Called by every lane of the wavefront. */
__device__ static inline void counts_flush_wave(DsmcCounts &cnt, unsigned long long *ctr, bool active)
{
    active = active && ctr != nullptr;
    for (;;) {
        uint64_t act = __ballot(active);
        if (act == 0) break;
        const int leader = __ffsll((long long)act) - 1;
        unsigned long long *lead = (unsigned long long *)__shfl((long long)ctr, leader);
        const bool mine = active && ctr == lead;
        unsigned long long v[8] = {cnt.esc, cnt.frozen, cnt.photo, cnt.depleted, cnt.capped, cnt.coll, cnt.steps,
                                   cnt.done};
#pragma unroll
        for (int i = 0; i < 8; i++) {
            unsigned long long x = mine ? v[i] : 0ull;
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) x += (unsigned long long)__shfl_xor((long long)x, o);
            if ((int)__lane_id() == leader && x) gadd(&lead[i], x);
        }
        active = active && !mine;
    }
    cnt.esc = cnt.frozen = cnt.photo = cnt.depleted = cnt.capped = cnt.coll = cnt.steps = cnt.done = 0;
}

/* Slots a lane holds for one wavefront-wide append (one atomic on the queue's
 * counter per flush instead of one per wavefront iteration: a single address
 * serves ~88M atomics/s on MI355X, and the service of a 1e7-particle round
 * would otherwise issue one per 64 of its ~1e7 slots). */
struct AppendBuf {
    int32_t s0, s1, s2, s3;
    int n;
};

__device__ static inline void append_flush(AppendBuf &A, int32_t *q, unsigned int *cnt)
{
    const uint64_t m0 = __ballot(A.n > 0);
    if (m0 == 0) return;
    const uint64_t m1 = __ballot(A.n > 1), m2 = __ballot(A.n > 2), m3 = __ballot(A.n > 3);
    const unsigned int c0 = __popcll(m0), c1 = __popcll(m1), c2 = __popcll(m2), c3 = __popcll(m3);
    const int leader = __ffsll((long long)m0) - 1;
    unsigned int base = 0;
    if ((int)__lane_id() == leader) base = atomicAdd(cnt, c0 + c1 + c2 + c3);
    base = __shfl(base, leader);
    const uint64_t lt = lanemask_lt();
    if (A.n > 0) q[base + __popcll(m0 & lt)] = A.s0;
    if (A.n > 1) q[base + c0 + __popcll(m1 & lt)] = A.s1;
    if (A.n > 2) q[base + c0 + c1 + __popcll(m2 & lt)] = A.s2;
    if (A.n > 3) q[base + c0 + c1 + c2 + __popcll(m3 & lt)] = A.s3;
    A.n = 0;
}

__device__ static inline void append_push(AppendBuf &A, bool pred, int32_t slot, int32_t *q, unsigned int *cnt)
{
    if (pred) {
        if (A.n == 0) A.s0 = slot;
        else if (A.n == 1) A.s1 = slot;
        else if (A.n == 2) A.s2 = slot;
        else A.s3 = slot;
        A.n++;
    }
    if (__ballot(A.n == 4)) append_flush(A, q, cnt);
}

/* ------------------------------------------------------------------ start */
/* Starts particles [0, n) of one run in slots slot0..slot0+n-1. */
template <int MODE, int ADV>
__global__ void __launch_bounds__(256)
start_kernel(DsmcEnv E, int32_t run_id, DsmcRec *rec, int64_t slot0, int64_t n, int32_t *flyq,
             unsigned int *fly_count)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    AppendBuf A;
    A.n = 0;
    for (int64_t b = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u); b < n; b += stride) {
        int64_t i = b + __lane_id();
        bool fly = false;
        int64_t slot = slot0 + i;
        if (i < n) {
            DsmcLane L;
            DsmcCounts cnt = {0, 0, 0, 0, 0, 0, 0, 0};
            lane_start<DsmcOpsDev, MODE, ADV>(L, E, i, cnt);
            counts_flush<DsmcOpsDev>(cnt, E);
            fly = L.phase == LANE_FLIGHT;
            if (fly) rec_store_all<MODE>(L, rec[slot], run_id);
        }
        append_push(A, fly, (int32_t)slot, flyq, fly_count);
    }
    append_flush(A, flyq, fly_count);
}

/* ------------------------------------------------------------------ flight */
/* Physical ballistic flights: at most 96 VGPRs, 5 waves per SIMD (measured
 * +5% over 4 on the Ceres 1e7 run); the other variants keep their registers. */
template <int MODE, int ADV>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MODE == 1 && ADV == 0 ? 5 : 1)))
flight_kernel(DsmcEnv E, DsmcRec *rec, DsmcQueues Q, int p, int refill_batch, int64_t step_budget)
{
    /* per-step tables and the deposit slot table -> LDS */
    extern __shared__ double lds_tabs[];
    DsmcCellTabs TB;
    {
        const int ntq = E.ftd + 1, nhr = E.hd + 2, ncs = 2 * (E.ffd + 1);
        for (int i = threadIdx.x; i < ntq + nhr + ncs; i += blockDim.x)
            lds_tabs[i] = i < ntq ? E.tq[i] : (i < ntq + nhr ? E.hr[i - ntq] : E.cs[i - ntq - nhr]);
        TB.tq = lds_tabs;
        TB.hr = lds_tabs + ntq;
        TB.cs = lds_tabs + ntq + nhr;
        dep_init();
        __syncthreads();
    }
    const unsigned int nq = Q.ctl[2 * p];           /* fly queue p is not appended to in this launch */
    unsigned int *head = &Q.ctl[2 * p + 1];
    const int32_t *flyq = Q.fly[p];
    const uint32_t lane = __lane_id();
    DsmcLane L;
    L.phase = LANE_IDLE;
    int32_t slot = -1;
    int64_t budget_end = 0;
    bool exhausted = false;
    for (;;) {
        /* refill idle lanes together (one atomic per wave) */
        uint64_t idle = __ballot(L.phase == LANE_IDLE);
        uint64_t flying = ~idle & __ballot(true);
        if (!exhausted && idle && (__popcll(idle) >= refill_batch || flying == 0)) {
            int leader = __ffsll((long long)idle) - 1;
            unsigned int base = 0;
            if ((int)lane == leader) base = atomicAdd(head, (unsigned int)__popcll(idle));
            base = __shfl(base, leader);
            if (base + (unsigned int)__popcll(idle) >= nq) exhausted = true;
            if (idle & (1ull << lane)) {
                unsigned int qi = base + __popcll(idle & lanemask_lt());
                if (qi < nq) {
                    slot = flyq[qi];
                    rec_load_flight<MODE, ADV>(L, rec[slot]);
                    budget_end = step_budget > 0 ? L.steps + step_budget : INT64_MAX;
                }
            }
            flying = __ballot(L.phase == LANE_FLIGHT);
        }
        if (flying == 0) {
            if (exhausted) break;
            continue;
        }
        /* step until enough lanes have ended their hop */
        for (;;) {
            /* two steps per vote of the wavefront (a lane whose hop ended in the
             * first one idles through the second): +1.6% on the Ceres 1e7 run */
#pragma unroll
            for (int u = 0; u < 2; u++) {
                if (L.phase == LANE_FLIGHT) {
                    lane_step<DsmcOpsLds, MODE, ADV>(L, E, TB);
                    if (L.phase == LANE_FLIGHT && L.steps >= budget_end) L.phase = LANE_IDLE + 8;
                }
            }
            uint64_t ended = __ballot(L.phase != LANE_FLIGHT && L.phase != LANE_IDLE);
            if (ended && (__popcll(ended | idle) >= refill_batch || __ballot(L.phase == LANE_FLIGHT) == 0)) break;
        }
        /* hand ended hops to the service queue, budget-stopped lanes to the next round */
        bool hop_end = L.phase == LANE_LANDED || L.phase == LANE_ENDED;
        bool paused = L.phase == LANE_IDLE + 8;
        if (hop_end || paused) {
            /* physical records keep an open visit across launches */
            if (hop_end || MODE == 0) lane_flush<DsmcOpsLds>(L, E);
            if (paused) L.phase = LANE_FLIGHT;
            rec_store_flight<MODE, ADV>(L, rec[slot]);
        }
        wave_append(hop_end, slot, Q.svc, &Q.ctl[4]);
        wave_append(paused, slot, Q.fly[1 - p], &Q.ctl[2 * (1 - p)]);
        if (hop_end || paused) L.phase = LANE_IDLE;
    }
    __syncthreads();
    dep_merge(E);
}

/* Physical ballistic flights (the default path of large single runs), event-
 * batched.  Every lane holds its cell's table entries in registers and takes
 * lean steps (lane_fast_phys: the push and a few comparisons) while nothing but
 * its step count and open visit change.  A step that needs more - a cell
 * boundary within the margins, a landing, escape, photolysis, the step cap or
 * the launch's budget, or the first deposit of a visit - parks the lane with its
 * push done; once `pend_batch` lanes of the wavefront are parked (or every
 * flying lane is), they run the rest of their step together (lane_tail_phys,
 * the code every lane of the wavefront used to pass through on almost every
 * step, since some lane crosses a cell boundary on most of them) and reload
 * their cell's entries.  Same particles, same steps, same deposits as
 * flight_kernel<1, 0>: a parked lane only waits.  The visits those steps close
 * are staged in LDS and issued by the whole wavefront (stage_issue: four lanes
 * per cell record with moments), 158 against 182 ms of flight kernel per 1e7-
 * particle Ceres run with one lane's four atomics per visit. */
/* lean steps between two votes of the wavefront: on the Ceres 1e7 run 2 / 4 /
 * 6 / 8 / 12 / 16 take 157 / 145 / 139 / 137 / 141 / 145 ms of flight kernel
 * (a parked lane idles for the rest of its group; the vote costs every lane) */
#define LEAN_STEPS 8
/* pools: their short hops park lanes every ~18 steps, so a vote costs more per
 * step (the 64-run batch: 4 / 6 / 8 / 10 / 12 / 16 / 20 lean steps between
 * votes, 1056 / 978 / 945 / 929 / 922 / 928 / 946 ms) */
#define LEAN_STEPS_POOL 12
#define LEAN_PARK_BATCH 4
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5)))
lean_flight_kernel(DsmcEnv E, DsmcRec *rec, DsmcQueues Q, int p, int refill_batch, int64_t step_budget,
                   int pend_batch)
{
    extern __shared__ double lds_tabs[];
    DsmcCellTabs TB;
    {
        const int ntq = E.ftd + 1, nhr = E.hd + 2, ncs = 2 * (E.ffd + 1);
        for (int i = threadIdx.x; i < ntq + nhr + ncs; i += blockDim.x)
            lds_tabs[i] = i < ntq ? E.tq[i] : (i < ntq + nhr ? E.hr[i - ntq] : E.cs[i - ntq - nhr]);
        TB.tq = lds_tabs;
        TB.hr = lds_tabs + ntq;
        TB.cs = lds_tabs + ntq + nhr;
        stage_init();
        __syncthreads();
    }
    const unsigned int nq = Q.ctl[2 * p];
    unsigned int *head = &Q.ctl[2 * p + 1];
    const int32_t *flyq = Q.fly[p];
    const uint32_t lane = __lane_id();
    const int64_t cap = E.max_steps > 0 ? E.max_steps : INT64_MAX;
    DsmcLane L;
    L.phase = LANE_IDLE;
    LeanLane N;
    N.pend = false;
    N.easy = false;
    N.stop = 0;
    int32_t slot = -1;
    int64_t budget_end = 0;
    bool exhausted = false;
    /* particle-steps taken in this launch, modulo 2^32 per lane (each particle's
     * steps at its store minus its steps at its load) */
    uint32_t taken = 0;
    for (;;) {
        uint64_t idle = __ballot(L.phase == LANE_IDLE);
        uint64_t flying = ~idle & __ballot(true);
        if (!exhausted && idle && (__popcll(idle) >= refill_batch || flying == 0)) {
            int leader = __ffsll((long long)idle) - 1;
            unsigned int base = 0;
            if ((int)lane == leader) base = atomicAdd(head, (unsigned int)__popcll(idle));
            base = __shfl(base, leader);
            if (base + (unsigned int)__popcll(idle) >= nq) exhausted = true;
            if (idle & (1ull << lane)) {
                unsigned int qi = base + __popcll(idle & lanemask_lt());
                if (qi < nq) {
                    slot = flyq[qi];
                    rec_load_flight<1, 0>(L, rec[slot]);
                    taken -= (uint32_t)L.steps;
                    budget_end = step_budget > 0 ? L.steps + step_budget : INT64_MAX;
                    N.stop = budget_end < cap ? budget_end : cap;
                    N.pend = false;
                    lean_refresh(N, L, E, TB);
                }
            }
            flying = __ballot(L.phase == LANE_FLIGHT);
        }
        if (flying == 0) {
            if (exhausted) break;
            continue;
        }
        for (;;) {
            /* `act` stays a lane mask across the unrolled steps */
            const bool fl = L.phase == LANE_FLIGHT;
            bool act = fl && !N.pend;
#pragma unroll
            for (int u = 0; u < LEAN_STEPS; u++)
                if (act) act = lane_fast_phys(L, E, N);
            N.pend = fl && !act;
            const uint64_t parked = __ballot(N.pend);
            if (parked && (__popcll(parked) >= pend_batch || parked == __ballot(fl))) {
                if (N.pend) {
                    lean_tail<DsmcOpsStage>(L, E, TB, N);
                    N.pend = false;
                    if (L.phase == LANE_FLIGHT) {
                        if (L.steps >= budget_end) L.phase = LANE_IDLE + 8;
                        else lean_refresh(N, L, E, TB);
                    }
                }
                stage_issue<true>();
            }
            uint64_t ended = __ballot(L.phase != LANE_FLIGHT && L.phase != LANE_IDLE);
            if (ended && (__popcll(ended | idle) >= refill_batch || __ballot(L.phase == LANE_FLIGHT) == 0)) break;
        }
        bool hop_end = L.phase == LANE_LANDED || L.phase == LANE_ENDED;
        bool paused = L.phase == LANE_IDLE + 8;
        if (hop_end || paused) {
            if (hop_end) lane_flush<DsmcOpsStage>(L, E);   /* a pause keeps the open visit */
            if (paused) L.phase = LANE_FLIGHT;
            rec_store_flight<1, 0>(L, rec[slot]);
            taken += (uint32_t)L.steps;
        }
        stage_issue();
        wave_append(hop_end, slot, Q.svc, &Q.ctl[4]);
        wave_append(paused, slot, Q.fly[1 - p], &Q.ctl[2 * (1 - p)]);
        if (hop_end || paused) L.phase = LANE_IDLE;
    }
    /* the launch's particle-steps into ctl[12..13] (64-bit, per run) */
    unsigned long long wsum = taken;
    for (int o = 32; o > 0; o >>= 1) wsum += __shfl_xor(wsum, o);
    if (lane == 0 && wsum) atomicAdd((unsigned long long *)&Q.ctl[12], wsum);
}

/* Flight kernel of a pool: as above, with each lane's run constants in registers
 * and every run's radial thresholds in LDS (deposits go straight to the run's
 * grid; the LDS slot table is for single runs). */
template <int MODE, int ADV>
__global__ void __launch_bounds__(256)
pool_flight_kernel(DsmcPoolDev P, DsmcQueues Q, int p, int refill_batch, int64_t step_budget)
{
    static_assert(MODE == 1 && ADV == 0, "pools run physical ballistic particles");
    extern __shared__ double lds_tabs[];
    const int ntq = P.ftd + 1, ncs = 2 * (P.ffd + 1);
    for (int i = threadIdx.x; i < ntq + ncs + P.nhr; i += blockDim.x)
        lds_tabs[i] = i < ntq ? P.tq[i] : (i < ntq + ncs ? P.cs[i - ntq] : P.hr[i - ntq - ncs]);
    __syncthreads();
    const double *lds_hr = lds_tabs + ntq + ncs;
    DsmcCellTabs TB;
    TB.tq = lds_tabs;
    TB.cs = lds_tabs + ntq;
    TB.hr = lds_hr;

    const unsigned int nq = Q.ctl[2 * p];
    unsigned int *head = &Q.ctl[2 * p + 1];
    const int32_t *flyq = Q.fly[p];
    DsmcRec *rec = P.rec;
    const uint32_t lane = __lane_id();
    DsmcLane L;
    L.phase = LANE_IDLE;
    DsmcStepEnv S = {};
    int32_t slot = -1;
    int64_t budget_end = 0;
    bool exhausted = false;
    for (;;) {
        uint64_t idle = __ballot(L.phase == LANE_IDLE);
        uint64_t flying = ~idle & __ballot(true);
        if (!exhausted && idle && (__popcll(idle) >= refill_batch || flying == 0)) {
            int leader = __ffsll((long long)idle) - 1;
            unsigned int base = 0;
            if ((int)lane == leader) base = atomicAdd(head, (unsigned int)__popcll(idle));
            base = __shfl(base, leader);
            if (base + (unsigned int)__popcll(idle) >= nq) exhausted = true;
            if (idle & (1ull << lane)) {
                unsigned int qi = base + __popcll(idle & lanemask_lt());
                if (qi < nq) {
                    slot = flyq[qi];
                    S = P.steps[rec[slot].run];
                    TB.hr = lds_hr + S.hr_off;
                    rec_load_flight<MODE, ADV>(L, rec[slot]);
                    budget_end = step_budget > 0 ? L.steps + step_budget : INT64_MAX;
                }
            }
            flying = __ballot(L.phase == LANE_FLIGHT);
        }
        if (flying == 0) {
            if (exhausted) break;
            continue;
        }
        S.ftd = P.ftd; S.ffd = P.ffd; S.hd = P.hd;    /* pool-wide: kept scalar */
        for (;;) {
            if (L.phase == LANE_FLIGHT) {
                lane_step<DsmcOpsDev, MODE, ADV>(L, S, TB);
                if (L.phase == LANE_FLIGHT && L.steps >= budget_end) L.phase = LANE_IDLE + 8;
            }
            uint64_t ended = __ballot(L.phase != LANE_FLIGHT && L.phase != LANE_IDLE);
            if (ended && (__popcll(ended | idle) >= refill_batch || __ballot(L.phase == LANE_FLIGHT) == 0)) break;
        }
        bool hop_end = L.phase == LANE_LANDED || L.phase == LANE_ENDED;
        bool paused = L.phase == LANE_IDLE + 8;
        if (hop_end || paused) {
            if (hop_end) lane_flush<DsmcOpsDev>(L, S);     /* pauses keep the open visit */
            if (paused) L.phase = LANE_FLIGHT;
            rec_store_flight<MODE, ADV>(L, rec[slot]);
        }
        wave_append(hop_end, slot, Q.svc, &Q.ctl[4]);
        wave_append(paused, slot, Q.fly[1 - p], &Q.ctl[2 * (1 - p)]);
        if (hop_end || paused) L.phase = LANE_IDLE;
    }
}

/* The physical landing service.  The run's environment is copied from device
 * memory once, at the top (its fields load together instead of one dependent
 * load per use).  phys_service is the out-of-line form (single runs: the flight
 * kernel's registers stay those of the Verlet step; the lane state travels
 * through the call frame); pools inline the body. */
template <class Ops, int ADV, bool FMN = false>
__device__ __forceinline__ void phys_service_body(DsmcLane &T, const DsmcEnv *Ep, DsmcCounts &cnt,
                                                  const DsmcCellTabs TB)
{
    DsmcEnv E = *Ep;
    /* the cell walks of a relaunch read the kernel's LDS copies of the tables */
    E.tq = TB.tq;
    E.cs = TB.cs;
    E.hr = TB.hr;
    lane_service<Ops, 1, ADV, FMN>(T, E, cnt);
}

template <class Ops, int ADV>
__device__ __attribute__((noinline)) void phys_service(DsmcLane &T, const DsmcEnv *Ep, DsmcCounts &cnt,
                                                       const DsmcCellTabs TB)
{
    phys_service_body<Ops, ADV>(T, Ep, cnt, TB);
}

/* The compat landing service (relaunch with the frozen scan, polar deposit,
 * outputs of a finished particle), out of line for the same reason. */
__device__ __attribute__((noinline)) void compat_service(DsmcLane &T, const DsmcEnv *Ep, DsmcCounts &cnt)
{
    const DsmcEnv E = *Ep;
    lane_service<DsmcOpsLds, 0, 0>(T, E, cnt);
}

/* A collision (physical collisional runs), out of line with only what it reads
 * and writes passed in registers: velocity, random-stream counter, cell. */
struct CollideOut {
    double vx, vy, vz, col_p;
    uint64_t k;
    int num_col;
};

__device__ __attribute__((noinline)) CollideOut phys_collide(double vx, double vy, double vz, double col_p, uint64_t k,
                                                             int64_t pid, int ht, int td, int fd, const DsmcEnv *Ep)
{
    const DsmcEnv E = *Ep;
    DsmcLane T;
    T.vx = vx; T.vy = vy; T.vz = vz; T.col_p = col_p;
    T.k = k; T.pid = pid; T.blk = ~0ull;
    T.ht = ht; T.td = td; T.fd = fd;
    T.num_col = 0;
    lane_collide<true>(T, E);
    CollideOut o = {T.vx, T.vy, T.vz, T.col_p, T.k, T.num_col};
    return o;
}

/* Lanes waiting for a collision that start a collision round in the flight
 * loop: single runs 48 (a round's latency, its loads of the cell's fit and the
 * run's environment, is paid once for more lanes: a 2e5-particle Moon run 265
 * against 343 ms of phase 2 with 16; Europa 302 against 377, Mercury 348
 * against 407), pools 16 (the served mix: no difference). */
#define COLLIDE_BATCH_SINGLE 48
#define COLLIDE_BATCH_POOL 16

__device__ __forceinline__ void lane_collide_lean(DsmcLane &L, const DsmcEnv *Ep)
{
    CollideOut o = phys_collide(L.vx, L.vy, L.vz, L.col_p, L.k, L.pid, L.ht, L.td, L.fd, Ep);
    L.vx = o.vx; L.vy = o.vy; L.vz = o.vz;
    L.k = o.k; L.blk = ~0ull;
    L.col_p = o.col_p;
    L.real_col_p = 0.0;
    L.num_col += o.num_col;
    L.phase = LANE_FLIGHT;
}

/* Physical particles: flight with the landing service inline.  A lane whose hop
 * ends lands (freeze-out or re-emission) or finishes its particle in the same
 * launch, so a particle makes as many hops per launch as its step budget
 * allows; only budget-paused particles wait for the next launch.  Lanes whose
 * hops ended wait (masked) until `svc_batch` lanes of the wavefront need the
 * service, so the service code runs for many lanes at once; lanes whose
 * particle finished take a new one from the queue at every service round (a
 * wavefront that waited for `svc_batch` idle lanes before refilling kept ~40%
 * of its lanes empty on the 64-run batch and ran 20% slower).  POOL: lanes hold
 * particles of different runs (step constants in registers, every run's radial
 * thresholds in LDS); otherwise one run's environment is the kernel argument and
 * deposits go through the LDS slot table. */
/* Pools: up to 168 VGPRs (3 waves per SIMD), so that the lanes' per-run step
 * constants do not spill inside the flight loop (a reload waits behind the
 * wavefront's outstanding deposit atomics); single runs: 128 (4 waves). */
template <bool POOL, int MODE, int ADV, bool STAGE, bool SLDS = false>
__device__ __forceinline__ void inline_service_loop(const DsmcEnv &E0, const DsmcEnv *Ed, const DsmcPoolDev &P,
                                                    DsmcRec *rec, const DsmcQueues &Q, int p, int svc_batch,
                                                    int64_t step_budget)
{
    static_assert(MODE == 1 || (!POOL && !ADV), "pools hold physical runs; compat inline is ballistic");
    extern __shared__ double lds_tabs[];
    const int ftd = POOL ? P.ftd : E0.ftd, ffd = POOL ? P.ffd : E0.ffd, hd = POOL ? P.hd : E0.hd;
    const int ntq = ftd + 1, ncs = 2 * (ffd + 1), nhr = POOL ? P.nhr : hd + 2;
    const double *gtq = POOL ? P.tq : E0.tq, *gcs = POOL ? P.cs : E0.cs, *ghr = POOL ? P.hr : E0.hr;
    for (int i = threadIdx.x; i < ntq + ncs + nhr; i += blockDim.x)
        lds_tabs[i] = i < ntq ? gtq[i] : (i < ntq + ncs ? gcs[i - ntq] : ghr[i - ntq - ncs]);
    /* SLDS: the pool's run table behind the cell tables (the lanes' loads of
     * their run's step constants for the cell logic then wait for no global
     * access, such as the deposit atomics in flight) */
    DsmcStepEnv *lds_steps = reinterpret_cast<DsmcStepEnv *>(lds_tabs + ntq + ncs + nhr);
    if constexpr (SLDS) {
        static_assert(sizeof(DsmcStepEnv) % 8 == 0, "the run table is copied in 8-byte words");
        const unsigned long long *src = reinterpret_cast<const unsigned long long *>(P.steps);
        unsigned long long *dst = reinterpret_cast<unsigned long long *>(lds_steps);
        const int nw = P.steps_lds * (int)(sizeof(DsmcStepEnv) / 8);
        for (int i = threadIdx.x; i < nw; i += blockDim.x) dst[i] = src[i];
    }
    static_assert(!STAGE || MODE == 1, "staged deposits are physical");
    if constexpr (STAGE) stage_init();
    else if constexpr (!POOL) dep_init();
    __syncthreads();
    const double *lds_hr = lds_tabs + ntq + ncs;
    DsmcCellTabs TB;
    TB.tq = lds_tabs;
    TB.cs = lds_tabs + ntq;
    TB.hr = lds_hr;
    if (POOL) rec = P.rec;

    const unsigned int nq = Q.ctl[2 * p];
    unsigned int *head = &Q.ctl[2 * p + 1];
    const int32_t *flyq = Q.fly[p];
    const uint32_t lane = __lane_id();
    /* physical ballistic particles take lean steps (lane_fast_phys) and run the
     * full step logic in batches of parked lanes, as lean_flight_kernel does
     * (1e6-particle Ceres run +33%, the 64 runs of the batch one by one +22%,
     * the 64-run pool +10%, +16% with moments).  A pool lane holds only the
     * lean steps' run constants (LeanEnv) and reads its run's full step
     * constants again for the cell logic: with all of them in registers the
     * lean steps spilled and the pool ran 22% slower */
    constexpr bool LEAN = MODE == 1 && ADV == 0;
    DsmcLane L;
    L.phase = LANE_IDLE;
    DsmcStepEnv S = {};
    LeanLane N;
    LeanEnv FE = {};          /* pools: the lean steps' run constants */
    N.pend = false;
    N.easy = false;
    N.stop = 0;
    int32_t slot = -1, run = 0;
    int64_t budget_end = 0;
    bool exhausted = false;
    for (;;) {
        uint64_t idle = __ballot(L.phase == LANE_IDLE);
        uint64_t flying = ~idle & __ballot(true);
        if (!exhausted && idle) {
            int leader = __ffsll((long long)idle) - 1;
            unsigned int base = 0;
            if ((int)lane == leader) base = atomicAdd(head, (unsigned int)__popcll(idle));
            base = __shfl(base, leader);
            if (base + (unsigned int)__popcll(idle) >= nq) exhausted = true;
            if (idle & (1ull << lane)) {
                unsigned int qi = base + __popcll(idle & lanemask_lt());
                if (qi < nq) {
                    slot = flyq[qi];
                    if constexpr (POOL && LEAN) {
                        run = rec[slot].run;
                        TB.hr = lds_hr + (SLDS ? lds_steps[run].hr_off : P.steps[run].hr_off);
                    } else if (POOL) {
                        run = rec[slot].run;
                        S = P.steps[run];
                        TB.hr = lds_hr + S.hr_off;
                    }
                    if (MODE == 0) rec_load_all<0>(L, rec[slot]);
                    else rec_load_phys(L, rec[slot]);
                    budget_end = step_budget > 0 ? L.steps + step_budget : INT64_MAX;
                    if constexpr (LEAN && POOL) {
                        const DsmcStepEnv Sr = SLDS ? lds_steps[run] : P.steps[run];
                        const int64_t cap = Sr.max_steps > 0 ? Sr.max_steps : INT64_MAX;
                        N.stop = budget_end < cap ? budget_end : cap;
                        N.pend = false;
                        lean_refresh(N, L, Sr, TB);
                    } else if constexpr (LEAN) {
                        const int64_t cap = E0.max_steps > 0 ? E0.max_steps : INT64_MAX;
                        N.stop = budget_end < cap ? budget_end : cap;
                        N.pend = false;
                        lean_refresh(N, L, E0, TB);
                    }
                }
            }
            flying = __ballot(L.phase != LANE_IDLE);    /* lanes with work: a loaded particle may wait for its service */
        }
        if (flying == 0) {
            if (exhausted) break;
            continue;
        }
        if (POOL && !LEAN) { S.ftd = P.ftd; S.ffd = P.ffd; S.hd = P.hd; }
        /* pools: the lean steps' run constants, read again (LDS) for each flight
         * phase rather than held through the service */
        if constexpr (POOL && LEAN) FE = lean_env(SLDS ? lds_steps[run] : P.steps[run]);
        for (;;) {
          if constexpr (LEAN) {
            const bool fl = L.phase == LANE_FLIGHT;
            bool act = fl && !N.pend;
#pragma unroll
            for (int u = 0; u < (POOL ? LEAN_STEPS_POOL : LEAN_STEPS); u++) {
                if (act) {
                    if constexpr (POOL) act = lane_fast_phys(L, FE, N);
                    else act = lane_fast_phys(L, E0, N);
                }
            }
            N.pend = fl && !act;
            const uint64_t parked = __ballot(N.pend);
            if (parked && (__popcll(parked) >= LEAN_PARK_BATCH || parked == __ballot(fl))) {
                if (N.pend) {
                    if constexpr (POOL) {
                        /* the run's step constants, read again for the cell logic */
                        DsmcStepEnv Sr = SLDS ? lds_steps[run] : P.steps[run];
                        Sr.ftd = P.ftd; Sr.ffd = P.ffd; Sr.hd = P.hd;
                        if constexpr (STAGE) lean_tail<DsmcOpsStage>(L, Sr, TB, N);
                        else lean_tail<DsmcOpsDev>(L, Sr, TB, N);
                        N.pend = false;
                        if (L.phase == LANE_FLIGHT) {
                            if (L.steps >= budget_end) L.phase = LANE_IDLE + 8;
                            else lean_refresh(N, L, Sr, TB);
                        }
                    } else {
                        if constexpr (STAGE) lean_tail<DsmcOpsStage>(L, E0, TB, N);
                        else lean_tail<DsmcOpsLds>(L, E0, TB, N);
                        N.pend = false;
                        if (L.phase == LANE_FLIGHT) {
                            if (L.steps >= budget_end) L.phase = LANE_IDLE + 8;
                            else lean_refresh(N, L, E0, TB);
                        }
                    }
                }
                if constexpr (STAGE) stage_issue<true>();
            }
          } else {
            /* two steps per vote, as in flight_kernel (+2-4% on pools, single
             * short-hop and collisional runs) */
#pragma unroll
            for (int u = 0; u < 2; u++) {
                if (L.phase == LANE_FLIGHT) {
                    if constexpr (POOL && STAGE) lane_step<DsmcOpsStage, 1, ADV>(L, S, TB);
                    else if constexpr (POOL) lane_step<DsmcOpsDev, 1, ADV>(L, S, TB);
                    else if constexpr (STAGE) lane_step<DsmcOpsStage, 1, ADV>(L, E0, TB);
                    else lane_step<DsmcOpsLds, MODE, ADV>(L, E0, TB);
                    if (L.phase == LANE_FLIGHT && L.steps >= budget_end) L.phase = LANE_IDLE + 8;
                }
            }
            /* collisions in a dense gas come every few steps: once COLLIDE_BATCH_*
             * lanes wait for one they take the lean path here, without ending
             * the flight loop; fewer wait for the service below */
            if (MODE == 1 && ADV &&
                __popcll(__ballot(L.phase == LANE_COLLIDE)) >= (POOL ? COLLIDE_BATCH_POOL : COLLIDE_BATCH_SINGLE) &&
                L.phase == LANE_COLLIDE)
                lane_collide_lean(L, POOL ? P.envs + run : Ed);
            if constexpr (STAGE) stage_issue();
          }
            uint64_t ended = __ballot(L.phase != LANE_FLIGHT && L.phase != LANE_IDLE);
            uint64_t waiting = ended | __ballot(L.phase == LANE_IDLE);
            if ((ended && __popcll(waiting) >= svc_batch) || waiting == __ballot(true)) break;
        }
        /* the remaining collisions: the lean path (the deposit visit stays open) */
        if (MODE == 1 && ADV && __ballot(L.phase == LANE_COLLIDE) && L.phase == LANE_COLLIDE)
            lane_collide_lean(L, POOL ? P.envs + run : Ed);
        /* landings and ended particles: service inline */
        DsmcCounts cnt = {0, 0, 0, 0, 0, 0, 0, 0};
        const bool svc = L.phase == LANE_LANDED || L.phase == LANE_ENDED || L.phase == LANE_COLLIDE;
        if (svc) {
            DsmcLane T = L;
            if constexpr (MODE == 0) compat_service(T, Ed, cnt);
            /* pools: the service inline (the lane state needs no call frame; the
             * 64-run batch 8.96e10 against 8.62e10 particle-steps/s); single runs
             * out of line (inline, the 1e6-particle Ceres run took 38 against 30 ms) */
            else if constexpr (POOL && STAGE) phys_service_body<DsmcOpsStage, ADV, true>(T, P.envs + run, cnt, TB);
            else if constexpr (POOL) phys_service_body<DsmcOpsDev, ADV, true>(T, P.envs + run, cnt, TB);
            else if constexpr (STAGE) phys_service<DsmcOpsStage, ADV>(T, Ed, cnt, TB);
            else phys_service<DsmcOpsLds, ADV>(T, Ed, cnt, TB);
            L = T;
            if constexpr (LEAN) {
                if (L.phase == LANE_FLIGHT) {       /* relaunched: a new hop from its new cell */
                    N.pend = false;
                    if constexpr (POOL) {
                        DsmcStepEnv Sr = SLDS ? lds_steps[run] : P.steps[run];
                        Sr.ftd = P.ftd; Sr.ffd = P.ffd; Sr.hd = P.hd;
                        lean_refresh_flags(N, L, Sr);
                    } else {
                        lean_refresh(N, L, E0, TB);
                    }
                }
            }
        }
        /* pools: every lane's lean bounds are computed again from its cell (LDS
         * tables) after the service round rather than held through it, so that
         * the service has their registers; a lane without a particle in flight
         * computes them for cell 0 and never reads them */
        if constexpr (POOL && LEAN) {
            const bool fl = L.phase == LANE_FLIGHT;
            const DsmcStepEnv &Sr = SLDS ? lds_steps[run] : P.steps[run];
            N.B = lean_bounds(fl ? L.td : 0, fl ? L.fd : 0, fl ? L.ht : 0, Sr, TB);
        }
        if constexpr (STAGE) stage_issue();
        const bool fin = svc && L.phase == LANE_DONE;
        if (__ballot(fin)) counts_flush_wave(cnt, fin ? (POOL ? P.envs[run].counters : E0.counters) : nullptr, fin);
        if (fin) L.phase = LANE_IDLE;
        /* budget-paused particles go to the next launch; a physical one takes its
         * open deposit visit with it (a launch boundary does not split a visit),
         * a compat one flushes it, as the split flight kernel does */
        bool paused = L.phase == LANE_IDLE + 8;
        if (paused) {
            L.phase = LANE_FLIGHT;
            if constexpr (MODE == 0) {
                lane_flush<DsmcOpsLds>(L, E0);
                rec_store_all<0>(L, rec[slot], 0);
            } else {
                rec_store_phys(L, rec[slot]);
            }
        }
        wave_append(paused, slot, Q.fly[1 - p], &Q.ctl[2 * (1 - p)]);
        if (paused) L.phase = LANE_IDLE;
    }
    if constexpr (!POOL && !STAGE) {
        __syncthreads();
        dep_merge(E0);
    }
}

/* STAGE: deposits to cell records staged and issued wavefront-wide (pools, and
 * single runs with velocity moments); single runs without moments keep the
 * LDS slot table for their count deposits. */
template <bool POOL, int ADV, bool STAGE, bool SLDS = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(POOL ? 3 : 4)))
phys_kernel(DsmcEnv E0, const DsmcEnv *Ed, DsmcPoolDev P, DsmcRec *rec, DsmcQueues Q, int p, int svc_batch,
            int64_t step_budget)
{
    inline_service_loop<POOL, 1, ADV, STAGE, SLDS>(E0, Ed, P, rec, Q, p, svc_batch, step_budget);
}

/* Compat ballistic particles (the reference's run_ballistic) with the landing
 * service inline: each of a particle's hops (hundreds; many of one step) would
 * otherwise cost a round of launches, so a small run's time would be its
 * most-hopping particle's hop count times a round's latency.  (Compat advanced
 * runs keep the split kernels: inline they were 10% slower.) */
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
compat_kernel(DsmcEnv E0, const DsmcEnv *Ed, DsmcRec *rec, DsmcQueues Q, int p, int svc_batch, int64_t step_budget)
{
    DsmcPoolDev none = {};
    inline_service_loop<false, 0, 0, false>(E0, Ed, none, rec, Q, p, svc_batch, step_budget);
}

/* ------------------------------------------------------------------ service */
/* One lane per queued slot.  POOL: the record names its run's environment.  A
 * single run's fate counters gather over the whole launch (one atomic per
 * counter and wavefront at its end); relaunched slots are appended four per
 * lane at a time. */
template <int MODE, int ADV, bool POOL>
__global__ void __launch_bounds__(256)
service_kernel(DsmcEnv E0, const DsmcEnv *envs, DsmcRec *rec, DsmcQueues Q, int p)
{
    const unsigned int n = Q.ctl[4];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    AppendBuf A;
    A.n = 0;
    DsmcCounts cnt = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t b = (int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u); b < (int64_t)n; b += stride) {
        int64_t i = b + __lane_id();
        bool fly = false, fin = false;
        int32_t slot = 0;
        unsigned long long *ctr = nullptr;
        if (i < n) {
            slot = Q.svc[i];
            int run = POOL ? rec[slot].run : 0;
            const DsmcEnv E = POOL ? envs[run] : E0;      /* one wide load per lane in a pool */
            DsmcLane L;
            rec_load_all<MODE>(L, rec[slot]);
            lane_service<DsmcOpsDev, MODE, ADV>(L, E, cnt);
            fly = L.phase == LANE_FLIGHT;
            fin = L.phase == LANE_DONE;
            ctr = E.counters;
            if (fly) rec_store_all<MODE>(L, rec[slot], run);
        }
        if (POOL && __ballot(fin)) counts_flush_wave(cnt, ctr, fin);
        append_push(A, fly, slot, Q.fly[1 - p], &Q.ctl[2 * (1 - p)]);
    }
    append_flush(A, Q.fly[1 - p], &Q.ctl[2 * (1 - p)]);
    if (!POOL) counts_flush_wave(cnt, E0.counters, true);
}

__global__ void fill0_kernel(unsigned long long *p, size_t n)
{
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0ull;
}

/* Adds a physical run's per-cell records (sum vx, sum vy, sum vz, count) to its
 * moments and counts.  The counts are sums of integers below 2^53,
 * exact in double precision. */
__global__ void __launch_bounds__(256) fold_cellacc_kernel(const double *acc, double *vmom, unsigned long long *vcount,
                                                           int64_t cells)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += stride) {
        const double *a = acc + 4 * c;
        const double n = a[3];
        if (n != 0.0) {
            vmom[3 * c + 0] += a[0];
            vmom[3 * c + 1] += a[1];
            vmom[3 * c + 2] += a[2];
            vcount[c] += (unsigned long long)n;
        }
    }
}

/* Counts of cell records (fourth word, an exact integer below 2^53) into vcount. */
__global__ void __launch_bounds__(256) fold_counts_kernel(const double *acc, unsigned long long *vcount, int64_t cells)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += stride) {
        const double n = acc[4 * c + 3];
        if (n != 0.0) vcount[c] += (unsigned long long)n;
    }
}

/* ------------------------------------------------------------------ reset */
/* The round before a launch with the landing service inline: the slots whose hop
 * ended (the service queue) join the next round's fly queue behind the paused
 * ones, which start first (the longest particles, the run's critical path).
 * The reset that follows adds their count. */
__global__ void handover_kernel(DsmcQueues Q, int p)
{
    const unsigned int n = Q.ctl[4], base = Q.ctl[2 * (1 - p)];
    int32_t *dst = Q.fly[1 - p] + base;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (int64_t)n; i += stride) dst[i] = Q.svc[i];
}

/* add_svc: the service queue was handed over to the next fly queue. */
__global__ void reset_kernel(unsigned int *ctl, int p, unsigned int *rounds, int add_svc)
{
    if (threadIdx.x == 0) {
        if (add_svc) ctl[2 * (1 - p)] += ctl[4];
        ctl[2 * p] = 0;
        ctl[2 * p + 1] = 0;
        ctl[4] = 0;
        ctl[5] = 0;
        rounds[0] += 1;
    }
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
    double Ri = sqrt(x * x + y * y + z * z);
    for (int k = 0; k < nsteps; k++) {
        /* the kernels' own step (dm_div_pos, dm_sqrt_pos inside) */
        const double Rn = verlet_push<false>(x, y, z, vx, vy, vz, ox, oy, oz, Ri, GM, h, c1, c2, nullptr);
        Ri = Rn;
        double st = Rn <= R ? 1.0 : (Rn >= max_height ? 2.0 : 0.0);
        if (traj) {
            double *t = traj + ((size_t)i * nsteps + k) * 7;
            t[0] = x; t[1] = y; t[2] = z; t[3] = vx; t[4] = vy; t[5] = vz; t[6] = st;
        }
    }
    s[0] = x; s[1] = y; s[2] = z; s[3] = vx; s[4] = vy; s[5] = vz; s[6] = ox; s[7] = oy; s[8] = oz;
}

/* Elementary functions of dsmc_math.h on the device (bitwise host/device check):
 * fn 0 sin, 1 cos, 2 asin, 3 atan2(x,y), 4 exp, 5 log, 6 sqrt, 7 x/y,
 * 8 dm_div_pos(x, y), 9 dm_sqrt_pos(x). */
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
    case 8: r = dm_div_pos(a, b); break;
    case 9: r = dm_sqrt_pos(a); break;
    default: r = a / b; break;
    }
    out[i] = r;
}

/* ------------------------------------------------------------------ launchers */
#define DISPATCH(KERNEL, ...)                                                                 \
    do {                                                                                      \
        if (mode == 0 && !adv) hipLaunchKernelGGL((KERNEL<0, 0>), __VA_ARGS__);               \
        else if (mode == 0) hipLaunchKernelGGL((KERNEL<0, 1>), __VA_ARGS__);                  \
        else if (!adv) hipLaunchKernelGGL((KERNEL<1, 0>), __VA_ARGS__);                       \
        else hipLaunchKernelGGL((KERNEL<1, 1>), __VA_ARGS__);                                 \
    } while (0)

static bool block_ok(int block) { return block >= 64 && block <= 256 && block % 64 == 0; }

extern "C" hipError_t dsmc_launch_start(const DsmcEnv *E, int32_t run_id, DsmcRec *rec, int64_t slot0, int64_t n,
                                        DsmcQueues Q, int p, int grid, hipStream_t st)
{
    const int mode = E->mode, adv = E->advanced;
    DISPATCH(start_kernel, dim3(grid), dim3(256), 0, st, *E, run_id, rec, slot0, n, Q.fly[p], &Q.ctl[2 * p]);
    return hipGetLastError();
}

extern "C" hipError_t dsmc_launch_flight(const DsmcEnv *E, DsmcRec *rec, DsmcQueues Q, int p, int grid, int block,
                                         int refill_batch, int64_t step_budget, int park_batch, hipStream_t st)
{
    if (!block_ok(block)) return hipErrorInvalidValue;
    const int mode = E->mode, adv = E->advanced;
    const size_t lds = sizeof(double) * (size_t)((E->ftd + 1) + (E->hd + 2) + 2 * (E->ffd + 1));
    if (mode == 1 && !adv)
        hipLaunchKernelGGL(lean_flight_kernel, dim3(grid), dim3(block), lds, st, *E, rec, Q, p, refill_batch,
                           step_budget, park_batch > 0 ? park_batch : LEAN_PARK_BATCH);
    else
        DISPATCH(flight_kernel, dim3(grid), dim3(block), lds, st, *E, rec, Q, p, refill_batch, step_budget);
    return hipGetLastError();
}

extern "C" hipError_t dsmc_launch_handover(DsmcQueues Q, int p, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(handover_kernel, dim3(grid), dim3(256), 0, st, Q, p);
    hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(64), 0, st, Q.ctl, p, Q.ctl + 8, 1);
    return hipGetLastError();
}

/* Blocks of `block` threads of the flight kernel (inline_svc: the kernel with
 * the landing service inline) that one CU holds at once, 0 on error. */
extern "C" int dsmc_flight_occupancy(const DsmcEnv *E, int inline_svc, int block)
{
    if (!block_ok(block)) return 0;
    int nb = 0;
    hipError_t e = hipErrorInvalidValue;
    const size_t lds = sizeof(double) * (size_t)((E->ftd + 1) + (E->hd + 2) + 2 * (E->ffd + 1));
    const int mode = E->mode, adv = E->advanced;
    if (inline_svc && mode == 1) {
        const bool stg = E->cellacc != nullptr;
        if (adv && stg) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, phys_kernel<false, 1, true>, block, lds);
        else if (adv) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, phys_kernel<false, 1, false>, block, lds);
        else if (stg) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, phys_kernel<false, 0, true>, block, lds);
        else e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, phys_kernel<false, 0, false>, block, lds);
    } else if (inline_svc && !adv) {
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, compat_kernel, block, lds);
    } else if (mode == 0 && !adv) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, flight_kernel<0, 0>, block, lds);
    else if (mode == 0) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, flight_kernel<0, 1>, block, lds);
    else if (!adv) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, lean_flight_kernel, block, lds);
    else e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, flight_kernel<1, 1>, block, lds);
    return e == hipSuccess ? nb : 0;
}

/* Blocks of `block` threads of a pool's flight kernel with the landing service
 * inline that one CU holds at once (its LDS as launched), 0 on error. */
extern "C" int dsmc_pool_occupancy(const DsmcPoolDev *P, int block)
{
    if (!block_ok(block)) return 0;
    const size_t lds = sizeof(double) * (size_t)((P->ftd + 1) + 2 * (P->ffd + 1) + P->nhr);
    int nb = 0;
    hipError_t e;
    if (P->adv) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, phys_kernel<true, 1, true>, block, lds);
    else if (P->steps_lds > 0)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, phys_kernel<true, 0, true, true>, block,
                                                         lds + (size_t)P->steps_lds * sizeof(DsmcStepEnv));
    else e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, phys_kernel<true, 0, true>, block, lds);
    return e == hipSuccess ? nb : 0;
}

extern "C" hipError_t dsmc_launch_service(const DsmcEnv *E, DsmcRec *rec, DsmcQueues Q, int p, int grid,
                                          hipStream_t st)
{
    const int mode = E->mode, adv = E->advanced;
    if (mode == 0 && !adv) hipLaunchKernelGGL((service_kernel<0, 0, false>), dim3(grid), dim3(256), 0, st, *E,
                                               nullptr, rec, Q, p);
    else if (mode == 0) hipLaunchKernelGGL((service_kernel<0, 1, false>), dim3(grid), dim3(256), 0, st, *E,
                                           nullptr, rec, Q, p);
    else if (!adv) hipLaunchKernelGGL((service_kernel<1, 0, false>), dim3(grid), dim3(256), 0, st, *E,
                                      nullptr, rec, Q, p);
    else return hipErrorInvalidValue;
    hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(64), 0, st, Q.ctl, p, Q.ctl + 8, 0);
    return hipGetLastError();
}

extern "C" hipError_t dsmc_launch_pool_flight(const DsmcPoolDev *P, DsmcQueues Q, int p, int grid, int block,
                                              int refill_batch, int64_t step_budget, hipStream_t st)
{
    if (!block_ok(block) || P->mode != 1 || P->adv) return hipErrorInvalidValue;
    const size_t lds = sizeof(double) * (size_t)((P->ftd + 1) + 2 * (P->ffd + 1) + P->nhr);
    hipLaunchKernelGGL((pool_flight_kernel<1, 0>), dim3(grid), dim3(block), lds, st, *P, Q, p, refill_batch,
                       step_budget);
    return hipGetLastError();
}

extern "C" hipError_t dsmc_launch_phys(const DsmcEnv *E, const DsmcEnv *Ed, const DsmcPoolDev *P, DsmcRec *rec,
                                       DsmcQueues Q, int p, int grid, int block, int svc_batch, int64_t step_budget,
                                       hipStream_t st)
{
    if (!block_ok(block)) return hipErrorInvalidValue;
    if (P) {
        if (P->mode != 1) return hipErrorInvalidValue;
        const size_t lds = sizeof(double) * (size_t)((P->ftd + 1) + 2 * (P->ffd + 1) + P->nhr);
        DsmcEnv none;
        memset(&none, 0, sizeof none);
        if (P->adv)
            hipLaunchKernelGGL((phys_kernel<true, 1, true>), dim3(grid), dim3(block), lds, st, none, (const DsmcEnv *)nullptr,
                               *P, rec, Q, p, svc_batch, step_budget);
        else if (P->steps_lds > 0)
            hipLaunchKernelGGL((phys_kernel<true, 0, true, true>), dim3(grid), dim3(block),
                               lds + (size_t)P->steps_lds * sizeof(DsmcStepEnv), st, none, (const DsmcEnv *)nullptr,
                               *P, rec, Q, p, svc_batch, step_budget);
        else
            hipLaunchKernelGGL((phys_kernel<true, 0, true>), dim3(grid), dim3(block), lds, st, none, (const DsmcEnv *)nullptr,
                               *P, rec, Q, p, svc_batch, step_budget);
    } else if (E->mode == 0) {
        if (!Ed || E->advanced) return hipErrorInvalidValue;
        const size_t lds = sizeof(double) * (size_t)((E->ftd + 1) + 2 * (E->ffd + 1) + (E->hd + 2));
        hipLaunchKernelGGL(compat_kernel, dim3(grid), dim3(block), lds, st, *E, Ed, rec, Q, p, svc_batch,
                           step_budget);
    } else {
        if (E->mode != 1 || !Ed) return hipErrorInvalidValue;
        const size_t lds = sizeof(double) * (size_t)((E->ftd + 1) + 2 * (E->ffd + 1) + (E->hd + 2));
        DsmcPoolDev none;
        memset(&none, 0, sizeof none);
#define PHYS_SINGLE(A, S) hipLaunchKernelGGL((phys_kernel<false, A, S>), dim3(grid), dim3(block), lds, st, *E, Ed, \
                                           none, rec, Q, p, svc_batch, step_budget)
        const bool stg = E->cellacc != nullptr;
        if (E->advanced && stg) PHYS_SINGLE(1, true);
        else if (E->advanced) PHYS_SINGLE(1, false);
        else if (stg) PHYS_SINGLE(0, true);
        else PHYS_SINGLE(0, false);
#undef PHYS_SINGLE
    }
    hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(64), 0, st, Q.ctl, p, Q.ctl + 8, 0);
    return hipGetLastError();
}

extern "C" hipError_t dsmc_launch_pool_service(const DsmcPoolDev *P, DsmcQueues Q, int p, int grid, hipStream_t st)
{
    if (P->mode != 1 || P->adv) return hipErrorInvalidValue;
    DsmcEnv none;
    memset(&none, 0, sizeof none);
    hipLaunchKernelGGL((service_kernel<1, 0, true>), dim3(grid), dim3(256), 0, st, none, P->envs, P->rec, Q, p);
    hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(64), 0, st, Q.ctl, p, Q.ctl + 8, 0);
    return hipGetLastError();
}

extern "C" hipError_t dsmc_launch_fill0(void *p, size_t n8, hipStream_t st)
{
    if (n8 == 0) return hipSuccess;
    size_t blocks = (n8 + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(fill0_kernel, dim3((unsigned)blocks), dim3(256), 0, st, (unsigned long long *)p, n8);
    return hipGetLastError();
}

extern "C" hipError_t dsmc_launch_fold(const DsmcEnv *E, int64_t cells, hipStream_t st)
{
    if (!E->cellacc || cells <= 0) return hipSuccess;
    if (!E->vmom || !E->vcount) return hipErrorInvalidValue;
    int64_t blocks = (cells + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(fold_cellacc_kernel, dim3((unsigned)blocks), dim3(256), 0, st, (const double *)E->cellacc,
                       E->vmom, E->vcount, cells);
    return hipGetLastError();
}

extern "C" hipError_t dsmc_launch_fold_counts(const double *acc, unsigned long long *vcount, int64_t cells,
                                               hipStream_t st)
{
    if (!acc || !vcount || cells <= 0) return hipErrorInvalidValue;
    int64_t blocks = (cells + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(fold_counts_kernel, dim3((unsigned)blocks), dim3(256), 0, st, acc, vcount, cells);
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
