/*
 * philox.h - Philox4x32-10 counter-based generator (Salmon, Moraes, Dror, Shaw,
 * "Parallel random numbers: as easy as 1, 2, 3", SC'11), written from the
 * published round function: two 32x32->64 multiplies per round with the
 * multipliers 0xD2511F53 / 0xCD9E8D57 and the Weyl key increments
 * 0x9E3779B9 / 0xBB67AE85, ten rounds.
 *
 * The DSMC draw stream of one particle is keyed by (seed, particle id, stream):
 *   block  = philox(ctr = {k >> 2, pid_lo, pid_hi, stream}, key = {seed_lo, seed_hi})
 *   word k = block[k & 3]
 * The reference draws with rand_r(); its stream is replaced by
 *   draw31(k) = word(k) >> 1          (uniform on [0, RAND_MAX = 2^31 - 1])
 * so that `(double)draw / RAND_MAX` and `draw % count` keep the reference's
 * arithmetic (ballistic.c:112-113,161-162, velocity.c:66).
 *
 * Compiles as C (gcc) and as HIP device/host code (hipcc).
 */
#ifndef DSMC_PHILOX_H
#define DSMC_PHILOX_H

#include <stdint.h>

#if defined(__HIPCC__)
#define DSMC_RNG_FN __host__ __device__ static inline
#else
#define DSMC_RNG_FN static inline
#endif

#define DSMC_PHILOX_M0 0xD2511F53u
#define DSMC_PHILOX_M1 0xCD9E8D57u
#define DSMC_PHILOX_W0 0x9E3779B9u
#define DSMC_PHILOX_W1 0xBB67AE85u
#define DSMC_RAND_MAX 2147483647

typedef struct { uint32_t w[4]; } dsmc_u32x4;

DSMC_RNG_FN dsmc_u32x4 dsmc_philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2,
                                          uint32_t c3, uint32_t k0, uint32_t k1)
{
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)DSMC_PHILOX_M0 * (uint64_t)c0;
        uint64_t p1 = (uint64_t)DSMC_PHILOX_M1 * (uint64_t)c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += DSMC_PHILOX_W0;
        k1 += DSMC_PHILOX_W1;
    }
    dsmc_u32x4 o;
    o.w[0] = c0; o.w[1] = c1; o.w[2] = c2; o.w[3] = c3;
    return o;
}

/* Block `blk` of the stream of particle `pid`. */
DSMC_RNG_FN dsmc_u32x4 dsmc_rng_block(uint64_t seed, uint64_t pid, uint32_t stream,
                                      uint64_t blk)
{
    return dsmc_philox4x32_10((uint32_t)blk, (uint32_t)pid, (uint32_t)(pid >> 32),
                              stream ^ ((uint32_t)(blk >> 32) << 8),
                              (uint32_t)seed, (uint32_t)(seed >> 32));
}

/* Word k of the stream (no caching; the kernels cache the current block). */
DSMC_RNG_FN uint32_t dsmc_rng_word(uint64_t seed, uint64_t pid, uint32_t stream, uint64_t k)
{
    dsmc_u32x4 b = dsmc_rng_block(seed, pid, stream, k >> 2);
    return b.w[k & 3];
}

/* rand_r() replacement: uniform integer on [0, 2^31 - 1]. */
DSMC_RNG_FN int32_t dsmc_draw31(uint64_t seed, uint64_t pid, uint32_t stream, uint64_t k)
{
    return (int32_t)(dsmc_rng_word(seed, pid, stream, k) >> 1);
}

/* Uniform double on the open interval (0, 1) from two consecutive words
 * (53 random bits, offset by half an ulp so that log() is finite). */
DSMC_RNG_FN double dsmc_u01_from_words(uint32_t hi, uint32_t lo)
{
    return ((double)(hi >> 5) * 67108864.0 + (double)(lo >> 6) + 0.5)
           * (1.0 / 9007199254740992.0);
}

#endif /* DSMC_PHILOX_H */
