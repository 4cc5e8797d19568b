/*
 * dsmc_math.h - deterministic FP64 elementary functions for the DSMC path.
 *
 * The device kernels and the host must produce bit-identical results for the
 * same inputs.  Vendor libms (glibc on the host, ocml on the device) agree only
 * to an ulp or so, so every transcendental the kernels evaluate goes through
 * the functions below.  They use only IEEE-754 operations that are correctly
 * rounded on both sides (+, -, *, /, sqrt, fma, rint), in a fixed order, so a
 * host build (gcc, -ffp-contract=off) and a gfx950 build (hipcc,
 * -ffp-contract=off) return the same bits.
 *
 * Methods (all textbook; Cody & Waite, "Software Manual for the Elementary
 * Functions", 1980, for the argument reductions):
 *   dm_sincos  reduction by pi/2 with a three-part Cody-Waite constant, then
 *              Taylor polynomials of sin (to r^17) and cos (to r^16) on
 *              |r| <= pi/4 (truncation < 2^-60 relative).  Domain |x| < 2^20.
 *   dm_atan2   octant reduction to t in [0,1], then atan(t) = atan(c_i) +
 *              atan((t - c_i)/(1 + t c_i)) with c_i = tan(i pi/16) (i = 0..4,
 *              atan(c_i) stored as a double-double), Taylor series of atan on
 *              |u| <= tan(pi/32) to u^17.
 *   dm_asin    asin(x) = atan2(x, sqrt((1-x)(1+x))).
 *   dm_exp     x = k ln2 + r, |r| <= ln2/2, Taylor series to r^13, scale by 2^k.
 *   dm_log     x = 2^e m, m in [sqrt(1/2), sqrt(2)), log(m) = 2 atanh(f/(2+f))
 *              written as f - s (f - R) (s = f/(2+f)), series to s^21.
 * Accuracy against correctly rounded results is about 1 ulp (tests measure it
 * against glibc on millions of arguments).
 *
 * Compiles as C (gcc) and as HIP host/device code (hipcc).
 */
#ifndef DSMC_MATH_H
#define DSMC_MATH_H

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define DM_FN __host__ __device__ static inline
#else
#define DM_FN static inline
#endif

/* pi/2, pi and 2 pi as three-part sums (hi exact double nearest, then the
 * rounded remainders); M_PI of the reference equals DM_PI_1. */
#define DM_PIO2_1 0x1.921fb54442d18p+0
#define DM_PIO2_2 0x1.1a62633145c07p-54
#define DM_PIO2_3 (-0x1.f1976b7ed8fbcp-110)
#define DM_PI_1 0x1.921fb54442d18p+1
#define DM_PI_2 0x1.1a62633145c07p-53
#define DM_TWO_OVER_PI 0x1.45f306dc9c883p-1
#define DM_LN2_1 0x1.62e42fefa39efp-1
#define DM_LN2_2 0x1.abc9e3b39803fp-56
#define DM_LN2_HI32 0x1.62e42p-1            /* ln2 with the low 32 mantissa bits cleared */
#define DM_LN2_LO32 0x1.fdf473de6af28p-22   /* ln2 - DM_LN2_HI32 */
#define DM_INV_LN2 0x1.71547652b82fep+0
#define DM_SQRT2 0x1.6a09e667f3bcdp+0

typedef union { double d; uint64_t u; } dm_bits;

DM_FN uint64_t dm_as_u64(double x) { dm_bits b; b.d = x; return b.u; }
DM_FN double dm_from_u64(uint64_t u) { dm_bits b; b.u = u; return b.d; }

/* x * 2^k for |k| <= 2100, rounding once. */
DM_FN double dm_ldexp(double x, int k)
{
    if (k > 1023) { x *= 0x1p1023; k -= 1023; if (k > 1023) k = 1023; }
    else if (k < -1022) { x *= 0x1p-600; k += 600; if (k < -1022) { x *= 0x1p-600; k += 600; } }
    if (k < -1022) k = -1022;
    return x * dm_from_u64((uint64_t)(k + 1023) << 52);
}

/* ---------------------------------------------------------------- sin / cos */
DM_FN double dm__sin_kern(double r)
{
    double z = r * r;
    double p = 0x1.952c77030ad4ap-49;          /*  1/17! */
    p = fma(p, z, -0x1.ae7f3e733b81fp-41);      /* -1/15! */
    p = fma(p, z, 0x1.6124613a86d09p-33);       /*  1/13! */
    p = fma(p, z, -0x1.ae64567f544e4p-26);      /* -1/11! */
    p = fma(p, z, 0x1.71de3a556c734p-19);       /*  1/9!  */
    p = fma(p, z, -0x1.a01a01a01a01ap-13);      /* -1/7!  */
    p = fma(p, z, 0x1.1111111111111p-7);        /*  1/5!  */
    p = fma(p, z, -0x1.5555555555555p-3);       /* -1/3!  */
    return fma(r * z, p, r);
}

DM_FN double dm__cos_kern(double r)
{
    double z = r * r;
    double p = 0x1.ae7f3e733b81fp-45;           /*  1/16! */
    p = fma(p, z, -0x1.93974a8c07c9dp-37);      /* -1/14! */
    p = fma(p, z, 0x1.1eed8eff8d898p-29);       /*  1/12! */
    p = fma(p, z, -0x1.27e4fb7789f5cp-22);      /* -1/10! */
    p = fma(p, z, 0x1.a01a01a01a01ap-16);       /*  1/8!  */
    p = fma(p, z, -0x1.6c16c16c16c17p-10);      /* -1/6!  */
    p = fma(p, z, 0x1.5555555555555p-5);        /*  1/4!  */
    p = fma(p, z, -0.5);
    return fma(z, p, 1.0);
}

DM_FN void dm_sincos(double x, double *s, double *c)
{
    if (!(fabs(x) < 0x1p20)) {   /* NaN, inf or outside the supported domain */
        *s = x - x; *c = x - x;
        if (fabs(x) < INFINITY) { *s = NAN; *c = NAN; }
        return;
    }
    double k = rint(x * DM_TWO_OVER_PI);
    double r = fma(-k, DM_PIO2_1, x);
    r = fma(-k, DM_PIO2_2, r);
    r = fma(-k, DM_PIO2_3, r);
    int q = (int)(int64_t)k & 3;
    double sr = dm__sin_kern(r), cr = dm__cos_kern(r);
    switch (q) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = -sr; break;
    case 2: *s = -sr; *c = -cr; break;
    default: *s = -cr; *c = sr; break;
    }
}

DM_FN double dm_sin(double x) { double s, c; dm_sincos(x, &s, &c); return s; }
DM_FN double dm_cos(double x) { double s, c; dm_sincos(x, &s, &c); return c; }

/* ---------------------------------------------------------------- atan2 / asin */
/* atan(t) for t in [0, 1]. */
DM_FN double dm__atan01(double t)
{
    double c = 0.0, ahi = 0.0, alo = 0.0, u;
    if (t >= 0.8206787908286604) {          /* tan(7 pi/32) */
        c = 1.0; ahi = 0x1.921fb54442d18p-1; alo = 0x1.1a62633145c07p-55;
    } else if (t >= 0.5345111359507917) {   /* tan(5 pi/32) */
        c = 0x1.561b82ab7f990p-1; ahi = 0x1.2d97c7f3321d2p-1; alo = -0x1.8f57cafebcf16p-58;
    } else if (t >= 0.3033466836073424) {   /* tan(3 pi/32) */
        c = 0x1.a827999fcef32p-2; ahi = 0x1.921fb54442d18p-2; alo = 0x1.c398861b78b55p-59;
    } else if (t >= 0.09849140335716425) {  /* tan(pi/32) */
        c = 0x1.975f5e0553158p-3; ahi = 0x1.921fb54442d18p-3; alo = 0x1.f93470dfef04ap-58;
    }
    if (c == 0.0) u = t;
    else u = (t - c) / fma(t, c, 1.0);
    double z = u * u;
    double p = 0x1.e1e1e1e1e1e1ep-5;          /*  1/17 */
    p = fma(p, z, -0x1.1111111111111p-4);      /* -1/15 */
    p = fma(p, z, 0x1.3b13b13b13b14p-4);       /*  1/13 */
    p = fma(p, z, -0x1.745d1745d1746p-4);      /* -1/11 */
    p = fma(p, z, 0x1.c71c71c71c71cp-4);       /*  1/9  */
    p = fma(p, z, -0x1.2492492492492p-3);      /* -1/7  */
    p = fma(p, z, 0x1.999999999999ap-3);       /*  1/5  */
    p = fma(p, z, -0x1.5555555555555p-2);      /* -1/3  */
    double r = fma(u * z, p, u);
    return ahi + (alo + r);
}

/* atan2(y, x) with the C99 special cases for zeros and infinities. */
DM_FN double dm_atan2(double y, double x)
{
    if (isnan(x) || isnan(y)) return x + y;
    double ax = fabs(x), ay = fabs(y), a;
    if (ay <= ax) {
        if (ax == 0.0) a = 0.0;
        else if (isinf(ax)) a = isinf(ay) ? 0x1.921fb54442d18p-1 : 0.0;
        else a = dm__atan01(ay / ax);
    } else {
        if (isinf(ay)) a = DM_PIO2_1;
        else a = (DM_PIO2_1 - dm__atan01(ax / ay)) + DM_PIO2_2;
    }
    if (signbit(x)) a = (DM_PI_1 - a) + DM_PI_2;
    return copysign(a, y);
}

DM_FN double dm_asin(double x)
{
    if (!(fabs(x) <= 1.0)) return NAN;
    return dm_atan2(x, sqrt((1.0 - x) * (1.0 + x)));
}

/* ---------------------------------------------------------------- exp / log */
DM_FN double dm_exp(double x)
{
    if (isnan(x)) return x;
    if (x > 709.782712893384) return INFINITY;
    if (x < -745.1332191019412) return 0.0;
    double k = rint(x * DM_INV_LN2);
    double r = fma(-k, DM_LN2_1, x);
    r = fma(-k, DM_LN2_2, r);
    double p = 0x1.6124613a86d09p-33;           /* 1/13! */
    p = fma(p, r, 0x1.1eed8eff8d898p-29);       /* 1/12! */
    p = fma(p, r, 0x1.ae64567f544e4p-26);       /* 1/11! */
    p = fma(p, r, 0x1.27e4fb7789f5cp-22);       /* 1/10! */
    p = fma(p, r, 0x1.71de3a556c734p-19);       /* 1/9!  */
    p = fma(p, r, 0x1.a01a01a01a01ap-16);       /* 1/8!  */
    p = fma(p, r, 0x1.a01a01a01a01ap-13);       /* 1/7!  */
    p = fma(p, r, 0x1.6c16c16c16c17p-10);       /* 1/6!  */
    p = fma(p, r, 0x1.1111111111111p-7);        /* 1/5!  */
    p = fma(p, r, 0x1.5555555555555p-5);        /* 1/4!  */
    p = fma(p, r, 0x1.5555555555555p-3);        /* 1/3!  */
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.0);
    p = fma(p, r, 1.0);
    return dm_ldexp(p, (int)k);
}

DM_FN double dm_log(double x)
{
    if (isnan(x)) return x;
    if (x < 0.0) return NAN;
    if (x == 0.0) return -INFINITY;
    if (isinf(x)) return x;
    int e = 0;
    if (x < 0x1p-1022) { x *= 0x1p54; e = -54; }
    uint64_t b = dm_as_u64(x);
    e += (int)((b >> 52) & 0x7ff) - 1023;
    double m = dm_from_u64((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (m > DM_SQRT2) { m *= 0.5; e += 1; }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double p = 0x1.8618618618618p-4;            /* 2/21 */
    p = fma(p, z, 0x1.af286bca1af28p-4);        /* 2/19 */
    p = fma(p, z, 0x1.e1e1e1e1e1e1ep-4);        /* 2/17 */
    p = fma(p, z, 0x1.1111111111111p-3);        /* 2/15 */
    p = fma(p, z, 0x1.3b13b13b13b14p-3);        /* 2/13 */
    p = fma(p, z, 0x1.745d1745d1746p-3);        /* 2/11 */
    p = fma(p, z, 0x1.c71c71c71c71cp-3);        /* 2/9  */
    p = fma(p, z, 0x1.2492492492492p-2);        /* 2/7  */
    p = fma(p, z, 0x1.999999999999ap-2);        /* 2/5  */
    p = fma(p, z, 0x1.5555555555555p-1);        /* 2/3  */
    double R = z * p;
    double logm = fma(-s, f - R, f);
    double de = (double)e;
    return fma(de, DM_LN2_HI32, fma(de, DM_LN2_LO32, logm));
}

/* ---------------------------------------------------------------- log10 / 10^y / x^(1/4) */
/* log10 x = log x / ln 10; 10^y = exp(y ln 10) (y ln 10 rounded once: a relative
 * error of about |y| ln 10 ulp, below 1e-14 for the vapour pressures of the
 * thermal model); x^(1/4) = sqrt(sqrt(x)). */
#define DM_LN10 0x1.26bb1bbb55516p+1
DM_FN double dm_log10(double x) { return dm_log(x) / DM_LN10; }
DM_FN double dm_pow10(double y) { return dm_exp(y * DM_LN10); }
DM_FN double dm_root4(double x) { return sqrt(sqrt(x)); }

/* ---------------------------------------------------------------- a / b and sqrt(x), in range */
/* The correctly rounded FP64 division and square root the compiler emits for
 * gfx950 wrap the same reciprocal / reciprocal-square-root refinement in range
 * scaling (v_div_scale, ldexp by 2^256 below 2^-767) and special-case fix-up
 * (v_div_fixup, the class test for 0 and infinity).  For positive normal
 * operands well inside the range - a quotient and operands between 2^-500 and
 * 2^500 - the scaling is by 2^0 and the fix-up returns its input, so these
 * refinements alone give the same bits with fewer dependent instructions.  On the
 * host: the IEEE operations.  For the push: GM / r^3 and |s|^2. */
DM_FN double dm_div_pos(double a, double b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-b, r, 1.0);
    r = __builtin_fma(r, e, r);
    const double q = a * r;
    const double rem = __builtin_fma(-b, q, a);
    return __builtin_fma(rem, r, q);
#else
    return a / b;
#endif
}

DM_FN double dm_sqrt_pos(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = y * 0.5;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
#else
    return sqrt(x);
#endif
}

#endif /* DSMC_MATH_H */
