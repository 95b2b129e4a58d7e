/* kmath.h -- double-precision elementary functions of the tracking kernels.
 *
 * One source, compiled for the gfx950 device by hipcc and for the host by
 * gcc, so that a host build of the same expressions reproduces the device
 * results bit for bit.  Every routine is built from IEEE-754 binary64
 * add / sub / mul / div / sqrt only (correctly rounded on both targets),
 * with floating-point contraction disabled by the build (-ffp-contract=off):
 * no fused multiply-add, no platform libm, no lookup of vendor math code.
 *
 * Methods (textbook, see J.-M. Muller, "Elementary Functions", 3rd ed.):
 *   sin/cos  Cody-Waite reduction by pi/2 split in three parts (33+33+53
 *            bits), remainder carried as a double-double; Taylor series of
 *            sin/cos on |r| <= pi/4 (truncation < 2^-62 relative); cos uses
 *            Dekker's exact square and a compensated 1 - r^2/2.
 *   atan     |x| < 3/16: Taylor series; 3/16 <= |x| <= 2: table of
 *            atan(j/8) (hi+lo) and atan(x) = atan(c) + atan((x-c)/(1+xc)),
 *            |t| <= 1/16; |x| > 2: pi/2 - atan(1/x).
 *   atan2    C99 special cases, quadrant fix with pi as hi+lo.
 *   asin/acos  through atan2 and sqrt((1-x)(1+x)).
 *   log      x = 2^k m, m in [sqrt(1/2), sqrt(2)), log(1+f) = 2 atanh(s),
 *            s = f/(2+f), written as f - (s f - s^3 P(s^2)).
 *   exp      x = k ln2 + r (ln2 split 42+53 bits), Taylor series on
 *            |r| <= ln2/2, scaled by 2^k.
 * Coefficients are the round-to-nearest binary64 values of the exact
 * rationals (-1)^k/(2k+1)!, 1/k!, 2/(2k+1); the tables of atan(j/8), pi/2
 * and ln2 come from kelp_amd/csrc/gen_kmath_tables.py (200-bit mpmath).
 * Accuracy against glibc 2.35 is measured in tests/test_kmath.py.
 *
 * Valid as C99 (the oracle's diagnostic build) and as HIP C++.
 */
#ifndef KELP_KMATH_H
#define KELP_KMATH_H

#include <stdint.h>

#if defined(__HIP__)
#define KM_FN __host__ __device__ static inline
#else
#define KM_FN static inline
#endif

#define KM_2_OVER_PI 0x1.45f306dc9c883p-1
#define KM_PIO2_1 0x1.921fb54400000p+0 /* 33 bits of pi/2 */
#define KM_PIO2_2 0x1.0b4611a600000p-34 /* next 33 bits */
#define KM_PIO2_3 0x1.3198a2e037073p-69 /* the rest, rounded */
#define KM_PIO2_HI 0x1.921fb54442d18p+0
#define KM_PIO2_LO 0x1.1a62633145c07p-54
#define KM_PI_HI 0x1.921fb54442d18p+1
#define KM_PI_LO 0x1.1a62633145c07p-53
#define KM_LN2_HI 0x1.62e42fefa3800p-1 /* 42 bits: k*LN2_HI exact, |k|<2^11 */
#define KM_LN2_LO 0x1.ef35793c76730p-45
#define KM_INV_LN2 0x1.71547652b82fep+0
#define KM_SQRT2 0x1.6a09e667f3bcdp+0
#define KM_RND_BIG 0x1.8p52 /* (v + BIG) - BIG rounds |v| < 2^51 to an integer */

KM_FN double km_fabs(double x) { return __builtin_fabs(x); }
KM_FN double km_copysign(double x, double s) { return __builtin_copysign(x, s); }
KM_FN double km_sqrt(double x) { return __builtin_sqrt(x); }

KM_FN uint64_t km_bits(double x)
{
    uint64_t u;
    __builtin_memcpy(&u, &x, sizeof u);
    return u;
}

KM_FN double km_from_bits(uint64_t u)
{
    double x;
    __builtin_memcpy(&x, &u, sizeof x);
    return x;
}

KM_FN int km_isnan(double x) { return x != x; }

/* Nearest integer (ties to even) of |v| < 2^51, exact. */
KM_FN double km_rint_small(double v)
{
    double t = v + KM_RND_BIG;
    return t - KM_RND_BIG;
}

/* Python float `%` for a positive divisor d (CPython float_rem): the exact
 * remainder fmod(x, d) (exactly representable, so every correct fmod returns
 * the same bits), then a rounded +d when it is negative, +0.0 when zero. */
KM_FN double km_pymod_pos(double x, double d)
{
    double m = __builtin_fmod(x, d);
    if (m != 0.0) {
        if (m < 0.0) m += d;
    } else {
        m = 0.0;
    }
    return m;
}

/* numerics/diffusion.py:334 reflect_depth: specular fold of z into
 * [top, bot]; a degenerate column collapses to top. */
KM_FN double km_reflect_depth(double z, double top, double bot)
{
    double L = bot - top;
    if (L <= 0.0) return top;
    double y = km_pymod_pos(z - top, 2.0 * L);
    if (y < 0.0) y += 2.0 * L;
    if (y > L) y = 2.0 * L - y;
    return top + y;
}

/* Exact x mod 360 with the sign of the divisor -- the remainder Python's
 * float `%` (and numpy's remainder) return: fmod, then +360 when the
 * remainder is negative, +0.0 when it is zero.  Exact for |x| < 2^50. */
KM_FN double km_pymod360(double x)
{
    if (!(km_fabs(x) < 0x1p50))
        return x - x; /* NaN for NaN/inf, like fmod */
    double q = (double)(int64_t)(x / 360.0);  /* trunc; may be off by one */
    double r = x - q * 360.0;                 /* exact: q*360 exact, Sterbenz */
    if (x >= 0.0) {
        if (r >= 360.0) r -= 360.0;
        else if (r < 0.0) r += 360.0;
    } else {
        if (r <= -360.0) r += 360.0;
        else if (r > 0.0) r -= 360.0;
    }
    /* r is now fmod(x, 360) exactly; apply the divisor-sign rule */
    if (r != 0.0) {
        if (r < 0.0) r += 360.0;
    } else {
        r = 0.0;
    }
    return r;
}

/* ---------------------------------------------------------------- sin/cos */

/* r = x - n*pi/2 as rh + rl; returns n.  |x| < 1.5e6. */
KM_FN int km_reduce_pio2(double x, double *rh, double *rl)
{
    double fn = km_rint_small(x * KM_2_OVER_PI);
    double r1 = x - fn * KM_PIO2_1; /* exact */
    double w = fn * KM_PIO2_2;      /* exact */
    double t = r1 - w;              /* two-sum of r1 and -w */
    double bb = t - r1;
    double e = (r1 - (t - bb)) + (-w - bb);
    e = e - fn * KM_PIO2_3;
    double hsum = t + e;
    *rl = e - (hsum - t);
    *rh = hsum;
    return (int)fn;
}

KM_FN double km_sin_kernel(double rh, double rl)
{
    double z = rh * rh;
    double p = -0x1.5555555555555p-3 +
               z * (0x1.1111111111111p-7 +
               z * (-0x1.a01a01a01a01ap-13 +
               z * (0x1.71de3a556c734p-19 +
               z * (-0x1.ae64567f544e4p-26 +
               z * (0x1.6124613a86d09p-33 +
               z * (-0x1.ae7f3e733b81fp-41 +
               z * 0x1.952c77030ad4ap-49))))));
    return rh + ((rh * z) * p + rl * (1.0 - 0.5 * z));
}

/* Dekker's exact product error a*b - fl(a*b), no fma. */
KM_FN double km_two_prod_err(double a, double b, double p)
{
    const double split = 134217729.0; /* 2^27 + 1 */
    double ca = split * a, cb = split * b;
    double ah = ca - (ca - a), al = a - ah;
    double bh = cb - (cb - b), bl = b - bh;
    return ((ah * bh - p) + ah * bl + al * bh) + al * bl;
}

KM_FN double km_cos_kernel(double rh, double rl)
{
    double zh = rh * rh;
    double zl = km_two_prod_err(rh, rh, zh);
    double q = 0x1.5555555555555p-5 +
               zh * (-0x1.6c16c16c16c17p-10 +
               zh * (0x1.a01a01a01a01ap-16 +
               zh * (-0x1.27e4fb7789f5cp-22 +
               zh * (0x1.1eed8eff8d898p-29 +
               zh * (-0x1.93974a8c07c9dp-37 +
               zh * (0x1.ae7f3e733b81fp-45 +
               zh * -0x1.6827863b97d97p-53))))));
    double hz = 0.5 * zh;
    double w = 1.0 - hz;
    double e = (1.0 - w) - hz; /* rounding error of w, exact */
    return w + (e + ((zh * zh) * q - (0.5 * zl + rh * rl)));
}

/* Arguments with |x| >= KM_TRIG_MAX lose the exactness of the reduction;
 * the tracking code never forms them (radians of lon/lat, 2*pi*u). */
#define KM_TRIG_MAX 1.5e6

KM_FN double km_sin(double x)
{
    if (!(km_fabs(x) < KM_TRIG_MAX)) {
        if (km_isnan(x) || km_fabs(x) == __builtin_inf()) return x - x;
    }
    double rh, rl;
    int n = km_reduce_pio2(x, &rh, &rl);
    switch (n & 3) {
    case 0: return km_sin_kernel(rh, rl);
    case 1: return km_cos_kernel(rh, rl);
    case 2: return -km_sin_kernel(rh, rl);
    default: return -km_cos_kernel(rh, rl);
    }
}

KM_FN double km_cos(double x)
{
    if (!(km_fabs(x) < KM_TRIG_MAX)) {
        if (km_isnan(x) || km_fabs(x) == __builtin_inf()) return x - x;
    }
    double rh, rl;
    int n = km_reduce_pio2(x, &rh, &rl);
    switch (n & 3) {
    case 0: return km_cos_kernel(rh, rl);
    case 1: return -km_sin_kernel(rh, rl);
    case 2: return -km_cos_kernel(rh, rl);
    default: return km_sin_kernel(rh, rl);
    }
}

/* sin and cos of one argument, one reduction; bit-identical to km_sin and
 * km_cos called separately. */
KM_FN void km_sincos(double x, double *s, double *c)
{
    if (!(km_fabs(x) < KM_TRIG_MAX)) {
        if (km_isnan(x) || km_fabs(x) == __builtin_inf()) {
            *s = x - x;
            *c = x - x;
            return;
        }
    }
    double rh, rl;
    int n = km_reduce_pio2(x, &rh, &rl);
    double sk = km_sin_kernel(rh, rl), ck = km_cos_kernel(rh, rl);
    switch (n & 3) {
    case 0: *s = sk; *c = ck; break;
    case 1: *s = ck; *c = -sk; break;
    case 2: *s = -sk; *c = -ck; break;
    default: *s = -ck; *c = sk; break;
    }
}

/* ------------------------------------------------------------------ atan */

KM_FN double km_atan_series_small(double x) /* |x| < 3/16 */
{
    double z = x * x;
    double p = -0x1.5555555555555p-2 +
               z * (0x1.999999999999ap-3 +
               z * (-0x1.2492492492492p-3 +
               z * (0x1.c71c71c71c71cp-4 +
               z * (-0x1.745d1745d1746p-4 +
               z * (0x1.3b13b13b13b14p-4 +
               z * (-0x1.1111111111111p-4 +
               z * (0x1.e1e1e1e1e1e1ep-5 +
               z * (-0x1.af286bca1af28p-5 +
               z * (0x1.8618618618618p-5 +
               z * (-0x1.642c8590b2164p-5 +
               z * (0x1.47ae147ae147bp-5 +
               z * -0x1.2f684bda12f68p-5)))))))))));
    return x + (x * z) * p;
}

KM_FN double km_atan_series_tiny(double t) /* |t| <= 1/16 */
{
    double z = t * t;
    double p = -0x1.5555555555555p-2 +
               z * (0x1.999999999999ap-3 +
               z * (-0x1.2492492492492p-3 +
               z * (0x1.c71c71c71c71cp-4 +
               z * (-0x1.745d1745d1746p-4 +
               z * (0x1.3b13b13b13b14p-4 +
               z * (-0x1.1111111111111p-4 +
               z * 0x1.e1e1e1e1e1e1ep-5))))));
    return t + (t * z) * p;
}

/* atan(a) for 0 <= a <= 2 */
KM_FN double km_atan_0_2(double a)
{
    if (a < 0.1875) return km_atan_series_small(a);
    int j = (int)(a * 8.0 + 0.5); /* 2..16 */
    double c = (double)j * 0.125;
    double t = (a - c) / (1.0 + a * c);
    double hi, lo;
    switch (j) {
    case 2: hi = 0x1.f5b75f92c80ddp-3; lo = 0x1.8ab6e3cf7afbdp-57; break;
    case 3: hi = 0x1.6f61941e4def1p-2; lo = -0x1.c63aae6f6e918p-56; break;
    case 4: hi = 0x1.dac670561bb4fp-2; lo = 0x1.a2b7f222f65e2p-56; break;
    case 5: hi = 0x1.1e00babdefeb4p-1; lo = -0x1.928df287a668fp-58; break;
    case 6: hi = 0x1.4978fa3269ee1p-1; lo = 0x1.2419a87f2a458p-56; break;
    case 7: hi = 0x1.700a7c5784634p-1; lo = -0x1.8c34d25aadef6p-56; break;
    case 8: hi = 0x1.921fb54442d18p-1; lo = 0x1.1a62633145c07p-55; break;
    case 9: hi = 0x1.b034f38649c88p-1; lo = -0x1.be88d6936f833p-55; break;
    case 10: hi = 0x1.cac7c57846f9ep-1; lo = 0x1.0dae13ad18a6bp-55; break;
    case 11: hi = 0x1.e24dd44c855d1p-1; lo = 0x1.f7ac612ab33d8p-55; break;
    case 12: hi = 0x1.f730bd281f69bp-1; lo = 0x1.007887af0cbbdp-56; break;
    case 13: hi = 0x1.04e67277a01d7p+0; lo = 0x1.7115496c13eb6p-57; break;
    case 14: hi = 0x1.0d38f2c5ba09fp+0; lo = -0x1.bd0dc231bfd70p-54; break;
    case 15: hi = 0x1.14b1dd5f90ce1p+0; lo = -0x1.212d570a63fa2p-56; break;
    default: hi = 0x1.1b6e192ebbe44p+0; lo = 0x1.b1b466a88828ep-54; break;
    }
    return hi + (lo + km_atan_series_tiny(t));
}

KM_FN double km_atan(double x)
{
    if (km_isnan(x)) return x;
    double a = km_fabs(x);
    double r;
    if (a <= 2.0) {
        r = km_atan_0_2(a);
    } else if (a < 0x1p60) {
        r = KM_PIO2_HI - (km_atan_0_2(1.0 / a) - KM_PIO2_LO);
    } else {
        r = KM_PIO2_HI;
    }
    return km_copysign(r, x);
}

KM_FN double km_atan2(double y, double x)
{
    if (km_isnan(x) || km_isnan(y)) return x + y;
    double ax = km_fabs(x), ay = km_fabs(y);
    const double inf = __builtin_inf();
    if (ay == 0.0) {
        if (__builtin_signbit(x)) return km_copysign(KM_PI_HI, y);
        return y; /* +-0 */
    }
    if (ax == 0.0) return km_copysign(KM_PIO2_HI, y);
    if (ax == inf) {
        if (ay == inf)
            return km_copysign(__builtin_signbit(x) ? 3.0 * 0x1.921fb54442d18p-1
                                                      : 0x1.921fb54442d18p-1, y);
        return __builtin_signbit(x) ? km_copysign(KM_PI_HI, y) : km_copysign(0.0, y);
    }
    if (ay == inf) return km_copysign(KM_PIO2_HI, y);
    double q = ay / ax;
    double a = km_atan(q); /* q may be inf (-> pi/2) or 0 */
    if (__builtin_signbit(x)) a = KM_PI_HI - (a - KM_PI_LO);
    return km_copysign(a, y);
}

KM_FN double km_asin(double x)
{
    double a = km_fabs(x);
    if (!(a <= 1.0)) return (x - x) / (x - x); /* NaN for |x| > 1 or NaN */
    if (a < 0x1p-27) return x;
    return km_atan2(x, km_sqrt((1.0 - x) * (1.0 + x)));
}

KM_FN double km_acos(double x)
{
    double a = km_fabs(x);
    if (!(a <= 1.0)) return (x - x) / (x - x);
    return km_atan2(km_sqrt((1.0 - x) * (1.0 + x)), x);
}

/* ------------------------------------------------------------------- log */

KM_FN double km_log(double x)
{
    if (km_isnan(x)) return x;
    if (x < 0.0) return (x - x) / (x - x);
    if (x == 0.0) return -__builtin_inf();
    if (x == __builtin_inf()) return x;
    int k = 0;
    if (x < 0x1p-1022) { /* subnormal */
        x *= 0x1p54;
        k = -54;
    }
    uint64_t u = km_bits(x);
    k += (int)((u >> 52) & 0x7ff) - 1023;
    double m = km_from_bits((u & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    if (m > KM_SQRT2) {
        m *= 0.5;
        k += 1;
    }
    double f = m - 1.0; /* exact */
    double s = f / (2.0 + f);
    double z = s * s;
    double p = 0x1.5555555555555p-1 +
               z * (0x1.999999999999ap-2 +
               z * (0x1.2492492492492p-2 +
               z * (0x1.c71c71c71c71cp-3 +
               z * (0x1.745d1745d1746p-3 +
               z * (0x1.3b13b13b13b14p-3 +
               z * (0x1.1111111111111p-3 +
               z * (0x1.e1e1e1e1e1e1ep-4 +
               z * (0x1.af286bca1af28p-4 +
               z * (0x1.8618618618618p-4 +
               z * 0x1.642c8590b2164p-4)))))))));
    double lm = f - (s * f - (s * z) * p);
    double dk = (double)k;
    return dk * KM_LN2_HI + (lm + dk * KM_LN2_LO);
}

/* ------------------------------------------------------------------- exp */

KM_FN double km_exp(double x)
{
    if (km_isnan(x)) return x;
    if (x > 709.782712893384) return __builtin_inf();
    if (x < -745.2) return 0.0;
    double fk = km_rint_small(x * KM_INV_LN2);
    double rh = x - fk * KM_LN2_HI; /* exact */
    double rl = -(fk * KM_LN2_LO);
    double r = rh + rl;
    rl = rl - (r - rh);
    double q = 0x1.0000000000000p-1 +
               r * (0x1.5555555555555p-3 +
               r * (0x1.5555555555555p-5 +
               r * (0x1.1111111111111p-7 +
               r * (0x1.6c16c16c16c17p-10 +
               r * (0x1.a01a01a01a01ap-13 +
               r * (0x1.a01a01a01a01ap-16 +
               r * (0x1.71de3a556c734p-19 +
               r * (0x1.27e4fb7789f5cp-22 +
               r * (0x1.ae64567f544e4p-26 +
               r * (0x1.1eed8eff8d898p-29 +
               r * (0x1.6124613a86d09p-33 +
               r * (0x1.93974a8c07c9dp-37 +
               r * 0x1.ae7f3e733b81fp-41))))))))))));
    double e = 1.0 + (r + ((r * r) * q + rl * (1.0 + r)));
    int k = (int)fk;
    if (k > 1023) {
        e *= 0x1p1023;
        k -= 1023;
    } else if (k < -1021) {
        e *= 0x1p-600;
        k += 600;
    }
    return e * km_from_bits((uint64_t)(k + 1023) << 52);
}

#endif /* KELP_KMATH_H */
