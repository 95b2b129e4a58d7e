/* kmath.h -- double-precision elementary functions of the tracking kernels.
 *
 * One source, compiled for the gfx950 device by hipcc and for the host by
 * gcc, so that a host build of the same expressions reproduces the device
 * results bit for bit.  Every routine is built from IEEE-754 binary64
 * add / sub / mul / div / sqrt only (correctly rounded on both targets),
 * with floating-point contraction disabled by the build (-ffp-contract=off):
 * no fused multiply-add (exact products by Dekker's splitting), no platform
 * libm, no lookup of vendor math code.
 *
 * Methods (textbook, see J.-M. Muller, "Elementary Functions", 3rd ed.):
 *   sin/cos  Cody-Waite reduction by pi/2 split in three parts (33+33+53
 *            bits), remainder carried as a double-double; then a table of
 *            sin/cos(j/64) as double-doubles and the addition formulas with
 *            the Taylor series of sin/cos on |d| <= 1/128, the leading
 *            terms in double-double arithmetic (Dekker's exact product):
 *            relative error before the final rounding below 2^-80, so the
 *            result is the correctly rounded sine / cosine except when the
 *            exact value lies within that distance of a rounding boundary.
 *   atan2    C99 special cases; the ratio of the smaller to the larger
 *            argument as a double-double, j = rint(64 t), atan(j/64) as a
 *            double-double from the table and atan u of the reduced
 *            u = (t - j/64) / (1 + t j/64), |u| <= 1/128, by its series
 *            with the leading terms in double-double; pi/2 - and pi -
 *            in double-double: relative error before the final rounding
 *            below 2^-84.  atan(x) = atan2(x, 1).
 *   asin/acos  atan2 of x and sqrt(1 - x^2), the root formed as a
 *            double-double from (1 - x)(1 + x), one rounding at the end.
 *   log      x = 2^k m, m in [sqrt(1/2), sqrt(2)), c_i = RN(1/(1 + i/128))
 *            for the nearest i, r = m c_i - 1 exact as a double-double
 *            (|r| <= 2^-7.5), log x = k ln2 - log(c_i) + log1p(r) with
 *            -log(c_i) tabulated as a double-double and log1p's terms up
 *            to r^3 in double-double, the rest in double: relative error
 *            before the final rounding below 2^-76 (near x = 1, where
 *            c_i = 1, relative to log1p(r) itself).
 *   exp      x = (64 k + j) ln2/64 + r (ln2/64 split 36+36+53 bits),
 *            2^(j/64) as a double-double from the table, exp(r) - 1 by its
 *            series with the terms to r^3 in double-double: relative
 *            error before the final rounding below 2^-85.
 *   pow      x > 0: exp of y log x with log x and the product in
 *            double-double.
 * Each is therefore the correctly rounded value except when the exact
 * value lies within its error bound of a rounding boundary.
 * Coefficients are the round-to-nearest binary64 values of the exact
 * rationals (-1)^k/(2k+1)!, 1/k!, (-1)^k/(2k+1), 1/k; the tables
 * (atan(j/64), 2^(j/64), log(c_i), sin/cos(j/64)) and the constants pi/2,
 * ln2 and ln2/64 come from kelp_amd/csrc/gen_kmath_tables.py (200-bit
 * mpmath).
 * Accuracy against glibc is measured in tests/test_oracle_pins.py.
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

/* Dekker's exact product error a*b - fl(a*b), no fma. */
KM_FN double km_two_prod_err(double a, double b, double p)
{
    const double split = 134217729.0; /* 2^27 + 1 */
    double ca = split * a, cb = split * b;
    double ah = ca - (ca - a), al = a - ah;
    double bh = cb - (cb - b), bl = b - bh;
    return ((ah * bh - p) + ah * bl + al * bh) + al * bl;
}

/* ------------------------------------------------------ double-double */

KM_FN void km_two_sum(double a, double b, double *s, double *e)
{
    double t = a + b;
    double bb = t - a;
    *e = (a - (t - bb)) + (b - bb);
    *s = t;
}

KM_FN void km_fast_two_sum(double a, double b, double *s, double *e) /* |a| >= |b| or a == 0 */
{
    double t = a + b;
    *e = b - (t - a);
    *s = t;
}

/* (ah + al)(bh + bl) as a normalised double-double (relative error < 2^-100) */
KM_FN void km_dd_mul(double ah, double al, double bh, double bl, double *ph, double *pl)
{
    double p = ah * bh;
    double e = km_two_prod_err(ah, bh, p) + (ah * bl + al * bh);
    km_fast_two_sum(p, e, ph, pl);
}

KM_FN void km_dd_add(double ah, double al, double bh, double bl, double *sh, double *sl)
{
    double s, e;
    km_two_sum(ah, bh, &s, &e);
    e += al + bl;
    km_fast_two_sum(s, e, sh, sl);
}

static const double KM_LOGT_C[91] = {
    0x1.6816816816817p+0, 0x1.642c8590b2164p+0, 0x1.6058160581606p+0,
    0x1.5c9882b931057p+0, 0x1.58ed2308158edp+0, 0x1.5555555555555p+0,
    0x1.51d07eae2f815p+0, 0x1.4e5e0a72f0539p+0, 0x1.4afd6a052bf5bp+0,
    0x1.47ae147ae147bp+0, 0x1.446f86562d9fbp+0, 0x1.4141414141414p+0,
    0x1.3e22cbce4a902p+0, 0x1.3b13b13b13b14p+0, 0x1.3813813813814p+0,
    0x1.3521cfb2b78c1p+0, 0x1.323e34a2b10bfp+0, 0x1.2f684bda12f68p+0,
    0x1.2c9fb4d812ca0p+0, 0x1.29e4129e4129ep+0, 0x1.27350b8812735p+0,
    0x1.2492492492492p+0, 0x1.21fb78121fb78p+0, 0x1.1f7047dc11f70p+0,
    0x1.1cf06ada2811dp+0, 0x1.1a7b9611a7b96p+0, 0x1.1811811811812p+0,
    0x1.15b1e5f75270dp+0, 0x1.135c81135c811p+0, 0x1.1111111111111p+0,
    0x1.0ecf56be69c90p+0, 0x1.0c9714fbcda3bp+0, 0x1.0a6810a6810a7p+0,
    0x1.0842108421084p+0, 0x1.0624dd2f1a9fcp+0, 0x1.0410410410410p+0,
    0x1.0204081020408p+0, 0x1.0000000000000p+0, 0x1.fc07f01fc07f0p-1,
    0x1.f81f81f81f820p-1, 0x1.f44659e4a4271p-1, 0x1.f07c1f07c1f08p-1,
    0x1.ecc07b301ecc0p-1, 0x1.e9131abf0b767p-1, 0x1.e573ac901e574p-1,
    0x1.e1e1e1e1e1e1ep-1, 0x1.de5d6e3f8868ap-1, 0x1.dae6076b981dbp-1,
    0x1.d77b654b82c34p-1, 0x1.d41d41d41d41dp-1, 0x1.d0cb58f6ec074p-1,
    0x1.cd85689039b0bp-1, 0x1.ca4b3055ee191p-1, 0x1.c71c71c71c71cp-1,
    0x1.c3f8f01c3f8f0p-1, 0x1.c0e070381c0e0p-1, 0x1.bdd2b899406f7p-1,
    0x1.bacf914c1bad0p-1, 0x1.b7d6c3dda338bp-1, 0x1.b4e81b4e81b4fp-1,
    0x1.b2036406c80d9p-1, 0x1.af286bca1af28p-1, 0x1.ac5701ac5701bp-1,
    0x1.a98ef606a63bep-1, 0x1.a6d01a6d01a6dp-1, 0x1.a41a41a41a41ap-1,
    0x1.a16d3f97a4b02p-1, 0x1.9ec8e951033d9p-1, 0x1.9c2d14ee4a102p-1,
    0x1.999999999999ap-1, 0x1.970e4f80cb872p-1, 0x1.948b0fcd6e9e0p-1,
    0x1.920fb49d0e229p-1, 0x1.8f9c18f9c18fap-1, 0x1.8d3018d3018d3p-1,
    0x1.8acb90f6bf3aap-1, 0x1.886e5f0abb04ap-1, 0x1.8618618618618p-1,
    0x1.83c977ab2beddp-1, 0x1.8181818181818p-1, 0x1.7f405fd017f40p-1,
    0x1.7d05f417d05f4p-1, 0x1.7ad2208e0ecc3p-1, 0x1.78a4c8178a4c8p-1,
    0x1.767dce434a9b1p-1, 0x1.745d1745d1746p-1, 0x1.724287f46debcp-1,
    0x1.702e05c0b8170p-1, 0x1.6e1f76b4337c7p-1, 0x1.6c16c16c16c17p-1,
    0x1.6a13cd1537290p-1
};
static const double KM_LOGT_H[91] = {
    -0x1.5d5bddf595f31p-2, -0x1.522ae0738a3d7p-2, -0x1.4718dc271c41cp-2,
    -0x1.3c25277333183p-2, -0x1.314f1e1d35ce3p-2, -0x1.269621134db91p-2,
    -0x1.1bf99635a6b95p-2, -0x1.1178e8227e47ap-2, -0x1.07138604d5864p-2,
    -0x1.f991c6cb3b37ap-3, -0x1.e530effe71013p-3, -0x1.d1037f2655e7bp-3,
    -0x1.bd087383bd8aap-3, -0x1.a93ed3c8ad9e5p-3, -0x1.95a5adcf70182p-3,
    -0x1.823c16551a3c0p-3, -0x1.6f0128b756ab9p-3, -0x1.5bf406b543db0p-3,
    -0x1.4913d8333b563p-3, -0x1.365fcb0159014p-3, -0x1.23d712a49c201p-3,
    -0x1.1178e8227e47ap-3, -0x1.fe89139dbd565p-4, -0x1.da7276384469ep-4,
    -0x1.b6ac88dad5b1dp-4, -0x1.9335e5d594988p-4, -0x1.700d30aeac0e8p-4,
    -0x1.4d3115d207eacp-4, -0x1.2aa04a44717a1p-4, -0x1.08598b59e3a06p-4,
    -0x1.ccb73cdddb2d0p-5, -0x1.894aa149fb34bp-5, -0x1.466aed42de3f9p-5,
    -0x1.0415d89e74440p-5, -0x1.8492528c8cac5p-6, -0x1.0205658935837p-6,
    -0x1.010157588de69p-7, 0x0.0p+0, 0x1.fe02a6b106799p-8,
    0x1.fc0a8b0fc03c4p-7, 0x1.7b91b07d5b126p-6, 0x1.f829b0e7832f8p-6,
    0x1.39e87b9febd68p-5, 0x1.77458f632dcffp-5, 0x1.b42dd711971b9p-5,
    0x1.f0a30c01162a8p-5, 0x1.16536eea37ae3p-4, 0x1.341d7961bd1d0p-4,
    0x1.51b073f06183cp-4, 0x1.6f0d28ae56b4ep-4, 0x1.8c345d6319b23p-4,
    0x1.a926d3a4ad562p-4, 0x1.c5e548f5bc743p-4, 0x1.e27076e2af2eap-4,
    0x1.fec9131dbeabcp-4, 0x1.0d77e7cd08e5bp-3, 0x1.1b72ad52f67a2p-3,
    0x1.29552f81ff521p-3, 0x1.371fc201e8f75p-3, 0x1.44d2b6ccb7d1cp-3,
    0x1.526e5e3a1b438p-3, 0x1.5ff3070a793d6p-3, 0x1.6d60fe719d21bp-3,
    0x1.7ab890210d907p-3, 0x1.87fa06520c911p-3, 0x1.9525a9cf456b6p-3,
    0x1.a23bc1fe2b561p-3, 0x1.af3c94e80bff3p-3, 0x1.bc286742d8cd4p-3,
    0x1.c8ff7c79a9a20p-3, 0x1.d5c216b4fbb94p-3, 0x1.e27076e2af2e8p-3,
    0x1.ef0adcbdc5935p-3, 0x1.fb9186d5e3e29p-3, 0x1.0402594b4d041p-2,
    0x1.0a324e27390e2p-2, 0x1.1058bf9ae4ad4p-2, 0x1.1675cababa60fp-2,
    0x1.1c898c16999fbp-2, 0x1.22941fbcf7966p-2, 0x1.2895a13de86a4p-2,
    0x1.2e8e2bae11d31p-2, 0x1.347dd9a987d56p-2, 0x1.3a64c556945eap-2,
    0x1.404308686a7e4p-2, 0x1.4618bc21c5ec2p-2, 0x1.4be5f957778a1p-2,
    0x1.51aad872df82ep-2, 0x1.5767717455a6cp-2, 0x1.5d1bdbf5809cap-2,
    0x1.62c82f2b9c796p-2
};
static const double KM_LOGT_L[91] = {
    -0x1.d5f75b9a23ae4p-59, -0x1.3840b263acb43p-56, -0x1.d8fb4c14c56eep-56,
    -0x1.152d81af5713ap-56, -0x1.22966f61a3c23p-56, -0x1.e0efadd9db02ap-56,
    0x1.e9575c2124912p-56, -0x1.b8ce2d07f1cb7p-56, 0x1.24e912b16ec8bp-60,
    -0x1.ecca0cdf30143p-58, 0x1.f7627ef82f3f0p-57, 0x1.3f3adb7b71cbcp-58,
    0x1.1165504ad749ep-59, -0x1.bcafa9de97202p-57, -0x1.8a16283fdbd1cp-57,
    -0x1.6dcd318f4187ep-57, 0x1.37967087859b9p-59, 0x1.1f5b44c0df7f7p-61,
    0x1.0d5604930f137p-58, -0x1.bea08d2dca256p-57, -0x1.51c7e9efae297p-57,
    0x1.0e63a5f01c693p-58, 0x1.ac9f4215f9394p-58, -0x1.401fa71733017p-58,
    0x1.002bf768e52d0p-58, 0x1.478a85704ccb7p-58, -0x1.a36a677b4c8b2p-59,
    -0x1.da7d0b1e10b2fp-60, -0x1.aea2c72d05c08p-58, 0x1.dd7009902bf32p-58,
    0x1.e48fb0500efd5p-59, 0x1.2ba0b44cfaee5p-59, 0x1.9badefe942718p-60,
    -0x1.c05cf1d753621p-59, 0x1.d192d0619fa68p-60, -0x1.27c8e8416e717p-60,
    -0x1.46662d417cecep-62, 0x0.0p+0, -0x1.e44b7e3711e7fp-67,
    -0x1.83092c5964281p-62, -0x1.6d80ab38e9430p-62, 0x1.33e3f04f1ef25p-60,
    -0x1.5bfa937f551b7p-59, 0x1.8d3ca87b92968p-63, 0x1.0a34531f67db5p-59,
    0x1.85f325c5bbacdp-59, 0x1.2189705cf74cap-58, -0x1.3599f227becbbp-58,
    -0x1.5b61c65e5741ap-58, -0x1.20db323097324p-59, -0x1.294d2f5668495p-58,
    -0x1.d7a16eab1e2adp-59, 0x1.2eb0bf7c0b0d9p-59, -0x1.61578001e015ap-60,
    -0x1.5746b9981b36cp-58, 0x1.9a5dc5e9030adp-57, -0x1.fbe7ee5c69946p-57,
    0x1.301771c407dc0p-57, 0x1.e6cb62af18a02p-62, 0x1.7d3d950f87e23p-59,
    -0x1.546ff8a470d3ap-57, -0x1.bc60efafc6f6cp-58, 0x1.d551d97132e87p-57,
    -0x1.1072534a57e7dp-57, -0x1.9f7fdbfa08d9ap-57, -0x1.26fb3e2b1d1dap-57,
    0x1.24dc46c1ea664p-57, 0x1.a3398064df33ep-57, 0x1.cfce744870f57p-58,
    -0x1.4f689f8434011p-57, -0x1.a37794d03657dp-58, -0x1.61578001e015ep-59,
    0x1.e8637950dc20dp-57, 0x1.355519b0de535p-57, -0x1.08ec217a5022dp-57,
    0x1.bdcfde8061c03p-56, 0x1.3f415699663ecp-63, 0x1.ce63eab883727p-61,
    0x1.9f1a39d500e3cp-56, -0x1.dbd7ac258a2bdp-58, 0x1.7ad24c13f040fp-56,
    -0x1.1e99b72bd7bf2p-57, -0x1.16ea62c048cfbp-56, 0x1.cbcd735d03424p-60,
    -0x1.f79f6c1059cdbp-57, -0x1.7a42642661c62p-61, -0x1.4b366b609027ap-58,
    -0x1.d8db0a7cc1543p-56, -0x1.fb2a49af933e8p-57, -0x1.7dc9c7c23801fp-56,
    -0x1.090a0dd59fe35p-58
};
static const double KM_SINT_SH[52] = {
    0x0.0p+0, 0x1.fffaaaaeeeed5p-7, 0x1.ffeaaaeeee86fp-6,
    0x1.7fdc01032fba9p-5, 0x1.ffaaaeeed4edbp-5, 0x1.3facb12d1755bp-4,
    0x1.7f701032550e4p-4, 0x1.bf1b78568391dp-4, 0x1.feaaeee86ee36p-4,
    0x1.1f0d3d7afceafp-3, 0x1.3eb312c5d66cbp-3, 0x1.5e44fcfa126f3p-3,
    0x1.7dc102fbaf2b5p-3, 0x1.9d252d0cec312p-3, 0x1.bc6f84edc6199p-3,
    0x1.db9e15fb5a5d0p-3, 0x1.faaeed4f31577p-3, 0x1.0cd00cef36436p-2,
    0x1.1c37d64c6b876p-2, 0x1.2b8ddc43eb49fp-2, 0x1.3ad129769d3d8p-2,
    0x1.4a00c9b0f3d20p-2, 0x1.591bc9fa2f597p-2, 0x1.682138a38d7f7p-2,
    0x1.7710255764214p-2, 0x1.85e7a12826949p-2, 0x1.94a6be9f546c5p-2,
    0x1.a34c91cc50ccap-2, 0x1.b1d8305321617p-2, 0x1.c048b17b140a3p-2,
    0x1.ce9d2e3d4a51fp-2, 0x1.dcd4c15329c9ap-2, 0x1.eaee8744b05f0p-2,
    0x1.f8e99e76abc97p-2, 0x1.0362939c69955p-1, 0x1.0a4021e9e1001p-1,
    0x1.110d0c4b69c3bp-1, 0x1.17c8e5f2eedb0p-1, 0x1.1e7343236574cp-1,
    0x1.250bb93788bbbp-1, 0x1.2b91dea88421ep-1, 0x1.32054b148bc4fp-1,
    0x1.386597456282bp-1, 0x1.3eb25d36cd53ap-1, 0x1.44eb381cf386bp-1,
    0x1.4b0fc46aab761p-1, 0x1.511f9fd7b351cp-1, 0x1.571a6966d59b3p-1,
    0x1.5cffc16bf8f0dp-1, 0x1.62cf49921ac79p-1, 0x1.6888a4e134b2fp-1,
    0x1.6e2b77c40bde1p-1
};
static const double KM_SINT_SL[52] = {
    0x0.0p+0, -0x1.2ab639a9f0776p-63, -0x1.cd406fb224ae2p-60,
    -0x1.599bdf46e997ap-59, -0x1.2d16d32684b69p-59, -0x1.921915299468bp-58,
    0x1.afc2d1800501ap-60, 0x1.e91841dea4cc8p-58, -0x1.afcb2bcc6f03bp-59,
    -0x1.6ef95099769a5p-57, 0x1.47d666b66cb91p-57, -0x1.6f443063f89b6p-57,
    0x1.5ab50e23c97c3p-59, 0x1.9c43d80b1137dp-58, 0x1.9c1a56a7b0cabp-57,
    -0x1.32e20d6cc6fc2p-57, -0x1.15d88508e32b8p-57, -0x1.9fb0a0c93e2b4p-56,
    0x1.46076fe0dcff4p-56, 0x1.1553899f2d807p-57, 0x1.03d550487839ap-63,
    0x1.823ba6bb08eadp-56, 0x1.7c74bac3fe0cbp-57, -0x1.d889202444aadp-56,
    -0x1.6ead7314bb6cep-57, 0x1.8a40e9b5face0p-56, -0x1.69ce13e683f58p-56,
    -0x1.a310e3b50cecdp-58, -0x1.ae242cb99f519p-56, 0x1.19fe6757e9fa7p-57,
    -0x1.2fc8a12dae298p-57, 0x1.0d4c6e171fd9ap-56, -0x1.789b43c9b027dp-58,
    0x1.9d950af2d00a3p-58, -0x1.2d8cd78397b01p-55, -0x1.6f643a13914f6p-55,
    0x1.d918998809981p-55, 0x1.35e57102e2488p-57, 0x1.22a3fa4f41d5ap-56,
    0x1.ea3d02457bccep-56, -0x1.fa371db216ab0p-55, 0x1.f6b42095a135bp-55,
    -0x1.10fada93b07a8p-56, -0x1.be570e1570fc0p-58, -0x1.3ed6c1e6a5505p-55,
    0x1.0da05738cc59cp-61, -0x1.5c0e861c48831p-55, 0x1.c843b4d0fb197p-58,
    0x1.96cb370eb578ap-55, -0x1.edd9855b6241ap-55, -0x1.6b7d37644d5e6p-55,
    -0x1.0e729857fad53p-56
};
static const double KM_SINT_CH[52] = {
    0x1.0000000000000p+0, 0x1.fff000155549fp-1, 0x1.ffc00155527d3p-1,
    0x1.ff7006bfdf99fp-1, 0x1.ff0015549f4d3p-1, 0x1.fe7034129ef6fp-1,
    0x1.fdc06bf7e6b9bp-1, 0x1.fcf0c800e99b1p-1, 0x1.fc015527d5bd3p-1,
    0x1.faf22263c4bd3p-1, 0x1.f9c340a7cc428p-1, 0x1.f874c2e1eecf6p-1,
    0x1.f706bdf9ece1cp-1, 0x1.f57948cff6797p-1, 0x1.f3cc7c3b3d16ep-1,
    0x1.f20073086649fp-1, 0x1.f01549f7deea1p-1, 0x1.ee0b1fbc0f11cp-1,
    0x1.ebe214f76efa8p-1, 0x1.e99a4c3a7cd83p-1, 0x1.e733ea0193d40p-1,
    0x1.e4af14b2a449cp-1, 0x1.e20bf49acd6c1p-1, 0x1.df4ab3ebd875ep-1,
    0x1.dc6b7eb995912p-1, 0x1.d96e82f71a9dcp-1, 0x1.d653f073e4040p-1,
    0x1.d31bf8d8d7c06p-1, 0x1.cfc6cfa52ad9fp-1, 0x1.cc54aa2b2972ep-1,
    0x1.c8c5bf8ce1a84p-1, 0x1.c51a48b8b175ep-1, 0x1.c1528065b7d50p-1,
    0x1.bd6ea310294f5p-1, 0x1.b96eeef58840ep-1, 0x1.b553a410c104ep-1,
    0x1.b11d04162a4c6p-1, 0x1.accb526f69de5p-1, 0x1.a85ed4373e02dp-1,
    0x1.a3d7d0352bdcfp-1, 0x1.9f368ed912f85p-1, 0x1.9a7b5a36a6514p-1,
    0x1.95a67e00cb1fdp-1, 0x1.90b84784ddaf7p-1, 0x1.8bb105a5dc900p-1,
    0x1.869108d77a6c6p-1, 0x1.8158a31916d5dp-1, 0x1.7c0827f09e54fp-1,
    0x1.769fec655211fp-1, 0x1.712046fa77678p-1, 0x1.6b898fa9efb5dp-1,
    0x1.65dc1fdeb8cbap-1
};
static const double KM_SINT_CL[52] = {
    0x0.0p+0, 0x1.28a28a03a5ef3p-55, -0x1.3b54492d89b5bp-55,
    -0x1.8b3b560648d5fp-56, 0x1.328387b99426fp-55, -0x1.cbf4337c96f97p-57,
    0x1.31902b535f8dbp-55, 0x1.ea3d786d186acp-57, 0x1.b68f35094efb8p-55,
    -0x1.52ace133a2769p-58, 0x1.c5b6b063b7462p-55, -0x1.c6514e1332b16p-55,
    -0x1.698c80c36dcb4p-55, 0x1.e3a0d3e03b1d4p-57, -0x1.21a3ad28a3494p-57,
    0x1.b940416c1984bp-56, 0x1.d3c1e99e5cafdp-55, -0x1.bfd2380bbc3b1p-59,
    -0x1.02f9f12ba543ep-55, -0x1.2264b1bc53ce8p-55, -0x1.6428b3546ce13p-55,
    -0x1.68ca02e8a6833p-55, -0x1.660aec7ef636bp-58, -0x1.e2d8a7e6736c4p-55,
    0x1.4b364776dcd35p-58, 0x1.ff61bd5d2039dp-55, -0x1.76236434bec37p-55,
    0x1.e60dd3089cbddp-56, 0x1.8b5b5508f2a0dp-55, 0x1.4ee162ba83a98p-57,
    0x1.ab3d1a1590123p-56, -0x1.1bbb43b9aa880p-57, -0x1.892111312e828p-55,
    0x1.31bbcc88c109dp-56, 0x1.45a3cc78fade0p-58, 0x1.8ff7947027a15p-58,
    0x1.1dd561efbc0c2p-56, 0x1.8fb6a8dd6b6ccp-55, 0x1.9be06385ec792p-57,
    -0x1.68dbaeca19669p-55, -0x1.1d200c5791606p-55, 0x1.722cfcc9fa7a9p-55,
    -0x1.0befda21f862dp-55, -0x1.0feb10ab93b87p-56, 0x1.863e03e9474c1p-55,
    0x1.338ffe2bfe9ddp-56, -0x1.de8b90b8228dep-57, -0x1.c73d6d72aee68p-57,
    -0x1.827d5cf8c68c5p-57, 0x1.425b0a5029c81p-55, 0x1.15ac786ccf4b2p-56,
    -0x1.97c1b47337c77p-58
};

#define KM_SIXTH_H 0x1.5555555555555p-3
#define KM_SIXTH_L 0x1.5555555555555p-57
#define KM_THIRD_H 0x1.5555555555555p-2
#define KM_THIRD_L 0x1.5555555555555p-56

/* sin and cos of r = rh + rl, |r| <= pi/4 (+ a rounding): r = j/64 + d,
 * sin r = S_j cos d + C_j sin d, cos r = C_j cos d - S_j sin d, with the
 * double-double sin/cos(j/64) of the table and
 *   sin d = d - d^3/6 + d^5 (1/5! - d^2/7! + d^4/9! - d^6/11!),
 *   cos d = 1 - d^2/2 + d^4 (1/4! - d^2/6! + d^4/8! - d^6/10!),
 * |d| <= 1/128: the first two terms of each in double-double, the rest in
 * double. */
KM_FN void km_sincos_dd(double rh, double rl, double *s_out, double *c_out)
{
    int neg = rh < 0.0;
    if (neg) {
        rh = -rh;
        rl = -rl;
    }
    int j = (int)km_rint_small(rh * 64.0);
    if (j > 51) j = 51;
    double dh0 = rh - (double)j * 0.015625; /* exact (Sterbenz, or j = 0) */
    double dh, dl;
    km_two_sum(dh0, rl, &dh, &dl);
    double zh = dh * dh;
    double zl = km_two_prod_err(dh, dh, zh) + 2.0 * dh * dl;
    /* sin d */
    double d3h, d3l;
    km_dd_mul(dh, dl, zh, zl, &d3h, &d3l);
    double t3h, t3l;
    km_dd_mul(d3h, d3l, -KM_SIXTH_H, -KM_SIXTH_L, &t3h, &t3l);
    double ps = ((zh * zh) * dh) * (0x1.1111111111111p-7 +
                zh * (-0x1.a01a01a01a01ap-13 +
                zh * (0x1.71de3a556c734p-19 +
                zh * -0x1.ae64567f544e4p-26)));
    double sdh, sdl;
    km_dd_add(t3h, t3l, ps, 0.0, &sdh, &sdl);
    km_dd_add(dh, dl, sdh, sdl, &sdh, &sdl);
    /* cos d */
    double q = (zh * zh) * (0x1.5555555555555p-5 +
               zh * (-0x1.6c16c16c16c17p-10 +
               zh * (0x1.a01a01a01a01ap-16 +
               zh * -0x1.27e4fb7789f5cp-22)));
    double cdh, cdl;
    km_dd_add(-0.5 * zh, -0.5 * zl, q, 0.0, &cdh, &cdl);
    km_dd_add(1.0, 0.0, cdh, cdl, &cdh, &cdl);
    /* the addition formulas */
    double Sh = KM_SINT_SH[j], Sl = KM_SINT_SL[j], Ch = KM_SINT_CH[j], Cl = KM_SINT_CL[j];
    double ah, al, bh, bl, sh, sl, ch, cl;
    km_dd_mul(Sh, Sl, cdh, cdl, &ah, &al);
    km_dd_mul(Ch, Cl, sdh, sdl, &bh, &bl);
    km_dd_add(ah, al, bh, bl, &sh, &sl);
    km_dd_mul(Ch, Cl, cdh, cdl, &ah, &al);
    km_dd_mul(Sh, Sl, sdh, sdl, &bh, &bl);
    km_dd_add(ah, al, -bh, -bl, &ch, &cl);
    double sv = sh + sl;
    *s_out = neg ? -sv : sv;
    *c_out = ch + cl;
}

/* Arguments with |x| >= KM_TRIG_MAX lose the exactness of the reduction;
 * the tracking code never forms them (radians of lon/lat, 2*pi*u). */
#define KM_TRIG_MAX 1.5e6

KM_FN double km_sin(double x)
{
    if (!(km_fabs(x) < KM_TRIG_MAX)) {
        if (km_isnan(x) || km_fabs(x) == __builtin_inf()) return x - x;
    }
    if (x == 0.0) return x; /* keeps the sign of zero */
    double rh, rl, sv, cv;
    int n = km_reduce_pio2(x, &rh, &rl);
    km_sincos_dd(rh, rl, &sv, &cv);
    switch (n & 3) {
    case 0: return sv;
    case 1: return cv;
    case 2: return -sv;
    default: return -cv;
    }
}

KM_FN double km_cos(double x)
{
    if (!(km_fabs(x) < KM_TRIG_MAX)) {
        if (km_isnan(x) || km_fabs(x) == __builtin_inf()) return x - x;
    }
    double rh, rl, sv, cv;
    int n = km_reduce_pio2(x, &rh, &rl);
    km_sincos_dd(rh, rl, &sv, &cv);
    switch (n & 3) {
    case 0: return cv;
    case 1: return -sv;
    case 2: return -cv;
    default: return sv;
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
    double rh, rl, sk, ck;
    int n = km_reduce_pio2(x, &rh, &rl);
    km_sincos_dd(rh, rl, &sk, &ck);
    switch (n & 3) {
    case 0: *s = sk; *c = ck; break;
    case 1: *s = ck; *c = -sk; break;
    case 2: *s = -sk; *c = -ck; break;
    default: *s = -ck; *c = sk; break;
    }
    if (x == 0.0) *s = x;
}

/* ------------------------------------------------------------------ atan */

static const double KM_ATANT_H[65] = {
    0x0.0p+0, 0x1.fff555bbb729bp-7, 0x1.ffd55bba97625p-6,
    0x1.7fb818430da2ap-5, 0x1.ff55bb72cfdeap-5, 0x1.3f59f0e7c559dp-4,
    0x1.7ee182602f10fp-4, 0x1.be39ebe6f07c3p-4, 0x1.fd5ba9aac2f6ep-4,
    0x1.1e1fafb043727p-3, 0x1.3d6eee8c6626cp-3, 0x1.5c9811e3ec26ap-3,
    0x1.7b97b4bce5b02p-3, 0x1.9a6a8e96c8626p-3, 0x1.b90d7529260a2p-3,
    0x1.d77d5df205736p-3, 0x1.f5b75f92c80ddp-3, 0x1.09dc597d86362p-2,
    0x1.18bf5a30bf178p-2, 0x1.278372057ef46p-2, 0x1.362773707ebccp-2,
    0x1.44aa436c2af0ap-2, 0x1.530ad9951cd4ap-2, 0x1.614840309cfe2p-2,
    0x1.6f61941e4def1p-2, 0x1.7d5604b63b3f7p-2, 0x1.8b24d394a1b25p-2,
    0x1.98cd5454d6b18p-2, 0x1.a64eec3cc23fdp-2, 0x1.b3a911da65c6cp-2,
    0x1.c0db4c94ec9f0p-2, 0x1.cde53432c1351p-2, 0x1.dac670561bb4fp-2,
    0x1.e77eb7f175a34p-2, 0x1.f40dd0b541418p-2, 0x1.0039c73c1a40cp-1,
    0x1.0657e94db30d0p-1, 0x1.0c6145b5b43dap-1, 0x1.1255d9bfbd2a9p-1,
    0x1.1835a88be7c13p-1, 0x1.1e00babdefeb4p-1, 0x1.23b71e2cc9e6ap-1,
    0x1.2958e59308e31p-1, 0x1.2ee628406cbcap-1, 0x1.345f01cce37bbp-1,
    0x1.39c391cd4171ap-1, 0x1.3f13fb89e96f4p-1, 0x1.445065b795b56p-1,
    0x1.4978fa3269ee1p-1, 0x1.4e8de5bb6ec04p-1, 0x1.538f57b89061fp-1,
    0x1.587d81f732fbbp-1, 0x1.5d58987169b18p-1, 0x1.6220d115d7b8ep-1,
    0x1.66d663923e087p-1, 0x1.6b798920b3d99p-1, 0x1.700a7c5784634p-1,
    0x1.748978fba8e0fp-1, 0x1.78f6bbd5d315ep-1, 0x1.7d528289fa093p-1,
    0x1.819d0b7158a4dp-1, 0x1.85d69576cc2c5p-1, 0x1.89ff5ff57f1f8p-1,
    0x1.8e17aa99cc05ep-1, 0x1.921fb54442d18p-1
};
static const double KM_ATANT_L[65] = {
    0x0.0p+0, -0x1.220c39d4dff50p-61, -0x1.5ec431444912cp-60,
    -0x1.86ef8f794f105p-63, -0x1.c934d86d23f1dp-60, 0x1.ac4ce285df847p-58,
    -0x1.cfb654c0c3d98p-58, 0x1.f7b8f29a05987p-58, -0x1.cd37686760c17p-59,
    -0x1.b485914dacf8cp-59, 0x1.61a3b0ce9281bp-57, -0x1.054ab2c010f3dp-58,
    0x1.347b0b4f881cap-58, 0x1.cf601e7b4348ep-59, 0x1.17b10d2e0e5abp-61,
    0x1.c648d1534597ep-57, 0x1.8ab6e3cf7afbdp-57, 0x1.62e47390cb865p-56,
    0x1.30ca4748b1bf9p-57, -0x1.077cdd36dfc81p-56, -0x1.963a544b672d8p-57,
    -0x1.5d5e43c55b3bap-56, -0x1.2566480884082p-57, -0x1.a725715711f00p-56,
    -0x1.c63aae6f6e918p-56, 0x1.69c885c2b249ap-56, 0x1.b6d0ba3748fa8p-56,
    0x1.9e6c988fd0a77p-56, -0x1.24dec1b50b7ffp-56, 0x1.ae187b1ca5040p-56,
    -0x1.cc1ce70934c34p-56, -0x1.a2cfa4418f1adp-56, 0x1.a2b7f222f65e2p-56,
    0x1.0e53dc1bf3435p-56, -0x1.a3992dc382a23p-57, -0x1.b32c949c9d593p-55,
    -0x1.d5b495f6349e6p-56, 0x1.974fa13b5404fp-58, -0x1.2bdaee1c0ee35p-58,
    0x1.c621cec00c301p-55, -0x1.928df287a668fp-58, 0x1.c421c9f38224ep-57,
    -0x1.09e73b0c6c087p-56, 0x1.c5d5e9ff0cf8dp-55, 0x1.1021137c71102p-55,
    -0x1.2304331d8bf46p-55, 0x1.ecf8b492644f0p-56, -0x1.f76d0163f79c8p-56,
    0x1.2419a87f2a458p-56, 0x1.4a33dbeb3796cp-55, -0x1.1bb74abda520cp-55,
    -0x1.5e5c9d8c5a950p-56, 0x1.0028e4bc5e7cap-57, -0x1.2b785350ee8c1p-57,
    -0x1.6ea6febe8bbbap-56, -0x1.a80386188c50ep-55, -0x1.8c34d25aadef6p-56,
    0x1.7b2a6165884a1p-59, 0x1.406a089803740p-55, 0x1.560821e2f3aa9p-55,
    -0x1.bf76229d3b917p-56, 0x1.6b66e7fc8b8c3p-57, -0x1.55b9a5e177a1bp-55,
    -0x1.ec182ab042f61p-56, 0x1.1a62633145c07p-55
};

KM_FN void km_dd_div(double ah, double al, double bh, double bl, double *qh, double *ql)
{
    /* (ah + al) / (bh + bl): q1 = ah / bh, the remainder (a - q1 b) formed
     * exactly to first order (ah - q1 bh is exact: q1 bh is within a
     * factor 2 of ah), q2 = remainder / bh */
    double q1 = ah / bh;
    double p = q1 * bh;
    double pe = km_two_prod_err(q1, bh, p);
    double r = (((ah - p) - pe) + al) - q1 * bl;
    km_fast_two_sum(q1, r / bh, qh, ql);
}

KM_FN void km_dd_sqrt(double ah, double al, double *sh, double *sl)
{
    /* sqrt(ah + al), ah > 0: s = sqrt(ah), then one Newton correction
     * (a - s^2) / (2 s) with s^2 formed exactly */
    double s = km_sqrt(ah);
    double p = s * s;
    double pe = km_two_prod_err(s, s, p);
    double r = ((ah - p) - pe) + al;
    km_fast_two_sum(s, r / (2.0 * s), sh, sl);
}

/* atan of t = th + tl, 0 <= t <= 1 (+ a rounding), as a double-double:
 * j = rint(64 t), c = j/64, atan t = atan c + atan u with
 * u = (t - c) / (1 + t c), |u| <= 1/128 (+), and
 *   atan u = u + u^3 (-1/3 + u^2 (1/5 - u^2/7 + u^4/9 - u^6/11 + u^8/13)),
 * u, u^2, u^3 and the -1/3 in double-double, the bracket in double:
 * relative error before the final rounding below 2^-84. */
KM_FN void km_atan01_dd(double th, double tl, double *rh, double *rl)
{
    int j = (int)km_rint_small(th * 64.0);
    if (j > 64) j = 64;
    double c = (double)j * 0.015625;
    double nh, nl;
    km_two_sum(th - c, tl, &nh, &nl); /* th - c exact (Sterbenz, or c = 0) */
    double ph = th * c;
    double pl = km_two_prod_err(th, c, ph) + tl * c;
    double dh, dl;
    km_two_sum(1.0, ph, &dh, &dl);
    km_fast_two_sum(dh, dl + pl, &dh, &dl);
    double uh, ul;
    km_dd_div(nh, nl, dh, dl, &uh, &ul);
    double zh = uh * uh;
    double zl = km_two_prod_err(uh, uh, zh) + 2.0 * uh * ul;
    double u3h, u3l;
    km_dd_mul(uh, ul, zh, zl, &u3h, &u3l);
    double q = zh * (0x1.999999999999ap-3 +
               zh * (-0x1.2492492492492p-3 +
               zh * (0x1.c71c71c71c71cp-4 +
               zh * (-0x1.745d1745d1746p-4 +
               zh * 0x1.3b13b13b13b14p-4))));
    double Ph, Pl, ch, cl, ah, al;
    km_dd_add(-KM_THIRD_H, -KM_THIRD_L, q, 0.0, &Ph, &Pl);
    km_dd_mul(u3h, u3l, Ph, Pl, &ch, &cl);
    km_dd_add(uh, ul, ch, cl, &ah, &al);
    km_dd_add(KM_ATANT_H[j], KM_ATANT_L[j], ah, al, rh, rl);
}

/* atan2 of y = yh + yl > 0 and x = xh + xl > 0 as a double-double (both of
 * moderate range: their products and squares neither overflow nor
 * underflow): the smaller over the larger, atan01, pi/2 minus it when
 * y > x. */
KM_FN void km_atan2_pos_dd(double yh, double yl, double xh, double xl, double *rh, double *rl)
{
    int swap = yh > xh;
    double th, tl, ah, al;
    if (swap)
        km_dd_div(xh, xl, yh, yl, &th, &tl);
    else
        km_dd_div(yh, yl, xh, xl, &th, &tl);
    km_atan01_dd(th, tl, &ah, &al);
    if (swap) km_dd_add(KM_PIO2_HI, KM_PIO2_LO, -ah, -al, &ah, &al);
    *rh = ah;
    *rl = al;
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
    double rh, rl;
    if (ay < ax * 0x1p-60 || ax < ay * 0x1p-60) {
        /* one far smaller: atan of a tiny ratio t is t (t^3/3 is below
         * half an ulp of t), or pi/2 - t */
        if (ay < ax) {
            rh = ay / ax;
            rl = 0.0;
        } else {
            km_fast_two_sum(KM_PIO2_HI, KM_PIO2_LO - ax / ay, &rh, &rl);
        }
    } else {
        /* both scaled by powers of two until the larger is in [1, 2) (the
         * smaller then in [2^-61, 2)): exact */
        double big = ax > ay ? ax : ay;
        int e = (int)((km_bits(big) >> 52) & 0x7ff) - 1023;
        if (e > 1000) {
            ax *= 0x1p-600;
            ay *= 0x1p-600;
            e -= 600;
        } else if (e < -1000) {
            ax *= 0x1p600;
            ay *= 0x1p600;
            big *= 0x1p600;
            e = (int)((km_bits(big) >> 52) & 0x7ff) - 1023;
        }
        double sc = km_from_bits((uint64_t)(1023 - e) << 52);
        km_atan2_pos_dd(ay * sc, 0.0, ax * sc, 0.0, &rh, &rl);
    }
    if (__builtin_signbit(x)) km_dd_add(KM_PI_HI, KM_PI_LO, -rh, -rl, &rh, &rl);
    return km_copysign(rh + rl, y);
}

KM_FN double km_atan(double x)
{
    if (km_isnan(x)) return x;
    return km_atan2(x, 1.0);
}

/* sqrt(1 - a^2), 0 <= a < 1, as a double-double: (1 - a)(1 + a) formed
 * exactly to second order, then km_dd_sqrt */
KM_FN void km_cosasin_dd(double a, double *sh, double *sl)
{
    double oh, ol, ph, pl, wh, wl;
    km_two_sum(1.0, -a, &oh, &ol);
    km_two_sum(1.0, a, &ph, &pl);
    km_dd_mul(oh, ol, ph, pl, &wh, &wl);
    km_dd_sqrt(wh, wl, sh, sl);
}

/* asin x = atan2(x, sqrt(1 - x^2)) and acos x = atan2(sqrt(1 - x^2), x)
 * in double-double, rounded once */
KM_FN double km_asin(double x)
{
    double a = km_fabs(x);
    if (!(a <= 1.0)) return (x - x) / (x - x); /* NaN for |x| > 1 or NaN */
    if (a < 0x1p-26) return x; /* x^3/6 is below half an ulp of x */
    if (a == 1.0) return km_copysign(KM_PIO2_HI, x);
    double sh, sl, rh, rl;
    km_cosasin_dd(a, &sh, &sl);
    km_atan2_pos_dd(a, 0.0, sh, sl, &rh, &rl);
    return km_copysign(rh + rl, x);
}

KM_FN double km_acos(double x)
{
    double a = km_fabs(x);
    if (!(a <= 1.0)) return (x - x) / (x - x);
    if (x == 1.0) return 0.0;
    if (x == -1.0) return KM_PI_HI;
    if (a < 0x1p-60) return KM_PIO2_HI;
    double sh, sl, rh, rl;
    km_cosasin_dd(a, &sh, &sl);
    km_atan2_pos_dd(sh, sl, a, 0.0, &rh, &rl);
    if (x < 0.0) km_dd_add(KM_PI_HI, KM_PI_LO, -rh, -rl, &rh, &rl);
    return rh + rl;
}

/* ------------------------------------------------------------------- log */

/* log x, x finite and > 0, as a double-double */
KM_FN void km_log_dd(double x, double *out_h, double *out_l)
{
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
    int i = (int)km_rint_small((m - 1.0) * 128.0); /* -37 .. 53, m - 1 exact */
    double c = KM_LOGT_C[i + 37];
    /* r = m c - 1 as a double-double: m c exact by Dekker, then p - 1 exact */
    double p = m * c;
    double pe = km_two_prod_err(m, c, p);
    double rh, rl;
    km_two_sum(p - 1.0, pe, &rh, &rl);
    /* log1p(r) = r - r^2/2 + r^3/3 + r^4 P(r) */
    double zh = rh * rh;
    double zl = km_two_prod_err(rh, rh, zh) + 2.0 * rh * rl;
    double r3h, r3l, t3h, t3l;
    km_dd_mul(rh, rl, zh, zl, &r3h, &r3l);
    km_dd_mul(r3h, r3l, KM_THIRD_H, KM_THIRD_L, &t3h, &t3l);
    double tail = (zh * zh) * (-0x1.0000000000000p-2 +
                  rh * (0x1.999999999999ap-3 +
                  rh * (-0x1.5555555555555p-3 +
                  rh * (0x1.2492492492492p-3 +
                  rh * (-0x1.0000000000000p-3 +
                  rh * (0x1.c71c71c71c71cp-4 +
                  rh * (-0x1.999999999999ap-4 +
                  rh * (0x1.745d1745d1746p-4 +
                  rh * -0x1.5555555555555p-4))))))));
    double sh, sl;
    km_dd_add(t3h, t3l, tail, 0.0, &sh, &sl);
    km_dd_add(-0.5 * zh, -0.5 * zl, sh, sl, &sh, &sl);
    km_dd_add(rh, rl, sh, sl, &sh, &sl);
    km_dd_add(KM_LOGT_H[i + 37], KM_LOGT_L[i + 37], sh, sl, &sh, &sl);
    if (k != 0) {
        /* k ln2 = k LN2_HI (exact: 42 bits times |k| < 2^11) + k LN2_LO (exact product) */
        double dk = (double)k;
        double kl = dk * KM_LN2_LO;
        double kle = km_two_prod_err(dk, KM_LN2_LO, kl);
        double kh, khl;
        km_fast_two_sum(dk * KM_LN2_HI, kl, &kh, &khl);
        km_dd_add(kh, khl + kle, sh, sl, &sh, &sl);
    }
    *out_h = sh;
    *out_l = sl;
}

KM_FN double km_log(double x)
{
    if (km_isnan(x)) return x;
    if (x < 0.0) return (x - x) / (x - x);
    if (x == 0.0) return -__builtin_inf();
    if (x == __builtin_inf()) return x;
    double h, l;
    km_log_dd(x, &h, &l);
    return h + l;
}

/* ------------------------------------------------------------------- exp */

static const double KM_EXPT_H[64] = {
    0x1.0000000000000p+0, 0x1.02c9a3e778061p+0, 0x1.059b0d3158574p+0,
    0x1.0874518759bc8p+0, 0x1.0b5586cf9890fp+0, 0x1.0e3ec32d3d1a2p+0,
    0x1.11301d0125b51p+0, 0x1.1429aaea92de0p+0, 0x1.172b83c7d517bp+0,
    0x1.1a35beb6fcb75p+0, 0x1.1d4873168b9aap+0, 0x1.2063b88628cd6p+0,
    0x1.2387a6e756238p+0, 0x1.26b4565e27cddp+0, 0x1.29e9df51fdee1p+0,
    0x1.2d285a6e4030bp+0, 0x1.306fe0a31b715p+0, 0x1.33c08b26416ffp+0,
    0x1.371a7373aa9cbp+0, 0x1.3a7db34e59ff7p+0, 0x1.3dea64c123422p+0,
    0x1.4160a21f72e2ap+0, 0x1.44e086061892dp+0, 0x1.486a2b5c13cd0p+0,
    0x1.4bfdad5362a27p+0, 0x1.4f9b2769d2ca7p+0, 0x1.5342b569d4f82p+0,
    0x1.56f4736b527dap+0, 0x1.5ab07dd485429p+0, 0x1.5e76f15ad2148p+0,
    0x1.6247eb03a5585p+0, 0x1.6623882552225p+0, 0x1.6a09e667f3bcdp+0,
    0x1.6dfb23c651a2fp+0, 0x1.71f75e8ec5f74p+0, 0x1.75feb564267c9p+0,
    0x1.7a11473eb0187p+0, 0x1.7e2f336cf4e62p+0, 0x1.82589994cce13p+0,
    0x1.868d99b4492edp+0, 0x1.8ace5422aa0dbp+0, 0x1.8f1ae99157736p+0,
    0x1.93737b0cdc5e5p+0, 0x1.97d829fde4e50p+0, 0x1.9c49182a3f090p+0,
    0x1.a0c667b5de565p+0, 0x1.a5503b23e255dp+0, 0x1.a9e6b5579fdbfp+0,
    0x1.ae89f995ad3adp+0, 0x1.b33a2b84f15fbp+0, 0x1.b7f76f2fb5e47p+0,
    0x1.bcc1e904bc1d2p+0, 0x1.c199bdd85529cp+0, 0x1.c67f12e57d14bp+0,
    0x1.cb720dcef9069p+0, 0x1.d072d4a07897cp+0, 0x1.d5818dcfba487p+0,
    0x1.da9e603db3285p+0, 0x1.dfc97337b9b5fp+0, 0x1.e502ee78b3ff6p+0,
    0x1.ea4afa2a490dap+0, 0x1.efa1bee615a27p+0, 0x1.f50765b6e4540p+0,
    0x1.fa7c1819e90d8p+0
};
static const double KM_EXPT_L[64] = {
    0x0.0p+0, -0x1.19083535b085dp-56, 0x1.d73e2a475b465p-55,
    0x1.186be4bb284ffp-57, 0x1.8a62e4adc610bp-54, 0x1.03a1727c57b53p-59,
    -0x1.6c51039449b3ap-54, -0x1.32fbf9af1369ep-54, -0x1.19041b9d78a76p-55,
    0x1.e5b4c7b4968e4p-55, 0x1.e016e00a2643cp-54, 0x1.dc775814a8495p-55,
    0x1.9b07eb6c70573p-54, 0x1.2bd339940e9d9p-55, 0x1.612e8afad1255p-55,
    0x1.0024754db41d5p-54, 0x1.6f46ad23182e4p-55, 0x1.32721843659a6p-54,
    -0x1.63aeabf42eae2p-54, -0x1.5e436d661f5e3p-56, 0x1.ada0911f09ebcp-55,
    -0x1.ef3691c309278p-58, 0x1.89b7a04ef80d0p-59, 0x1.3c1a3b69062f0p-56,
    0x1.d4397afec42e2p-56, -0x1.4b309d25957e3p-54, -0x1.07abe1db13cadp-55,
    0x1.9bb2c011d93adp-54, 0x1.6324c054647adp-54, 0x1.ba6f93080e65ep-54,
    -0x1.383c17e40b497p-54, -0x1.bb60987591c34p-54, -0x1.bdd3413b26456p-54,
    -0x1.bbe3a683c88abp-57, -0x1.16e4786887a99p-55, -0x1.0245957316dd3p-54,
    -0x1.41577ee04992fp-55, 0x1.05d02ba15797ep-56, -0x1.d4c1dd41532d8p-54,
    -0x1.fc6f89bd4f6bap-54, 0x1.6e9f156864b27p-54, 0x1.5cc13a2e3976cp-55,
    -0x1.75fc781b57ebcp-57, -0x1.d185b7c1b85d1p-54, 0x1.c7c46b071f2bep-56,
    -0x1.359495d1cd533p-54, -0x1.d2f6edb8d41e1p-54, 0x1.0fac90ef7fd31p-54,
    0x1.7a1cd345dcc81p-54, -0x1.2805e3084d708p-57, -0x1.5584f7e54ac3bp-56,
    0x1.23dd07a2d9e84p-55, 0x1.11065895048ddp-55, 0x1.2884dff483cadp-54,
    0x1.503cbd1e949dbp-56, -0x1.cbc3743797a9cp-54, 0x1.2ed02d75b3707p-55,
    0x1.c2300696db532p-54, -0x1.1a5cd4f184b5cp-54, 0x1.39e8980a9cc8fp-55,
    -0x1.e9c23179c2893p-54, 0x1.dc7f486a4b6b0p-54, 0x1.9d3e12dd8a18bp-54,
    0x1.74853f3a5931ep-55
};

#define KM_LN2_64_1 0x1.62e42fefa0000p-7 /* 36 bits: n * LN2_64_1 exact, |n| < 2^17 */
#define KM_LN2_64_2 0x1.cf79abc9e0000p-46 /* next 36 bits */
#define KM_LN2_64_3 0x1.d9cc01f97b57ap-85
#define KM_64_OVER_LN2 0x1.71547652b82fep+6

/* exp(xh + xl) = 2^k 2^(j/64) exp(r): n = rint(xh 64/ln2) = 64 k + j,
 * r = x - n ln2/64 as a double-double (ln2/64 in three parts), |r| <=
 * ln2/128 (+), and exp(r) - 1 = r + r^2/2 + r^3/6 + r^4 (1/4! + ... +
 * r^5/9!), the first three terms in double-double: the mantissa
 * 2^(j/64) exp(r) as a double-double with relative error below 2^-85,
 * returned with k.  |xh| < 746. */
KM_FN void km_exp_dd(double xh, double xl, double *mh, double *ml, int *k_out)
{
    double fn = km_rint_small(xh * KM_64_OVER_LN2);
    int n = (int)fn;
    int j = n & 63;
    *k_out = (n - j) / 64;
    double r1 = xh - fn * KM_LN2_64_1; /* exact */
    double rh, rl;
    km_two_sum(r1, -(fn * KM_LN2_64_2), &rh, &rl); /* the product exact */
    rl += xl - fn * KM_LN2_64_3;
    km_fast_two_sum(rh, rl, &rh, &rl);
    double zh = rh * rh;
    double zl = km_two_prod_err(rh, rh, zh) + 2.0 * rh * rl;
    double r3h, r3l, t3h, t3l;
    km_dd_mul(rh, rl, zh, zl, &r3h, &r3l);
    km_dd_mul(r3h, r3l, KM_SIXTH_H, KM_SIXTH_L, &t3h, &t3l);
    double tail = (zh * zh) * (0x1.5555555555555p-5 +
                  rh * (0x1.1111111111111p-7 +
                  rh * (0x1.6c16c16c16c17p-10 +
                  rh * (0x1.a01a01a01a01ap-13 +
                  rh * (0x1.a01a01a01a01ap-16 +
                  rh * 0x1.71de3a556c734p-19)))));
    double eh, el;
    km_dd_add(t3h, t3l, tail, 0.0, &eh, &el);
    km_dd_add(0.5 * zh, 0.5 * zl, eh, el, &eh, &el);
    km_dd_add(rh, rl, eh, el, &eh, &el); /* exp(r) - 1 */
    double Th = KM_EXPT_H[j], Tl = KM_EXPT_L[j];
    double ph, pl;
    km_dd_mul(Th, Tl, eh, el, &ph, &pl);
    km_dd_add(Th, Tl, ph, pl, mh, ml);
}

/* 2^k (mh + ml), rounded once where the result is normal */
KM_FN double km_scale_dd(double mh, double ml, int k)
{
    if (k > 1023) {
        mh *= 0x1p1023;
        ml *= 0x1p1023;
        k -= 1023;
    } else if (k < -1021) {
        /* subnormal range: scaled in two steps (may round twice) */
        double m = (mh + ml) * 0x1p-600;
        return m * km_from_bits((uint64_t)(k + 600 + 1023) << 52);
    }
    return (mh + ml) * km_from_bits((uint64_t)(k + 1023) << 52);
}

KM_FN double km_exp(double x)
{
    if (km_isnan(x)) return x;
    if (x > 709.782712893384) return __builtin_inf();
    if (x < -745.2) return 0.0;
    double mh, ml;
    int k;
    km_exp_dd(x, 0.0, &mh, &ml, &k);
    return km_scale_dd(mh, ml, k);
}

/* x^y for x > 0 finite and y finite: exp(y log x) with log x and the
 * product in double-double (relative error of the result before its
 * rounding below 2^-70 |y log x| + 2^-84); other arguments through
 * exp(y log x) in plain double. */
KM_FN double km_pow(double x, double y)
{
    if (!(x > 0.0) || x == __builtin_inf() || km_isnan(y) || km_fabs(y) == __builtin_inf())
        return km_exp(y * km_log(x));
    if (y == 0.0 || x == 1.0) return 1.0;
    double lh, ll, zh, zl;
    km_log_dd(x, &lh, &ll);
    km_dd_mul(lh, ll, y, 0.0, &zh, &zl);
    if (zh > 709.782712893384) return __builtin_inf();
    if (zh < -745.2) return 0.0;
    double mh, ml;
    int k;
    km_exp_dd(zh, zl, &mh, &ml, &k);
    return km_scale_dd(mh, ml, k);
}

#endif /* KELP_KMATH_H */
