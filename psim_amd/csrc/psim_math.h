/*
 * FP64 transcendentals used on the phonon path, written from their textbook
 * series so that the HIP kernels and the CPU oracle round identically
 * (vendor libms differ in the last ulp, which would split trajectories).
 * Every function uses only IEEE add/mul/div, floor, frexp and ldexp, all
 * exact or correctly rounded on both gfx950 and x86-64; build with
 * -ffp-contract=off on both sides.
 *
 * Methods:
 *   psim_log    : x = m 2^e with m in [sqrt(1/2), sqrt(2)); log(m) = 2 atanh(s),
 *                 s = (m-1)/(m+1), atanh Taylor series to s^21, recombined as
 *                 f - (f^2/2 - s (f^2/2 + R)) (the "hfsq" arrangement of
 *                 Tang / fdlibm); e ln2 added as a two-word constant.
 *   psim_exp    : Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, Taylor
 *                 series of e^r to r^13, scaled by 2^k.
 *   psim_sinh   : Taylor series to x^19 for |x| < 1, (e^x - e^-x)/2 above.
 *   psim_cos2pi : cos(2 pi u) with exact quadrant reduction of 4u, Taylor
 *                 series of sin and cos on |a| <= pi/4.
 *   psim_pow4   : (x x)(x x).
 * Error bounds against glibc are checked by the tests (a few ulp).
 */
#ifndef PSIM_MATH_H
#define PSIM_MATH_H

#include "psim_rng.h" /* PSIM_QUAL */

#if !defined(__HIPCC__)
#include <math.h>
#endif

#define PSIM_PI 3.1415926535897932384626433832795028841971693993751058209749445923
/* ln 2 split so that k * PSIM_LN2_HI is exact for |k| < 2^20. */
#define PSIM_LN2_HI 0x1.62e42fee00000p-1
#define PSIM_LN2_LO 0x1.a39ef35793c76p-33
#define PSIM_INV_LN2 0x1.71547652b82fep+0
/* pi/2 as a two-word constant. */
#define PSIM_PIO2_HI 0x1.921fb54442d18p+0
#define PSIM_PIO2_LO 0x1.1a62633145c07p-54

PSIM_QUAL double psim_log(double x) {
    if (!(x > 0.0)) { return (x == 0.0) ? -__builtin_inf() : __builtin_nan(""); }
    if (x == __builtin_inf()) { return x; }
    int e = 0;
    double m = frexp(x, &e); /* m in [0.5, 1) */
    if (m < 0.70710678118654752440) {
        m = m * 2.0;
        e -= 1;
    }
    const double f = m - 1.0; /* exact */
    const double s = f / (2.0 + f);
    const double z = s * s;
    /* R = 2 (z/3 + z^2/5 + ... + z^10/21) */
    double p = 2.0 / 21.0;
    p = p * z + 2.0 / 19.0;
    p = p * z + 2.0 / 17.0;
    p = p * z + 2.0 / 15.0;
    p = p * z + 2.0 / 13.0;
    p = p * z + 2.0 / 11.0;
    p = p * z + 2.0 / 9.0;
    p = p * z + 2.0 / 7.0;
    p = p * z + 2.0 / 5.0;
    p = p * z + 2.0 / 3.0;
    const double R = z * p;
    const double hfsq = 0.5 * f * f;
    const double dk = (double)e;
    return dk * PSIM_LN2_HI - ((hfsq - (s * (hfsq + R) + dk * PSIM_LN2_LO)) - f);
}

PSIM_QUAL double psim_exp(double x) {
    if (x != x) { return x; }
    if (x > 709.782712893384) { return __builtin_inf(); }
    if (x < -745.1332191019412) { return 0.0; }
    const double kd = floor(x * PSIM_INV_LN2 + 0.5);
    const int k = (int)kd;
    const double r = (x - kd * PSIM_LN2_HI) - kd * PSIM_LN2_LO;
    /* e^r - 1 - r = r^2 (1/2! + r/3! + ... + r^11/13!) */
    double p = 1.0 / 6227020800.0; /* 1/13! */
    p = p * r + 1.0 / 479001600.0;
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    const double em1 = r + (r * r) * p;
    return ldexp(1.0 + em1, k);
}

PSIM_QUAL double psim_sinh(double x) {
    const double a = fabs(x);
    double r;
    if (a < 1.0) {
        const double z = a * a;
        /* sinh(a) = a + a z (1/3! + z/5! + ... + z^8/19!) */
        double p = 1.0 / 121645100408832000.0; /* 1/19! */
        p = p * z + 1.0 / 355687428096000.0; /* 1/17! */
        p = p * z + 1.0 / 1307674368000.0; /* 1/15! */
        p = p * z + 1.0 / 6227020800.0; /* 1/13! */
        p = p * z + 1.0 / 39916800.0; /* 1/11! */
        p = p * z + 1.0 / 362880.0; /* 1/9! */
        p = p * z + 1.0 / 5040.0; /* 1/7! */
        p = p * z + 1.0 / 120.0; /* 1/5! */
        p = p * z + 1.0 / 6.0; /* 1/3! */
        r = a + (a * z) * p;
    } else if (a < 22.0) {
        const double e = psim_exp(a);
        r = 0.5 * (e - 1.0 / e);
    } else {
        r = 0.5 * psim_exp(a);
    }
    return (x < 0.0) ? -r : r;
}

/* cos(2 pi u); exact reduction of 4u into a quadrant and |f| <= 1/2. */
PSIM_QUAL double psim_cos2pi(double u) {
    const double t = 4.0 * u; /* exact */
    double q = floor(t);
    double f = t - q; /* exact, in [0, 1) */
    if (f > 0.5) {
        f = f - 1.0;
        q = q + 1.0;
    }
    const double a = f * PSIM_PIO2_HI + f * PSIM_PIO2_LO; /* |a| <= pi/4 */
    const double z = a * a;
    /* cos a = 1 - z/2 + z^2 (1/4! - z/6! + ... + z^6/16!) */
    double pc = 1.0 / 20922789888000.0; /* 1/16! */
    pc = pc * z - 1.0 / 87178291200.0; /* 1/14! */
    pc = pc * z + 1.0 / 479001600.0; /* 1/12! */
    pc = pc * z - 1.0 / 3628800.0; /* 1/10! */
    pc = pc * z + 1.0 / 40320.0; /* 1/8! */
    pc = pc * z - 1.0 / 720.0; /* 1/6! */
    pc = pc * z + 1.0 / 24.0; /* 1/4! */
    /* sin a = a + a z (-1/3! + z/5! - ... + z^7/17!) */
    double ps = 1.0 / 355687428096000.0; /* 1/17! */
    ps = ps * z - 1.0 / 1307674368000.0; /* 1/15! */
    ps = ps * z + 1.0 / 6227020800.0; /* 1/13! */
    ps = ps * z - 1.0 / 39916800.0; /* 1/11! */
    ps = ps * z + 1.0 / 362880.0; /* 1/9! */
    ps = ps * z - 1.0 / 5040.0; /* 1/7! */
    ps = ps * z + 1.0 / 120.0; /* 1/5! */
    ps = ps * z - 1.0 / 6.0; /* 1/3! */
    const double c = (1.0 - 0.5 * z) + (z * z) * pc;
    const double s = a + (a * z) * ps;
    const long long qi = ((long long)q) & 3;
    return (qi == 0) ? c : (qi == 1) ? -s : (qi == 2) ? -c : s;
}

PSIM_QUAL double psim_pow4(double x) {
    const double x2 = x * x;
    return x2 * x2;
}

#endif /* PSIM_MATH_H */
