/* omath.h -- FP64 elementary functions and the Philox4x32-10 counter RNG of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  These routines exist so that the CPU restatement and the HIP
 * library evaluate exp/log with the same sequence of IEEE-754 basic operations (+ - * /, rint,
 * no contraction), which makes every table index and McICA bit reproducible bit for bit on
 * both sides.  The HIP library carries its own device copy of the same algorithm
 * (woof_nam_amd/csrc/wr_math.h); a GPU test checks the two agree bit for bit and a CPU test
 * checks this copy against libm to a few ulp.
 *
 * exp: Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, with ln2 split into a 42-bit head
 *      (k*head exact for |k| < 2^11) and a tail; e^r from its Taylor series through r^13
 *      (coefficients 1/n! rounded to double), summed as 1 + (r + r^2 q(r)).
 * log: x = 2^e m, m in [sqrt(1/2), sqrt(2)); with f = m - 1 and s = f/(2+f),
 *      log(m) = 2 atanh(s) = f - (f^2/2 - s (f^2/2 + R(s^2))), R from the atanh series
 *      2 s^2/3 + 2 s^4/5 + ... through s^20.
 * pow: exp(y log x) for x > 0 (the only use: RRTMG's CO2/N2O column adjustments).
 */
#ifndef WOOFRAD_OMATH_H
#define WOOFRAD_OMATH_H

#include <math.h>
#include <stdint.h>
#include <string.h>

#define OM_LN2_HI 0x1.62e42fefa3800p-1
#define OM_LN2_LO 0x1.ef35793c76730p-45
#define OM_INV_LN2 0x1.71547652b82fep+0
#define OM_SQRT2 0x1.6a09e667f3bcdp+0

static inline double om_pow2i(int k) /* 2^k for -1022 <= k <= 1023 */
{
    uint64_t b = (uint64_t)(k + 1023) << 52;
    double d;
    memcpy(&d, &b, 8);
    return d;
}

static inline double om_exp(double x)
{
    if (x != x) return x;
    if (x > 709.782712893384) return INFINITY;
    if (x < -745.1332191019412) return 0.0;
    const double kf = rint(x * OM_INV_LN2);
    const double r = (x - kf * OM_LN2_HI) - kf * OM_LN2_LO;
    double q = 0x1.6124613a86d09p-33;          /* 1/13! */
    q = q * r + 0x1.1eed8eff8d898p-29;         /* 1/12! */
    q = q * r + 0x1.ae64567f544e4p-26;         /* 1/11! */
    q = q * r + 0x1.27e4fb7789f5cp-22;         /* 1/10! */
    q = q * r + 0x1.71de3a556c734p-19;         /* 1/9!  */
    q = q * r + 0x1.a01a01a01a01ap-16;         /* 1/8!  */
    q = q * r + 0x1.a01a01a01a01ap-13;         /* 1/7!  */
    q = q * r + 0x1.6c16c16c16c17p-10;         /* 1/6!  */
    q = q * r + 0x1.1111111111111p-7;          /* 1/5!  */
    q = q * r + 0x1.5555555555555p-5;          /* 1/4!  */
    q = q * r + 0x1.5555555555555p-3;          /* 1/3!  */
    q = q * r + 0x1.0000000000000p-1;          /* 1/2!  */
    const double p = 1.0 + (r + (r * r) * q);
    const int k = (int)kf;
    if (k > 1023) return (p * om_pow2i(1023)) * 2.0;
    if (k < -1022) return (p * om_pow2i(k + 600)) * om_pow2i(-600);
    return p * om_pow2i(k);
}

static inline double om_log(double x)
{
    if (x != x || x < 0.0) return NAN;
    if (x == 0.0) return -INFINITY;
    if (x == INFINITY) return x;
    uint64_t b;
    memcpy(&b, &x, 8);
    int eadj = 0;
    if ((b >> 52) == 0) { /* subnormal */
        x = x * 0x1p54;
        memcpy(&b, &x, 8);
        eadj = -54;
    }
    int e = (int)((b >> 52) & 0x7ff) - 1023 + eadj;
    const uint64_t mb = (b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
    double m;
    memcpy(&m, &mb, 8);
    if (m > OM_SQRT2) {
        m = m * 0.5;
        e += 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double R = 0x1.8618618618618p-4;           /* 2/21 */
    R = R * z + 0x1.af286bca1af28p-4;          /* 2/19 */
    R = R * z + 0x1.e1e1e1e1e1e1ep-4;          /* 2/17 */
    R = R * z + 0x1.1111111111111p-3;          /* 2/15 */
    R = R * z + 0x1.3b13b13b13b14p-3;          /* 2/13 */
    R = R * z + 0x1.745d1745d1746p-3;          /* 2/11 */
    R = R * z + 0x1.c71c71c71c71cp-3;          /* 2/9  */
    R = R * z + 0x1.2492492492492p-2;          /* 2/7  */
    R = R * z + 0x1.999999999999ap-2;          /* 2/5  */
    R = R * z + 0x1.5555555555555p-1;          /* 2/3  */
    R = R * z;
    const double hfsq = (0.5 * f) * f;
    const double ef = (double)e;
    return ef * OM_LN2_HI + (f - (hfsq - (s * (hfsq + R) + ef * OM_LN2_LO)));
}

static inline double om_pow(double x, double y) { return om_exp(y * om_log(x)); }

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
 * SC'11): 10 rounds of the 4x32 multiply-xor round with key schedule Weyl increments. */
static inline void om_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

#endif
