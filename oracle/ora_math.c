/*
 * ora_math.c -- exp, log and Philox4x32-10 for the oracle (test infrastructure).
 *
 * The reference calls libm exp/log (HH2_test.h:265-276, HHI_buffer.h:173-180,
 * HH2_test.cpp:861-864, HHI_buffer.cpp:739-740, NGS:16629/16647,
 * poisson_noise.cpp:14).  libm's results are not specified bit for bit, so the
 * oracle and the GPU path both use the functions below, written with only
 * IEEE-754 +, -, *, /, fma and rint, which round identically on x86-64 and
 * gfx950.  The same functions are linked into oracle/_ref in place of libm's
 * when the reference classes are checked bit for bit.
 *
 * exp: table method (Tang, ACM TOMS 15(2), 1989): x = k*ln2/64 + r,
 *      exp(x) = 2^(k/64) * (1 + p(r)), |r| <= ln2/128, 2^(j/64) as a
 *      double-double table, p the degree-6 Taylor polynomial, scaling by ldexp
 *      (IEEE scaleB: a single rounding, also into the subnormal range).
 *      Error ~0.505 ulp.
 * log: log(1+f) = f - (hfsq - s*(hfsq + R(s^2))), s = f/(2+f), the reduction of
 *      the fdlibm log (Sun Microsystems, 1993); R is the Taylor series of
 *      2*atanh(s)/s - 2 in z = s^2 to degree 11.
 * Philox4x32-10: Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
 *      as 1, 2, 3", SC'11.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "hvc_oracle.h"
#include "ora_exp2_table.h"

static inline uint64_t d2u(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
static inline double u2d(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }

double ora_exp(double x)
{
    if (x != x) return x;                         /* NaN */
    if (x > 710.0) x = 710.0;                     /* beyond: overflow to inf / underflow to 0 */
    if (x < -746.0) x = -746.0;
    double kd = rint(x * ORA_INV_LN2_64);
    int k = (int)kd;
    double r = fma(-kd, ORA_LN2_64_HI, x);
    r = fma(-kd, ORA_LN2_64_LO, r);
    double z = r * r;
    /* p(r) = r + r^2/2 + r^3/6 + r^4/24 + r^5/120 + r^6/720 */
    double p = fma(r, 1.0 / 720.0, 1.0 / 120.0);
    p = fma(r, p, 1.0 / 24.0);
    p = fma(r, p, 1.0 / 6.0);
    p = fma(r, p, 0.5);
    p = fma(z, p, r);
    int j = k & 63;
    int e = (k - j) / 64;                          /* floor(k / 64) */
    double th = ORA_EXP2_T[2 * j];
    double tl = ORA_EXP2_T[2 * j + 1];
    double y = th + fma(th, p, tl);               /* y in [1/e^(1/128), 2*e^(1/128)) */
    if (e >= -1021 && e <= 1023)                   /* normal or overflowing result: y * 2^e exact */
        return y * u2d((uint64_t)(e + 1023) << 52);
    return ldexp(y, e);                            /* one rounding, also for subnormal results */
}

double ora_log(double x)
{
    if (x != x) return x;
    if (x < 0.0) return u2d(0x7ff8000000000000ull);
    if (x == 0.0) return u2d(0xfff0000000000000ull);
    if (x == u2d(0x7ff0000000000000ull)) return x;
    int e = 0;
    uint64_t u = d2u(x);
    if ((u >> 52) == 0) {                          /* subnormal: scale by 2^54 */
        x = x * 18014398509481984.0;
        u = d2u(x);
        e = -54;
    }
    e += (int)(u >> 52) - 1023;
    double m = u2d((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull); /* [1,2) */
    if (m > 1.4142135623730951) {                 /* m in [sqrt(1/2), sqrt(2)] */
        m = m * 0.5;
        e += 1;
    }
    double f = m - 1.0;                            /* exact (Sterbenz) */
    double s = f / (2.0 + f);
    double z = s * s;
    /* R(z) = sum_{k=1..11} 2 z^k / (2k+1) */
    double R = 2.0 / 23.0;
    R = fma(z, R, 2.0 / 21.0);
    R = fma(z, R, 2.0 / 19.0);
    R = fma(z, R, 2.0 / 17.0);
    R = fma(z, R, 2.0 / 15.0);
    R = fma(z, R, 2.0 / 13.0);
    R = fma(z, R, 2.0 / 11.0);
    R = fma(z, R, 2.0 / 9.0);
    R = fma(z, R, 2.0 / 7.0);
    R = fma(z, R, 2.0 / 5.0);
    R = fma(z, R, 2.0 / 3.0);
    R = z * R;
    double hfsq = 0.5 * f * f;
    double ed = (double)e;
    double lo = fma(s, hfsq + R, ed * ORA_LN2_LO);
    return fma(ed, ORA_LN2_HI, f - (hfsq - lo));
}

void ora_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        if (round > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void ora_white_indices(uint32_t seed_lo, uint32_t seed_hi, uint32_t trial,
                       uint32_t neuron, uint32_t step, int idx[4])
{
    uint32_t ctr[4] = {step, neuron, trial, ORA_STREAM_RA_WHITE};
    uint32_t key[2] = {seed_lo, seed_hi};
    uint32_t o[4];
    ora_philox4x32(ctr, key, o);
    for (int q = 0; q < 4; ++q) idx[q] = (int)(((uint64_t)o[q] * 30u) >> 32);
}

uint32_t ora_poisson_u32(uint32_t seed_lo, uint32_t seed_hi, uint32_t trial,
                         uint32_t stream, uint32_t neuron, uint32_t draw)
{
    uint32_t ctr[4] = {draw >> 2, neuron, trial, stream};
    uint32_t key[2] = {seed_lo, seed_hi};
    uint32_t o[4];
    ora_philox4x32(ctr, key, o);
    return o[draw & 3u];
}
