// hvc_math.hip.h -- device exp/log and counter-based noise for the HVC engine.
//
// The reference evaluates libm exp/log (HH2_test.h:265-276, HHI_buffer.h:173-180,
// poisson_noise.cpp:14).  The engine defines its transcendental functions with
// IEEE-754 +, -, *, /, fma and rint only, so every result is reproducible bit for
// bit on any IEEE machine:
//   exp: x = k*ln2/64 + r, exp(x) = 2^(k/64) (1 + p(r)), |r| <= ln2/128, with a
//        double-double table of 2^(j/64) and p the degree-6 Taylor polynomial
//        (Tang's table method, ACM TOMS 15(2), 1989), scaled with ldexp;
//        error ~0.505 ulp.
//   log: log(1+f) = f - (f^2/2 - s (f^2/2 + R(s^2))), s = f/(2+f) (the fdlibm
//        reduction), R the Taylor series of 2 atanh(s)/s - 2 to degree 11.
// Noise: Philox4x32-10 (Salmon et al., SC'11).  One call per (RA neuron, step)
// gives the four Debraband point-distribution indices (HH2_test.cpp:640-643);
// Poisson draws take one 32-bit word each (poisson_noise.cpp:9-21).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hvc_exp2_table.h"

namespace hvc {

static __constant__ double kExp2Table[128] = HVC_EXP2_TABLE_INIT;

constexpr uint32_t kStreamRAWhite = 0x52410000u;
constexpr uint32_t kStreamIPoisson = 0x49000000u;
constexpr uint32_t kStreamRAPoisson = 0x52500000u;

__device__ __forceinline__ double u2d(uint64_t u) { return __longlong_as_double((long long)u); }
__device__ __forceinline__ uint64_t d2u(double x) { return (uint64_t)__double_as_longlong(x); }

__device__ __forceinline__ double dexp(double x)
{
    // NaN passes through the comparisons, the reduction and the polynomial
    x = x > 710.0 ? 710.0 : x;
    x = x < -746.0 ? -746.0 : x;
    double kd = __builtin_rint(x * HVC_INV_LN2_64);
    int k = (int)kd;
    double r = __builtin_fma(-kd, HVC_LN2_64_HI, x);
    r = __builtin_fma(-kd, HVC_LN2_64_LO, r);
    double z = r * r;
    double p = __builtin_fma(r, 1.0 / 720.0, 1.0 / 120.0);
    p = __builtin_fma(r, p, 1.0 / 24.0);
    p = __builtin_fma(r, p, 1.0 / 6.0);
    p = __builtin_fma(r, p, 0.5);
    p = __builtin_fma(z, p, r);
    int j = k & 63;
    int e = k >> 6;  // floor(k / 64)
    double th = kExp2Table[2 * j];
    double tl = kExp2Table[2 * j + 1];
    double y = th + __builtin_fma(th, p, tl);
    return __builtin_ldexp(y, e);  // v_ldexp_f64: one rounding, also for subnormal results
}

// a / B for a divisor B > 0 known at compile time, with Y = RN(1/B): q = RN(a Y),
// r = a - B q (exact, fma), result RN(q + r Y) -- Markstein's correction
// (Markstein 1990, IBM J. Res. Dev. 34(1)).  For 2^-900 <= |q| < 2^1000:
// |q - a/B| < 1.5 ulp, r is exact, and the fma result differs from a/B by at
// most 3 * 2^-54 ulp while a/B lies at least ulp / (4 B_odd) from every
// rounding boundary (B_odd = odd part of B's significand); for B_odd < 2^51 --
// every divisor used here -- the result is the correctly rounded quotient, bit
// for bit a / B.  For a = +-0 the result is a zero with a's sign, as a / B is.
// Any other quotient (tiny, huge, infinite or NaN) is recorded in the guard and
// the caller repeats the step with IEEE division.  No branch, so the compiler
// can interleave the many independent quotients of a stage.
struct DivGuard {
    int emin = 0;      // least binary exponent of a quotient (frexp convention, 0 for zero)
    double big = 0.0;  // sum of |q| 2^-1000: >= 1 or NaN once a quotient is huge or not finite
    __device__ __forceinline__ void note(double q)
    {
        emin = min(emin, __builtin_amdgcn_frexp_exp(q));
        big = __builtin_fma(__builtin_fabs(q), 0x1p-1000, big);
    }
    // a value k that enters tableau terms (n k) / d with n / d in [2^-7, 2^5]
    // and |n| < 2^36 only (DP8 stages 7-12 and the final combination): noting
    // k once stands for noting each such quotient.  k = 0 or 2^-892 <= |k| <
    // 2^985 keeps n k finite and every quotient in [2^-899, 2^990) or zero.
    __device__ __forceinline__ void note_k(double k)
    {
        emin = min(emin, __builtin_amdgcn_frexp_exp(k) - 8);
        big = __builtin_fma(__builtin_fabs(k), 0x1p-985, big);
    }
    __device__ __forceinline__ bool failed() const { return emin < -899 || !(big < 1.0); }
};

// Markstein's correction with the rounded reciprocal Y = RN(1 / B)
__device__ __forceinline__ double cdiv_plain(double a, double B, double Y)
{
    double q = a * Y;
    double r = __builtin_fma(-q, B, a);
    return __builtin_copysign(__builtin_fma(r, Y, q), a);
}

__device__ __forceinline__ double cdiv_fast(double a, double B, double Y, DivGuard& g)
{
    g.note(a * Y);
    return cdiv_plain(a, B, Y);
}

// the same quotient with the IEEE fallback applied per value
__device__ __forceinline__ double cdiv_checked(double a, double B)
{
    DivGuard g;
    double q = cdiv_fast(a, B, 1.0 / B, g);
    return g.failed() ? a / B : q;
}
#define HVC_CDIV(a, B) ::hvc::cdiv_checked((a), (double)(B))

__device__ __forceinline__ double dlog(double x)
{
    if (x != x) return x;
    if (x < 0.0) return u2d(0x7ff8000000000000ull);
    if (x == 0.0) return u2d(0xfff0000000000000ull);
    if (x == u2d(0x7ff0000000000000ull)) return x;
    int e = 0;
    uint64_t u = d2u(x);
    if ((u >> 52) == 0) {
        x = x * 18014398509481984.0;
        u = d2u(x);
        e = -54;
    }
    e += (int)(u >> 52) - 1023;
    double m = u2d((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e += 1;
    }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double R = 2.0 / 23.0;
    R = __builtin_fma(z, R, 2.0 / 21.0);
    R = __builtin_fma(z, R, 2.0 / 19.0);
    R = __builtin_fma(z, R, 2.0 / 17.0);
    R = __builtin_fma(z, R, 2.0 / 15.0);
    R = __builtin_fma(z, R, 2.0 / 13.0);
    R = __builtin_fma(z, R, 2.0 / 11.0);
    R = __builtin_fma(z, R, 2.0 / 9.0);
    R = __builtin_fma(z, R, 2.0 / 7.0);
    R = __builtin_fma(z, R, 2.0 / 5.0);
    R = __builtin_fma(z, R, 2.0 / 3.0);
    R = z * R;
    double hfsq = 0.5 * f * f;
    double ed = (double)e;
    double lo = __builtin_fma(s, hfsq + R, ed * HVC_LN2_LO);
    return __builtin_fma(ed, HVC_LN2_HI, f - (hfsq - lo));
}

struct U4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox4x32(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                         uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        if (round > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint32_t lo0 = 0xD2511F53u * c0;
        uint32_t hi0 = __umulhi(0xD2511F53u, c0);
        uint32_t lo1 = 0xCD9E8D57u * c2;
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c2);
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    return U4{c0, c1, c2, c3};
}

__device__ __forceinline__ int point_index(uint32_t w) { return (int)(((uint64_t)w * 30u) >> 32); }

__device__ __forceinline__ uint32_t poisson_u32(uint32_t seed_lo, uint32_t seed_hi, uint32_t trial,
                                                uint32_t stream, uint32_t neuron, uint32_t draw)
{
    U4 o = philox4x32(draw >> 2, neuron, trial, stream, seed_lo, seed_hi);
    uint32_t q = draw & 3u;
    return q == 0 ? o.x : (q == 1 ? o.y : (q == 2 ? o.z : o.w));
}

}  // namespace hvc
