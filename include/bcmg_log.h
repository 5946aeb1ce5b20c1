/*
 * bcmg_log.h - portable natural logarithm used for the matching weights.
 *
 * The matching weight of an off-diagonal entry is W = log(|AH| / (0.999 m))
 * (reference: src/AMG/matching.cu:386-396). The weights decide which edges are
 * matched, so the host restatement and the gfx950 kernels must produce the
 * SAME double for the same input. Neither glibc's log nor the device math
 * library give that guarantee across the two sides, so both call this one
 * function, which uses only IEEE-754 add, sub, mul and div (correctly rounded
 * on x86-64 SSE2 and on CDNA) in a fixed order. Translation units that use it
 * must be compiled without floating-point contraction (-ffp-contract=off), or
 * the pragma below takes care of it for clang.
 *
 * Method: the argument reduction and recombination of Sun's fdlibm e_log.c
 * (x = 2^k (1+f), 1+f in (sqrt(2)/2, sqrt(2)], s = f/(2+f),
 *  log(1+f) = f - hfsq + s (hfsq + R), hfsq = f^2/2), with R evaluated from the
 * Taylor series of 2 atanh(s) - 2 s = s * sum_{k>=1} 2 s^(2k) / (2k+1),
 * truncated after k = 11 (|s| <= 0.1716, truncation < 2^-58 relative).
 * ln 2 is split into a 31-bit head (so k * head is exact) and a tail.
 * Accuracy is checked against libm in the CPU tests (<= 1 ulp).
 */
#ifndef BCMG_LOG_H
#define BCMG_LOG_H

#include <stdint.h>

#if defined(__HIPCC__)
#define BCMG_HD __host__ __device__
#else
#define BCMG_HD
#endif

#if defined(__clang__)
#pragma clang fp contract(off)
#endif

static inline BCMG_HD double bcmg_log(double x)
{
    const double ln2_hi = 0x1.62e42fee00000p-1;
    const double ln2_lo = 0x1.a39ef35793c76p-33;
    const double sqrt2 = 0x1.6a09e667f3bcdp+0;
    /* c_k = 2 / (2k + 1), k = 1..11 */
    const double c1 = 0x1.5555555555555p-1;
    const double c2 = 0x1.999999999999ap-2;
    const double c3 = 0x1.2492492492492p-2;
    const double c4 = 0x1.c71c71c71c71cp-3;
    const double c5 = 0x1.745d1745d1746p-3;
    const double c6 = 0x1.3b13b13b13b14p-3;
    const double c7 = 0x1.1111111111111p-3;
    const double c8 = 0x1.e1e1e1e1e1e1ep-4;
    const double c9 = 0x1.af286bca1af28p-4;
    const double c10 = 0x1.8618618618618p-4;
    const double c11 = 0x1.642c8590b2164p-4;

    if (!(x > 0.0)) {
        if (x == 0.0) return -__builtin_inf();
        return __builtin_nan("");          /* negative or NaN */
    }
    if (x == __builtin_inf()) return x;

    uint64_t bits = __builtin_bit_cast(uint64_t, x);
    int k = 0;
    if ((bits >> 52) == 0) {               /* subnormal: scale by 2^54 */
        x = x * 0x1p54;
        bits = __builtin_bit_cast(uint64_t, x);
        k = -54;
    }
    k += (int)(bits >> 52) - 1023;
    bits = (bits & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull;
    double m = __builtin_bit_cast(double, bits);   /* m in [1, 2) */
    if (m > sqrt2) {
        m = m * 0.5;
        k += 1;
    }
    const double f = m - 1.0;              /* exact (Sterbenz) */
    const double s = f / (2.0 + f);
    const double z = s * s;
    double R = c11;
    R = c10 + z * R;
    R = c9 + z * R;
    R = c8 + z * R;
    R = c7 + z * R;
    R = c6 + z * R;
    R = c5 + z * R;
    R = c4 + z * R;
    R = c3 + z * R;
    R = c2 + z * R;
    R = c1 + z * R;
    R = z * R;
    const double hfsq = 0.5 * (f * f);
    const double dk = (double)k;
    return dk * ln2_hi - ((hfsq - (s * (hfsq + R) + dk * ln2_lo)) - f);
}

#endif /* BCMG_LOG_H */
