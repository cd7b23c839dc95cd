// Correctly rounded double quotients without a division instruction
// sequence (long on the device): x / 6 for the cubic B-spline weights of
// src/filtered_noise.h:33-61 (seven per sample), and a / b by the grid and
// line spacings. Host and device share these definitions
// (tests/cpp/div_selftest.cpp checks them against the division on the host).
#pragma once

#if defined(__HIPCC__)
#define QEW_HD __host__ __device__
#else
#define QEW_HD
#endif

namespace qew_div {

// q0 = RN(x RN(1/6)) is within one ulp of x / 6, the remainder x - 6 q0 is
// exact in one fma, and the correction q0 + e RN(1/6), rounded once, is
// RN(x / 6) whenever that quotient is normal or zero (Markstein's correction
// step; in the normal range x / 6 is never a rounding tie, since x is a
// multiple of 4 or 8 units of the quotient's last place). The sign is
// restored for x = -0.
QEW_HD inline double div6(double x) {
    constexpr double r = 1.0 / 6.0;
    const double q = x * r;
    const double e = __builtin_fma(-q, 6.0, x);
    return __builtin_copysign(__builtin_fma(e, r, q), x);
}

// div6 for numerators that can be tiny: a subnormal x / 6 can be an exact
// tie, x - 6 q = +-3 2^-1074, which div6 rounds toward zero; when q is then
// odd, the even neighbour one unit further from zero is the rounded quotient.
QEW_HD inline double div6_sub(double x) {
    const double q = div6(x);
    const double e = __builtin_fma(-q, 6.0, x);
    const long long b = __builtin_bit_cast(long long, q);
    return (__builtin_fabs(e) == 0x3p-1074 && (b & 1)) ? __builtin_bit_cast(double, b + 1) : q;
}

// a / b for b > 0 given rb = RN(1 / b): the same correction step, exact for
// b = 1 (rb = 1, no correction) and correctly rounded for any b whenever the
// quotient is normal or zero (Markstein's theorem: rb within half an ulp of
// 1 / b and a rb within one ulp of a / b).
QEW_HD inline double divp(double a, double b, double rb) {
    const double q = a * rb;
    const double e = __builtin_fma(-q, b, a);
    return __builtin_copysign(__builtin_fma(e, rb, q), a);
}

}  // namespace qew_div
