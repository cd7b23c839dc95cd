// The division-free quotients of the device sampler (qew_div::div6 for
// numerators whose quotient is normal or zero, qew_div::div6_sub for any
// numerator, qew_div::divp by a spacing) against the correctly rounded
// division, on the host:
// random significands in every binade from the subnormals up, both signs;
// every subnormal multiple of 2^-1074 below 2^20 units; the seven weight
// numerators of src/filtered_noise.h:33-61 at random and at tiny t; divp
// over random quotients of normal magnitude.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../qew_amd/csrc/qew_div.hpp"

static uint64_t s = 88172645463325252ull;
static uint64_t rnd() {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return s;
}
static long n = 0, bad_any = 0, bad_normal = 0;

// div6_sub must equal the division everywhere; div6 wherever the quotient
// is normal or zero
static void check(double x) {
    const double a = x / 6.0, b = qew_div::div6(x), c = qew_div::div6_sub(x);
    ++n;
    if (std::memcmp(&a, &c, 8) != 0) {
        if (bad_any < 5) std::printf("div6_sub(%a) = %a, x / 6 = %a\n", x, c, a);
        ++bad_any;
    }
    if ((a == 0.0 || std::fabs(a) >= 0x1p-1022) && std::memcmp(&a, &b, 8) != 0) {
        if (bad_normal < 5) std::printf("div6(%a) = %a, x / 6 = %a\n", x, b, a);
        ++bad_normal;
    }
}

int main() {
    for (int e = -1074; e < 8; ++e)
        for (int k = 0; k < 1000; ++k) {
            const uint64_t m = rnd();
            double x;
            if (e < -1022)
                x = std::ldexp((double)(m >> (64 - (e + 1075))), -1074);
            else
                x = std::ldexp(1.0 + (double)(m >> 12) * 0x1p-52, e);
            check(x);
            check(-x);
        }
    for (long m = 0; m < (1L << 20); ++m) check(std::ldexp((double)m, -1074));
    for (long k = 0; k < 1000000; ++k) {
        double t = (double)(rnd() >> 11) * 0x1p-53;
        if (k % 4 == 1) t = std::ldexp(1.0 + (double)(rnd() >> 12) * 0x1p-52, -(int)(rnd() % 1074));
        const double t2 = t * t, t3 = t2 * t, omt = 1.0 - t;
        check(omt * omt * omt);
        check(4.0 - 6.0 * t2 + 3.0 * t3);
        check(1.0 + 3.0 * t + 3.0 * t2 - 3.0 * t3);
        check(t3);
        check(-12.0 * t + 9.0 * t2);
        check(3.0 + 6.0 * t - 9.0 * t2);
        check(3.0 * t2);
    }
    check(0.0);
    check(-0.0);
    // divp: a / b from RN(1 / b) for spacings b > 0, correctly rounded when
    // the quotient is normal or zero (b = 1 exactly, random b, b with an
    // all-ones significand)
    for (long k = 0; k < 2000000; ++k) {
        double b = std::ldexp(1.0 + (double)(rnd() >> 12) * 0x1p-52, (int)(rnd() % 41) - 20);
        if (k % 7 == 0) {
            uint64_t bits;
            std::memcpy(&bits, &b, 8);
            bits |= 0x000FFFFFFFFFFFFFull;
            std::memcpy(&b, &bits, 8);
        }
        if (k % 11 == 0) b = 1.0;
        double a = std::ldexp(1.0 + (double)(rnd() >> 12) * 0x1p-52, (int)(rnd() % 81) - 40);
        if (k % 3 == 0) a = (double)(int64_t)(rnd() >> 40) * std::ldexp(1.0, -(int)(rnd() % 30));
        if (k & 1) a = -a;
        const double d = a / b, q = qew_div::divp(a, b, 1.0 / b);
        ++n;
        if (std::memcmp(&d, &q, 8) != 0) {
            if (bad_normal < 5) std::printf("divp(%a, %a) = %a, a / b = %a\n", a, b, q, d);
            ++bad_normal;
        }
    }
    std::printf("%ld values: %ld div6_sub mismatches, %ld div6 mismatches with a normal quotient\n", n,
                bad_any, bad_normal);
    return (bad_any || bad_normal) ? 1 : 0;
}
