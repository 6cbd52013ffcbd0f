// exact_div_check.hip -- host program: cgli::div_r against the IEEE
// division, bit for bit, on random and structured operands.  Built and run
// by test_exact_div.py.
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "cgli_internal.hpp"

static uint64_t splitmix(uint64_t &s)
{
    uint64_t z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// a double with a random sign and mantissa and an exponent in [emin, emax]
static double draw(uint64_t &s, int emin, int emax)
{
    const uint64_t b = splitmix(s);
    const uint64_t e = (uint64_t)(emin + (int)((b >> 53) % (uint64_t)(emax - emin + 1)) + 1023);
    const uint64_t bits = (b & 0x000fffffffffffffULL) | (e << 52) | ((b >> 52) & 1ULL) << 63;
    double x;
    std::memcpy(&x, &bits, 8);
    return x;
}

static long bad = 0, total = 0;
static bool plain = false;   // the uncorrected product x (1/d) instead: must fail
static void check(double x, double d)
{
    const double a = plain ? x * (1.0 / d) : cgli::div_r(x, d, 1.0 / d), b = x / d;
    total++;
    if (std::memcmp(&a, &b, 8) != 0) {
        if (bad < 5 && !plain) std::printf("mismatch x=%a d=%a div_r=%a div=%a\n", x, d, a, b);
        bad++;
    }
}
static void checkf(float x, float d)
{
    const float a = plain ? x * (1.0f / d) : cgli::div_r(x, d, 1.0f / d), b = x / d;
    total++;
    if (std::memcmp(&a, &b, 4) != 0) {
        if (bad < 5 && !plain) std::printf("mismatch (float) x=%a d=%a div_r=%a div=%a\n", x, d, a, b);
        bad++;
    }
}
// a float with a random sign and mantissa and an exponent in [emin, emax]
static float drawf(uint64_t &s, int emin, int emax)
{
    const uint64_t b = splitmix(s);
    const uint32_t e = (uint32_t)(emin + (int)((b >> 40) % (uint64_t)(emax - emin + 1)) + 127);
    const uint32_t bits = (uint32_t)(b & 0x7fffffULL) | (e << 23) | (uint32_t)((b >> 32) & 1ULL) << 31;
    float x;
    std::memcpy(&x, &bits, 4);
    return x;
}

int main(int argc, char **argv)
{
    const long n = argc > 1 ? std::atol(argv[1]) : 1000000;
    plain = argc > 2 && std::strcmp(argv[2], "plain") == 0;
    uint64_t s = 20240611;
    for (long t = 0; t < n; t++) check(draw(s, -250, 250), draw(s, -250, 250));
    // operands of one binade (the quotient near 1, where rounding is tight)
    for (long t = 0; t < n; t++) check(draw(s, 0, 0), draw(s, 0, 0));
    // divisors a few ulps from a power of two, and exact quotients
    for (int e = -60; e <= 60; e++)
        for (int k = -64; k <= 64; k++) {
            const double d = std::ldexp(1.0, e) * (1.0 + k * 0x1p-52);
            for (int m = 1; m <= 64; m++) {
                check(d * m, d);
                check(draw(s, -8, 8), d);
            }
        }
    // binary32: random exponents over +-30, one binade
    for (long t = 0; t < n; t++) checkf(drawf(s, -30, 30), drawf(s, -30, 30));
    for (long t = 0; t < n; t++) checkf(drawf(s, 0, 0), drawf(s, 0, 0));
    std::printf("checked %ld mismatches %ld\n", total, bad);
    return bad != 0;
}
