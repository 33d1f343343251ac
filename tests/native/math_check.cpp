// Host checks of arithmetic helpers of the device code:
//  - swg::cdiv (swg_math.h) against IEEE division: random quotients over the
//    normal range for every constant divisor the drift uses, bit for bit;
//  - the compile-time table of RSHELL powers (swg_level.h) against the same
//    powers evaluated at run time;
//  - swg::pow_third (the cube roots of the hyperbolic first guess) against
//    x**THIRD in 80-bit long double: the double nearest to powl's value,
//    wherever powl's value is not within 2^-8 ulp of a rounding midpoint
//    (there an 80-bit evaluation cannot decide the rounding); outside
//    [2^-1000, 2^1000] it must equal pow(x, THIRD).
// Prints "ok <count>" or the first mismatches.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../swiftest_amd/csrc/swg_level.h"

static uint64_t st = 0x9E3779B97F4A7C15ull;
static uint64_t next()
{
    st ^= st << 13;
    st ^= st >> 7;
    st ^= st << 17;
    return st;
}

int main(int argc, char **argv)
{
    const long n = argc > 1 ? atol(argv[1]) : 2000000;
    const double cs[] = {182, 132, 90, 56, 30, 12, 210, 156, 110, 72, 42, 20, 6, 3, 39916800.0};
    long bad = 0, tot = 0;
    for (double c : cs) {
        const double r = 1.0 / c;
        for (long i = 0; i < n; i++) {
            const uint64_t u = next();
            const int ex = (int)(next() % 1960) - 980;
            double x = std::ldexp(1.0 + (double)(u >> 12) * 0x1p-52, ex);
            if (i % 4 == 0) x = (double)(int64_t)(next() % 1000000007) * std::ldexp(1.0, ex / 4);
            if (u & 1) x = -x;
            const double a = x / c, b = swg::cdiv(x, c, r);
            uint64_t ia, ib;
            std::memcpy(&ia, &a, 8);
            std::memcpy(&ib, &b, 8);
            tot++;
            if (ia != ib && bad++ < 5) std::printf("mismatch c=%g x=%a ieee=%a cdiv=%a\n", c, x, a, b);
        }
    }
    constexpr swg::ShellPowers tab{};
    for (int k = 0; k < swg::MAXLEV + 2; k++) {
        volatile double base = swg::RSHELL;  // run-time evaluation
        volatile int kk = k;
        const double sq = swg::powi(base, 2 * kk);
        double r = 1.0;
        for (int i = 1; i <= kk; i++) r = r * base;
        uint64_t a, b, c, e;
        std::memcpy(&a, &sq, 8);
        std::memcpy(&b, &tab.sq[k], 8);
        std::memcpy(&c, &r, 8);
        std::memcpy(&e, &tab.r[k], 8);
        tot += 2;
        if (a != b && bad++ < 5) std::printf("mismatch powi(RSHELL, %d): %a vs table %a\n", 2 * k, sq, tab.sq[k]);
        if (c != e && bad++ < 5) std::printf("mismatch RSHELL^%d: %a vs table %a\n", k, r, tab.r[k]);
    }
    long skipped = 0;
    for (long i = 0; i < 4 * n; i++) {
        const int ex = (i & 1) ? (int)(next() % 2000) - 1000 : (int)(next() % 40) - 20;
        const double x = std::ldexp(1.0 + (double)(next() >> 12) * 0x1p-52, ex);
        const long double L = powl((long double)x, (long double)swg::THIRD);
        const double yr = (double)L;
        const long double half = (long double)(std::nextafter(yr, INFINITY) - yr) / 2;
        if (half - fabsl(L - (long double)yr) < half * 0x1p-8L) {
            skipped++;
            continue;
        }
        const double y = swg::pow_third(x);
        tot++;
        if (y != yr && bad++ < 5) std::printf("mismatch pow_third(%a) = %a, nearest %a\n", x, y, yr);
    }
    // exact cubes with |ln x| small enough that x**THIRD rounds to the cube root
    const double cubes[][2] = {{1.0, 1.0}, {8.0, 2.0}, {27.0, 3.0}, {0.125, 0.5}};
    for (const auto &c : cubes) {
        const double y = swg::pow_third(c[0]);
        tot++;
        if (y != c[1] && bad++ < 5) std::printf("mismatch pow_third(%a) = %a, expected %a\n", c[0], y, c[1]);
    }
    const double edge[] = {0.0, 0x1p-1074, 0x1p-1001, 0x1p1001, INFINITY, NAN};
    for (double x : edge) {
        const double y = swg::pow_third(x);
        const double g = std::pow(x, swg::THIRD);
        uint64_t a, b;
        std::memcpy(&a, &y, 8);
        std::memcpy(&b, &g, 8);
        tot++;
        if (a != b && bad++ < 5) std::printf("mismatch pow_third(%a) = %a, pow %a\n", x, y, g);
    }
    if (skipped > n / 64) {
        std::printf("too many near-midpoint cases skipped: %ld\n", skipped);
        return 1;
    }
    if (bad) {
        std::printf("bad %ld of %ld\n", bad, tot);
        return 1;
    }
    std::printf("ok %ld\n", tot);
    return 0;
}
