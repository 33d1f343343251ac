// Host checks of arithmetic helpers of the device code:
//  - swg::cdiv (swg_math.h) against IEEE division: random quotients over the
//    normal range for every constant divisor the drift uses, bit for bit;
//  - the compile-time table of RSHELL powers (swg_level.h) against the same
//    powers evaluated at run time.
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
    if (bad) {
        std::printf("bad %ld of %ld\n", bad, tot);
        return 1;
    }
    std::printf("ok %ld\n", tot);
    return 0;
}
