// Host build of the tracer's sincos_near / exp_near / log_pos / rcp
// (blackhole_amd/csrc/kerr_dev.h) against glibc: prints, per function, the
// number of arguments and the largest distance in ulps.  The host build
// evaluates the same FMA sequences as the device (fma is exact on both; rcp
// is a correctly rounded division here, the device's is <= 1 ulp).
#include "../../blackhole_amd/csrc/kerr_dev.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>

static int64_t ulps(double a, double b) {
    if (a == b) return 0;
    int64_t ia, ib;
    std::memcpy(&ia, &a, 8);
    std::memcpy(&ib, &b, 8);
    if (ia < 0) ia = INT64_MIN - ia;
    if (ib < 0) ib = INT64_MIN - ib;
    return ia > ib ? ia - ib : ib - ia;
}

int main(int argc, char** argv) {
    const long n = argc > 1 ? std::atol(argv[1]) : 2000000;
    std::mt19937_64 g(12345);
    std::uniform_real_distribution<double> theta(-10.0, 10.0), wide(-1.0e5, 1.0e5), ex(-700.0, 700.0),
        small(-1e-3, 1e-3);
    int64_t ms = 0, mc = 0, me = 0, ml = 0;
    for (long i = 0; i < n; ++i) {
        const double x = (i % 4 == 0) ? wide(g) : (i % 4 == 1) ? small(g) : theta(g);   // random: ~1e-6 from a zero at best
        double s, c;
        bhx::sincos_near(x, &s, &c);
        ms = std::max(ms, ulps(s, std::sin(x)));
        mc = std::max(mc, ulps(c, std::cos(x)));
        const double y = ex(g);
        me = std::max(me, ulps(bhx::exp_near(y), std::exp(y)));
        // log over the normal range: random exponent and mantissa
        const double m = 1.0 + (g() >> 12) * 0x1p-52;
        const double z = std::ldexp(m, static_cast<int>(g() % 2040) - 1020);
        ml = std::max(ml, ulps(bhx::log_pos(z), std::log(z)));
    }
    // Next to the zeros (multiples of pi/2, and the pi/4 quadrant boundaries):
    // the two-term reduction drops k (pi/2 - hi - lo) ~ k 2e-33, so there the
    // bound is absolute: |ours - glibc| <= 1 ulp + |k| 4e-33.  Prints the
    // largest excess over that bound (0: within it).
    double excess = 0.0;
    for (int j = -256; j <= 256; ++j) {
        for (double d : {-1e-12, 0.0, 1e-12}) {
            const double x = j * 0.78539816339744830962 + d;
            const double k = std::rint(x * 0.63661977236758134308);
            double s, c;
            bhx::sincos_near(x, &s, &c);
            const double bs = std::fabs(std::nextafter(std::sin(x), 2.0) - std::sin(x)) + std::fabs(k) * 4e-33;
            const double bc = std::fabs(std::nextafter(std::cos(x), 2.0) - std::cos(x)) + std::fabs(k) * 4e-33;
            excess = std::max(excess, std::fabs(s - std::sin(x)) - bs);
            excess = std::max(excess, std::fabs(c - std::cos(x)) - bc);
        }
    }
    std::printf("sin %ld %lld\ncos %ld %lld\nexp %ld %lld\nlog %ld %lld\nboundary %d %d\n", n, (long long)ms, n,
                (long long)mc, n, (long long)me, n, (long long)ml, 513 * 3, excess > 0.0 ? 1 : 0);
    return 0;
}
