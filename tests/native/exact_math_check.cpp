// Host-side check of the device math replicas (twins_amd/csrc/exact_math.h) against what ns-3 itself
// computes on x86-64: libm's pow (glibc's FMA variant on an FMA host) and GetSeconds summed in x87
// long double. Built with g++ -ffp-contract=off; exits non-zero on any bit difference.
#include <cmath>
#include <cstdio>
#include <random>
#include "../../twins_amd/csrc/exact_math.h"

static double seconds_x87(int64_t ps)
{
    typedef unsigned __int128 u128;
    const u128 inv = ~(u128)0 / 1000000000000ull + 1;
    bool neg = ps < 0;
    u128 a = neg ? (u128)(-(__int128)ps) : (u128)ps;
    u128 v = a * (uint64_t)(inv >> 64) + ((a * (uint64_t)inv) >> 64);
    long double r = (long double)(uint64_t)(v >> 64);
    r += (long double)(uint64_t)v / 18446744073709551616.0L;
    return (double)(neg ? -r : r);
}

int main(int argc, char **argv)
{
    long n = argc > 1 ? atol(argv[1]) : 2000000;
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> u(0, 1);
    long bad = 0, tot = 0;
    auto cmp = [&](double x, double y) {
        double a = tw_pow(x, y), b = std::pow(x, y);
        tot++;
        if (memcmp(&a, &b, 8) != 0 && !(std::isnan(a) && std::isnan(b)))
        {
            if (bad < 10)
                printf("pow(%a, %a): %a vs libm %a\n", x, y, a, b);
            bad++;
        }
    };
    // Cubic's calls: K = cbrt((W_max - W) / C) and (t - K)^3, t and K in seconds
    for (long i = 0; i < n; i++)
    {
        cmp((double)(uint32_t)(rng() % 100000 + 1) / 0.4, 1 / 3.);
        cmp(u(rng) * 20.0, 3);
        cmp(std::ldexp(u(rng), -(int)(rng() % 40)), 3);
    }
    // general arguments
    for (long i = 0; i < n / 4; i++)
    {
        double x = std::ldexp(u(rng) + 0.5, (int)(rng() % 200) - 100), y = (u(rng) - 0.5) * 40;
        cmp(x, y);
        cmp(-x, std::floor(y));
    }
    // results near and below the smallest normal (the underflow path)
    for (long i = 0; i < n / 4; i++)
    {
        double x = std::ldexp(u(rng) + 0.5, -(int)(rng() % 60) - 40), y = 10 + u(rng) * 12;
        cmp(x, y);
        cmp(-x, std::floor(y));
    }
    double sp[] = {0.0, -0.0, 1.0, -1.0, INFINITY, -INFINITY, NAN, 2.0, 0.5, 1e-310, 1e308};
    for (double x : sp)
        for (double y : sp)
            cmp(x, y);
    long sbad = 0;
    std::uniform_int_distribution<int64_t> ti(-20000000000000ll, 20000000000000ll);
    for (long i = 0; i < n; i++)
    {
        int64_t ps = i < 1000000 ? (int64_t)i : ti(rng);
        double a = tw_seconds(ps), b = seconds_x87(ps);
        if (memcmp(&a, &b, 8) != 0)
        {
            if (sbad < 10)
                printf("seconds(%lld): %a vs x87 %a\n", (long long)ps, a, b);
            sbad++;
        }
    }
    printf("pow: %ld of %ld differ; seconds: %ld of %ld differ\n", bad, tot, sbad, n);
    return bad || sbad;
}
