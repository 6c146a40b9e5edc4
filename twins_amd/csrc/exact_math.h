// Exact arithmetic of the ns-3 models, for gfx950 device code (and for host-side checks of the
// same source against the host's libm / x87):
//   tw_pow      glibc's pow (ARM optimized-routines' algorithm, glibc >= 2.28), as its FMA build
//               evaluates it: the contractions GCC makes under -mfma are written out as fma();
//               every other operation must stay unfused (build with -ffp-contract=off)
//   tw_seconds  ns3::Time::GetSeconds() at picosecond resolution: the Q64.64 product with
//               int64x64_t::Invert(10^12), summed in x87 extended precision, then rounded to double
//   tw_mrg_*    ns-3's MRG32k3a (RngStream::RandU01) and UniformRandomVariable::GetInteger
// Reference counterparts: src/nsmath.h:31-300 and src/rng.h:29-56.
#pragma once
#include <cstdint>
#include <cstring>

#if defined(__HIPCC__)
#define TW_HD __host__ __device__ inline
#else
#define TW_HD inline
#endif
#if defined(__HIP_DEVICE_COMPILE__)
#define TW_POW_TABLE static __constant__ const
#else
#define TW_POW_TABLE static const
#endif

#include "pow_tables.h"

namespace twm
{
TW_HD uint64_t as_u64(double f)
{
    uint64_t i;
    memcpy(&i, &f, 8);
    return i;
}
TW_HD double as_f64(uint64_t i)
{
    double f;
    memcpy(&f, &i, 8);
    return f;
}
TW_HD uint32_t top12(double x) { return (uint32_t)(as_u64(x) >> 52); }

// log(x) = k ln2 + log(c) + log1p(z/c - 1) as hi + *tail; x = 2^k z, z in [OFF, 2 OFF)
TW_HD double log_hi_lo(uint64_t ix, double *tail)
{
    const uint64_t OFF = 0x3fe6955500000000ull;
    uint64_t tmp = ix - OFF;
    int i = (int)((tmp >> 45) & 127);
    double kd = (double)(int)((int64_t)tmp >> 52);
    double z = as_f64(ix - (tmp & (0xfffull << 52)));
    double invc = LOG_TAB[i][0], logc = LOG_TAB[i][1], logctail = LOG_TAB[i][2];
    double r = __builtin_fma(z, invc, -1.0); // exact: 1/c has few bits
    double t1 = __builtin_fma(kd, LOG_LN2HI, logc);
    double t2 = t1 + r;
    double lo1 = __builtin_fma(kd, LOG_LN2LO, logctail);
    double lo2 = t1 - t2 + r;
    double ar = LOG_POLY[0] * r;
    double ar2 = r * ar;
    double ar3 = r * ar2;
    double hi = t2 + ar2;
    double lo3 = __builtin_fma(ar, r, -ar2);
    double lo4 = t2 - hi + ar2;
    double q = __builtin_fma(ar2, __builtin_fma(r, LOG_POLY[6], LOG_POLY[5]), __builtin_fma(r, LOG_POLY[4], LOG_POLY[3]));
    q = __builtin_fma(ar2, q, __builtin_fma(r, LOG_POLY[2], LOG_POLY[1]));
    double lo = __builtin_fma(ar3, q, lo1 + lo2 + lo3 + lo4);
    double y = hi + lo;
    *tail = hi - y + lo;
    return y;
}

// the scaling when 2^(k/N) leaves the normal range
TW_HD double exp_special(double tmp, uint64_t sbits, uint64_t ki)
{
    if ((ki & 0x80000000u) == 0)
    {
        double scale = as_f64(sbits - (1009ull << 52));
        return 0x1p1009 * __builtin_fma(scale, tmp, scale);
    }
    // (k < 0: glibc's FMA build leaves scale * tmp unfused here, as this host's libm shows;
    // Cubic's arguments never come this close to underflow)
    sbits += 1022ull << 52;
    double scale = as_f64(sbits);
    double st = scale * tmp;
    double y = scale + st;
    if ((y < 0 ? -y : y) < 1.0)
    {
        double one = y < 0.0 ? -1.0 : 1.0;
        double lo = scale - y + st;
        double hi = one + y;
        lo = one - hi + y + lo;
        y = (hi + lo) - one;
        if (y == 0.0)
            y = as_f64(sbits & 0x8000000000000000ull);
    }
    return 0x1p-1022 * y;
}

// sign * exp(x + xtail)
TW_HD double exp_hi_lo(double x, double xtail, uint32_t sign_bias)
{
    uint32_t abstop = top12(x) & 0x7ff;
    if (abstop - 0x3c9u >= 0x408u - 0x3c9u) // |x| < 2^-54 or |x| >= 512 (or inf/nan)
    {
        if (abstop - 0x3c9u >= 0x80000000u)
        {
            double one = 1.0 + x;
            return sign_bias ? -one : one;
        }
        if (abstop >= 0x409u) // |x| >= 1024
        {
            if (as_u64(x) >> 63)
                return sign_bias ? -0.0 : 0.0;
            double inf = as_f64(0x7ff0000000000000ull);
            return sign_bias ? -inf : inf;
        }
        abstop = 0;
    }
    double kd = __builtin_fma(EXP_INVLN2N, x, EXP_SHIFT);
    uint64_t ki = as_u64(kd);
    kd -= EXP_SHIFT;
    double r = __builtin_fma(kd, EXP_NEGLN2LON, __builtin_fma(kd, EXP_NEGLN2HIN, x));
    r += xtail;
    uint64_t idx = 2 * (ki & 127);
    uint64_t top = (ki + sign_bias) << 45;
    double tail = as_f64(EXP_TAB[idx]);
    uint64_t sbits = EXP_TAB[idx + 1] + top;
    double r2 = r * r;
    double tmp = __builtin_fma(r2 * r2, __builtin_fma(r, EXP_POLY[3], EXP_POLY[2]),
                               __builtin_fma(r2, __builtin_fma(r, EXP_POLY[1], EXP_POLY[0]), tail + r));
    if (abstop == 0)
        return exp_special(tmp, sbits, ki);
    double scale = as_f64(sbits);
    return __builtin_fma(scale, tmp, scale);
}

// 0: not an integer, 1: odd, 2: even
TW_HD int int_kind(uint64_t iy)
{
    int e = (int)(iy >> 52 & 0x7ff);
    if (e < 0x3ff)
        return 0;
    if (e > 0x3ff + 52)
        return 2;
    if (iy & ((1ull << (0x3ff + 52 - e)) - 1))
        return 0;
    return (iy & (1ull << (0x3ff + 52 - e))) ? 1 : 2;
}
TW_HD bool zero_inf_nan(uint64_t i) { return 2 * i - 1 >= 2 * 0x7ff0000000000000ull - 1; }
} // namespace twm

TW_HD double tw_pow(double x, double y)
{
    using namespace twm;
    const uint64_t ONE = 0x3ff0000000000000ull, INF = 0x7ff0000000000000ull;
    uint32_t sign_bias = 0;
    uint64_t ix = as_u64(x), iy = as_u64(y);
    uint32_t topx = top12(x), topy = top12(y);
    if (topx - 0x001u >= 0x7ffu - 0x001u || (topy & 0x7ff) - 0x3beu >= 0x43eu - 0x3beu)
    {
        if (zero_inf_nan(iy))
        {
            if (2 * iy == 0 || ix == ONE)
                return 1.0;
            if (2 * ix > 2 * INF || 2 * iy > 2 * INF)
                return x + y;
            if (2 * ix == 2 * ONE)
                return 1.0;
            if ((2 * ix < 2 * ONE) == !(iy >> 63))
                return 0.0;
            return y * y;
        }
        if (zero_inf_nan(ix))
        {
            double x2 = x * x;
            if ((ix >> 63) && int_kind(iy) == 1)
                x2 = -x2;
            return (iy >> 63) ? 1 / x2 : x2;
        }
        if (ix >> 63)
        {
            int k = int_kind(iy);
            if (k == 0)
                return as_f64(0x7ff8000000000000ull);
            if (k == 1)
                sign_bias = 0x800 << 7;
            ix &= 0x7fffffffffffffffull;
            topx &= 0x7ff;
        }
        if ((topy & 0x7ff) - 0x3beu >= 0x43eu - 0x3beu)
        {
            if (ix == ONE)
                return 1.0;
            if ((topy & 0x7ff) < 0x3be)
                return ix > ONE ? 1.0 + y : 1.0 - y;
            return (ix > ONE) == (topy < 0x800) ? as_f64(INF) : 0.0;
        }
        if (topx == 0) // subnormal x: normalize
        {
            ix = as_u64(x * 0x1p52) & 0x7fffffffffffffffull;
            ix -= 52ull << 52;
        }
    }
    double lo;
    double hi = log_hi_lo(ix, &lo);
    double ehi = y * hi;
    double elo = __builtin_fma(y, lo, __builtin_fma(y, hi, -ehi));
    return exp_hi_lo(ehi, elo, sign_bias);
}

// ns3::Time::GetSeconds() of a time in picoseconds. The Q64.64 quotient v = |ps| * Invert(10^12)
// (UmulByInvert: ah * bh + (ah * bl) >> 64); Invert(10^12) = floor(2^128 / 10^12) + 1 since ns-3
// rounds the inverse up when v * 10^12 falls short of 2^128. GetDouble sums the words in x87
// extended precision: v / 2^64 rounded to 64 significant bits, then to 53 (both to nearest even).
TW_HD double tw_seconds(int64_t ps)
{
    typedef unsigned __int128 u128;
    const uint64_t INV_HI = 18446744ull, INV_LO = 1359701234448236304ull;
    if (ps == 0)
        return 0.0;
    bool neg = ps < 0;
    uint64_t a = neg ? 0 - (uint64_t)ps : (uint64_t)ps;
    u128 v = (u128)a * INV_HI + (((u128)a * INV_LO) >> 64);
    uint64_t hi = (uint64_t)(v >> 64), lo = (uint64_t)v;
    int msb = hi ? 127 - __builtin_clzll(hi) : 63 - __builtin_clzll(lo);
    // round to `keep` significant bits, to nearest even
    auto round_to = [&](int keep) {
        if (msb < keep)
            return;
        int drop = msb - keep + 1;
        u128 half = (u128)1 << (drop - 1), rem = v & (((u128)1 << drop) - 1);
        v >>= drop;
        if (rem > half || (rem == half && (v & 1)))
            v += 1;
        v <<= drop;
        uint64_t h = (uint64_t)(v >> 64), l = (uint64_t)v;
        msb = h ? 127 - __builtin_clzll(h) : 63 - __builtin_clzll(l);
    };
    round_to(64);
    round_to(53);
    int sh = msb > 52 ? msb - 52 : 0;
    double d = (double)(uint64_t)(v >> sh); // < 2^53: exact
    int e = sh - 64;                       // value = d * 2^e
    d *= twm::as_f64((uint64_t)(1023 + e) << 52); // 2^e, normal for any time of a simulation
    return neg ? -d : d;
}

// ns-3's MRG32k3a stream state and draws (in doubles; every intermediate is an integer below 2^53)
struct TwMrg
{
    double s[6];
};
TW_HD double tw_mrg_u01(TwMrg &g)
{
    const double m1 = 4294967087.0, m2 = 4294944443.0;
    double p1 = 1403580.0 * g.s[1] - 810728.0 * g.s[0];
    int32_t k = (int32_t)(p1 / m1);
    p1 -= k * m1;
    if (p1 < 0.0)
        p1 += m1;
    g.s[0] = g.s[1];
    g.s[1] = g.s[2];
    g.s[2] = p1;
    double p2 = 527612.0 * g.s[5] - 1370589.0 * g.s[3];
    k = (int32_t)(p2 / m2);
    p2 -= k * m2;
    if (p2 < 0.0)
        p2 += m2;
    g.s[3] = g.s[4];
    g.s[4] = g.s[5];
    g.s[5] = p2;
    const double norm = 1.0 / 4294967088.0;
    return p1 > p2 ? (p1 - p2) * norm : (p1 - p2 + m1) * norm;
}
// UniformRandomVariable::GetInteger(0, n - 1)
TW_HD uint32_t tw_mrg_integer(TwMrg &g, uint32_t n)
{
    double u = tw_mrg_u01(g);
    double lo = 0.0, hi = (double)(n - 1) + 1.0;
    return (uint32_t)(lo + u * (hi - lo));
}
