// Host-side MRG32k3a stream states (ns-3's RngStream), shared by the TCP and RDMA set-up code.
#pragma once
#include <cstdint>
#include <cstring>

// ------------------------------------------------------------------ MRG32k3a streams (src/rng.h:60-159)
// RngStream(seed, 2^63 + stream, run): the seed vector times A^(stream 2^127 + run 2^76)
typedef unsigned __int128 u128;
struct Mat3
{
    uint64_t a[3][3];
};
static inline Mat3 mat_mul(const Mat3 &x, const Mat3 &y, uint64_t m)
{
    Mat3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
        {
            u128 v = 0;
            for (int k = 0; k < 3; k++)
                v += (u128)x.a[i][k] * y.a[k][j] % m;
            r.a[i][j] = (uint64_t)(v % m);
        }
    return r;
}
static inline void mat_apply(const Mat3 &x, uint64_t *v, uint64_t m)
{
    uint64_t r[3];
    for (int i = 0; i < 3; i++)
    {
        u128 s = 0;
        for (int k = 0; k < 3; k++)
            s += (u128)x.a[i][k] * v[k] % m;
        r[i] = (uint64_t)(s % m);
    }
    memcpy(v, r, sizeof r);
}
// powers A^(2^(e + b)), b = 0..63, computed once per (A, e)
struct Pow2Table
{
    Mat3 p[64];
    Pow2Table(Mat3 a, uint64_t m, int e)
    {
        for (int i = 0; i < e; i++)
            a = mat_mul(a, a, m);
        for (int b = 0; b < 64; b++)
        {
            p[b] = a;
            a = mat_mul(a, a, m);
        }
    }
    void jump(uint64_t n, uint64_t *v, uint64_t m) const
    {
        for (int b = 0; b < 64; b++)
            if ((n >> b) & 1)
                mat_apply(p[b], v, m);
    }
};
static inline void mrg_stream6(uint32_t seed, uint64_t stream, uint64_t run, double *out)
{
    const uint64_t m1 = 4294967087ull, m2 = 4294944443ull;
    static const Mat3 A1 = {{{0, 1, 0}, {0, 0, 1}, {m1 - 810728, 1403580, 0}}};
    static const Mat3 A2 = {{{0, 1, 0}, {0, 0, 1}, {m2 - 1370589, 0, 527612}}};
    static const Pow2Table s1(A1, m1, 127), s2(A2, m2, 127), r1(A1, m1, 76), r2(A2, m2, 76);
    uint64_t v1[3] = {seed, seed, seed}, v2[3] = {seed, seed, seed};
    uint64_t target = (1ull << 63) + stream;
    s1.jump(target, v1, m1);
    s2.jump(target, v2, m2);
    r1.jump(run, v1, m1);
    r2.jump(run, v2, m2);
    for (int i = 0; i < 3; i++)
        out[i] = (double)v1[i], out[3 + i] = (double)v2[i];
}
