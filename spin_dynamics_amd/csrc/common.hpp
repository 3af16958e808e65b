// common.hpp -- shared device helpers for the gfx950 path: complex FP64 on
// double2, xoroshiro128+ streams, launch geometry.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spc {

constexpr int kWave = 64;      // CDNA wavefront
constexpr int kMaxKrylov = 32; // templates are instantiated for 0..kMaxKrylov
constexpr int kMaxNuclei = 64;

// ------------------------------------------------------------ complex FP64
__host__ __device__ __forceinline__ double2 cz(double re, double im) { return make_double2(re, im); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cscale(double s, double2 a) { return make_double2(s * a.x, s * a.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
// acc + a*b
__device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 acc)
{
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(-a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(a.y, b.x, acc.y);
    return acc;
}
// acc + conj(a)*b
__device__ __forceinline__ double2 cfma_conj(double2 a, double2 b, double2 acc)
{
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(-a.y, b.x, acc.y);
    return acc;
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double cabs2(double2 a) { return fma(a.x, a.x, a.y * a.y); }
__device__ __forceinline__ double2 cdiv(double2 a, double2 b)
{
    // Smith's algorithm
    if (fabs(b.x) >= fabs(b.y)) {
        double r = b.y / b.x, den = b.x + b.y * r;
        return make_double2((a.x + a.y * r) / den, (a.y - a.x * r) / den);
    }
    double r = b.x / b.y, den = b.x * r + b.y;
    return make_double2((a.x * r + a.y) / den, (a.y * r - a.x) / den);
}

// Non-temporal (streaming) load / store of one complex value.
__device__ __forceinline__ double2 ntload(const double2 *p)
{
    return make_double2(__builtin_nontemporal_load(&p->x), __builtin_nontemporal_load(&p->y));
}
__device__ __forceinline__ void ntstore(double2 *p, double2 v)
{
    __builtin_nontemporal_store(v.x, &p->x);
    __builtin_nontemporal_store(v.y, &p->y);
}

// ----------------------------------------------------------- xoroshiro128+
// Same stream as the reference generator (m_random.f90:90-198, 280-340).
struct Rng {
    uint64_t s0, s1;
    double stored;
    int have;
};

__device__ __forceinline__ uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

__device__ __forceinline__ uint64_t rng_next(Rng &r)
{
    uint64_t a = r.s0, b = r.s1;
    uint64_t res = a + b;
    b ^= a;
    r.s0 = rotl64(a, 55) ^ b ^ (b << 14);
    r.s1 = rotl64(b, 36);
    return res;
}

__device__ inline void rng_jump(Rng &r)
{
    const uint64_t jmp[2] = {0xbeac0467eba5facbull, 0xd86b048b86aa9922ull};
    uint64_t t0 = 0, t1 = 0;
    for (int i = 0; i < 2; ++i)
        for (int b = 0; b < 64; ++b) {
            if (jmp[i] & (1ull << b)) {
                t0 ^= r.s0;
                t1 ^= r.s1;
            }
            (void)rng_next(r);
        }
    r.s0 = t0;
    r.s1 = t1;
}

__device__ inline void rng_set_seed(Rng &r, int64_t a, int64_t b)
{
    r.s0 = (uint64_t)a;
    r.s1 = (uint64_t)b;
    r.stored = 0.0;
    r.have = 0;
    rng_jump(r);
}

__device__ __forceinline__ double rng_unif01(Rng &r)
{
    uint64_t x = rng_next(r);
    x = (1023ull << 52) | (x >> 12);
    return __longlong_as_double((long long)x) - 1.0;
}

__device__ inline void rng_two_normals(Rng &r, double &o0, double &o1)
{
    double a, b, s;
    for (;;) {
        a = 2.0 * rng_unif01(r) - 1.0;
        b = 2.0 * rng_unif01(r) - 1.0;
        s = a * a + b * b;
        if (s < 1.0 && s > 0.0) break;
    }
    double f = sqrt(-2.0 * log(s) / s);
    o0 = a * f;
    o1 = b * f;
}

__device__ inline double rng_normal(Rng &r)
{
    if (r.have) {
        r.have = 0;
        return r.stored;
    }
    double a, b;
    rng_two_normals(r, a, b);
    r.stored = b;
    r.have = 1;
    return a;
}

// rng%circle (m_random.f90:280-297) draws a pair (a, b) uniform in the unit
// disc by rejection (2 unif01 per try, accepted when a^2 + b^2 <= 1) and
// returns the point on the unit circle at twice its angle, computed here.
__device__ inline void circle_point(double a, double b, double &x, double &y)
{
    const double s = a * a + b * b;
    x = (a * a - b * b) / s;
    y = 2.0 * a * b / s;
}

__device__ inline void rng_sphere(Rng &r, double radius, double *o)
{
    double a, b, s;
    for (;;) {
        a = 2.0 * rng_unif01(r) - 1.0;
        b = 2.0 * rng_unif01(r) - 1.0;
        s = a * a + b * b;
        if (s <= 1.0) break;
    }
    double t = sqrt(1.0 - s);
    o[0] = 2.0 * a * t * radius;
    o[1] = 2.0 * b * t * radius;
    o[2] = (1.0 - 2.0 * s) * radius;
}

// ------------------------------------------------------------- geometry
// Workgroups are dealt round-robin to the 8 XCDs; remap so that each XCD
// works on one contiguous range of logical blocks (shared L2 for the gathers).
__device__ __forceinline__ int xcd_remap(int bid, int nblocks)
{
    if ((nblocks & 7) != 0) return bid;
    return (bid & 7) * (nblocks >> 3) + (bid >> 3);
}

__device__ __forceinline__ int wave_uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

} // namespace spc
