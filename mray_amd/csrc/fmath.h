// Elementary functions of the path's sampling code (sin, cos, acos, asin,
// atan2, exp, log, atanh, cosh) in float, written with + - * / sqrt, fma,
// rint and bit operations only.  The gfx950 kernels and the CPU oracle both
// call these, so the same inputs give bit-identical results on both sides:
// a path's random decisions (a ray grazing a box edge, a cosine-weighted
// bounce) cannot diverge through a last-bit difference between the GPU's and
// the C library's functions, which would shift every later refill of the
// reference schedule.
//
// Methods: Cody-Waite range reduction (Cody and Waite, "Software Manual for
// the Elementary Functions", 1980); Taylor polynomials for sin / cos on
// |r| <= pi/4 and for log through log(1+f) = 2 atanh(f / (2+f)); polynomials
// for exp and atan fitted by this repository (fmath_fit.py: least squares on
// Chebyshev nodes).  Accuracy against double-precision references is
// checked in tests/test_fmath.py (a few ulp over the ranges the path uses).
// Plain C++ besides the HIP qualifiers, so the oracle compiles it with g++.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

#if defined(__HIPCC__) || defined(__HIP__)
#define MRAY_FM __host__ __device__ inline
#else
#define MRAY_FM inline
#endif

namespace mray
{
namespace fm
{

MRAY_FM uint32_t Bits(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
MRAY_FM float FromBits(uint32_t u)
{
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
MRAY_FM float Fma(float a, float b, float c) { return fmaf(a, b, c); }
MRAY_FM float Rint(float x) { return rintf(x); }    // nearest, ties to even

constexpr float kInf = __builtin_huge_valf();
constexpr float kPiHi = 3.14159274101257324f;       // float(pi)
constexpr float kPiLo = -8.74227766e-08f;           // pi - kPiHi
constexpr float kPio2Hi = 1.57079637050628662f;
constexpr float kPio2Lo = -4.37113883e-08f;
constexpr float kPio4Hi = 0.785398185253143311f;
constexpr float kPio4Lo = -2.18556941e-08f;
// pi/2 in three parts for the sin / cos reduction
constexpr float kRedA = 1.57079637050628662f;
constexpr float kRedB = -4.37113883e-08f;
constexpr float kRedC = -1.77635684e-15f;
constexpr float kLn2Hi = 0.693145751953125f;        // 0x3f317200: trailing zeros
constexpr float kLn2Lo = 1.42860677e-06f;

// 2^k for |k| <= 150 as two exact powers (the product rounds once, into the
// subnormal range when it must)
MRAY_FM float Scale2(float x, int k)
{
    int k1 = k / 2;
    int k2 = k - k1;
    return (x * FromBits(uint32_t(k1 + 127) << 23)) * FromBits(uint32_t(k2 + 127) << 23);
}

// e^x: x = k ln2 + r, |r| <= ln2/2; e^r = 1 + r + r^2 P(r)
MRAY_FM float Exp(float x)
{
    if(x != x) return x;
    if(x > 88.7228394f) return kInf;
    if(x < -103.972084f) return 0.0f;
    float k = Rint(x * 1.44269502f);
    float r = Fma(-k, kLn2Hi, x);
    r = Fma(-k, kLn2Lo, r);
    float p = 1.98992755e-04f;
    p = Fma(p, r, 1.39336439e-03f);
    p = Fma(p, r, 8.33329838e-03f);
    p = Fma(p, r, 4.16664667e-02f);
    p = Fma(p, r, 1.66666672e-01f);
    p = Fma(p, r, 0.5f);
    float e = Fma(r * r, p, r) + 1.0f;
    return Scale2(e, int(k));
}

// natural log: x = 2^e m, m in [sqrt(1/2), sqrt(2)], f = m - 1,
// log(1 + f) = 2 s (1 + z/3 + z^2/5 + ...), s = f / (2 + f), z = s^2
MRAY_FM float Log(float x)
{
    if(x != x) return x;
    if(x < 0.0f) return __builtin_nanf("");
    if(x == 0.0f) return -kInf;
    if(x == kInf) return x;
    int e = 0;
    uint32_t b = Bits(x);
    if(b < 0x00800000u)
    {
        x *= 8388608.0f;                            // 2^23: subnormal to normal
        b = Bits(x);
        e = -23;
    }
    e += int((b >> 23) & 0xFFu) - 127;
    float m = FromBits((b & 0x007FFFFFu) | 0x3F800000u);
    if(m > 1.41421354f)
    {
        m *= 0.5f;
        e += 1;
    }
    float f = m - 1.0f;
    float s = f / (2.0f + f);
    float z = s * s;
    float R = 0.0909090936f;                        // 1/11
    R = Fma(R, z, 0.111111112f);                    // 1/9
    R = Fma(R, z, 0.142857149f);                    // 1/7
    R = Fma(R, z, 0.200000003f);                    // 1/5
    R = Fma(R, z, 0.333333343f);                    // 1/3
    float s2 = s + s;
    float lm = Fma(s2, z * R, s2);
    float fe = float(e);
    return Fma(fe, kLn2Hi, Fma(fe, kLn2Lo, lm));
}

// sin and cos together: x = k pi/2 + r, |r| <= pi/4
MRAY_FM void SinCos(float x, float& s, float& c)
{
    if(!(x - x == 0.0f))                            // inf or nan
    {
        s = c = __builtin_nanf("");
        return;
    }
    float k = Rint(x * 0.636619747f);
    float r = Fma(-k, kRedA, x);
    r = Fma(-k, kRedB, r);
    r = Fma(-k, kRedC, r);
    float r2 = r * r;
    float ps = 2.75573188e-06f;                     // 1/9!
    ps = Fma(ps, r2, -1.98412701e-04f);             // -1/7!
    ps = Fma(ps, r2, 8.33333377e-03f);              // 1/5!
    ps = Fma(ps, r2, -0.166666672f);                // -1/3!
    float sr = Fma(r * r2, ps, r);
    float pc = -2.75573200e-07f;                    // -1/10!
    pc = Fma(pc, r2, 2.48015876e-05f);              // 1/8!
    pc = Fma(pc, r2, -1.38888892e-03f);             // -1/6!
    pc = Fma(pc, r2, 4.16666679e-02f);              // 1/4!
    pc = Fma(pc, r2, -0.5f);
    float cr = Fma(r2, pc, 1.0f);
    int q = int(k) & 3;
    if(q == 0) { s = sr; c = cr; }
    else if(q == 1) { s = cr; c = -sr; }
    else if(q == 2) { s = -sr; c = -cr; }
    else { s = -cr; c = sr; }
}
MRAY_FM float Sin(float x) { float s, c; SinCos(x, s, c); return s; }
MRAY_FM float Cos(float x) { float s, c; SinCos(x, s, c); return c; }

// atan(u) = u + u^3 P(u^2), |u| <= tan(pi/8)
MRAY_FM float AtanCore(float u)
{
    float z = u * u;
    float p = 2.74611376e-02f;
    p = Fma(p, z, -5.67722842e-02f);
    p = Fma(p, z, 7.51583278e-02f);
    p = Fma(p, z, -9.07263234e-02f);
    p = Fma(p, z, 1.11100323e-01f);
    p = Fma(p, z, -1.42856807e-01f);
    p = Fma(p, z, 1.99999988e-01f);
    p = Fma(p, z, -3.33333343e-01f);
    return Fma(u * z, p, u);
}
// atan on [0, 1]: above tan(pi/8), atan(t) = pi/4 + atan((t - 1) / (t + 1))
MRAY_FM float Atan01(float t)
{
    if(t <= 0.414213568f) return AtanCore(t);
    return (kPio4Hi + AtanCore((t - 1.0f) / (t + 1.0f))) + kPio4Lo;
}

MRAY_FM float Atan2(float y, float x)
{
    if(x != x || y != y) return x + y;
    const bool yNeg = (Bits(y) >> 31) != 0u, xNeg = (Bits(x) >> 31) != 0u;
    const float ay = fabsf(y), ax = fabsf(x);
    float a;
    if(ay == 0.0f && ax == 0.0f) a = xNeg ? kPiHi : 0.0f;
    else
    {
        if(ay == kInf && ax == kInf) a = kPio4Hi;
        else if(ay > ax) a = (kPio2Hi - Atan01(ax / ay)) + kPio2Lo;
        else a = Atan01(ay / ax);
        if(xNeg) a = (kPiHi - a) + kPiLo;
    }
    return yNeg ? -a : a;
}
MRAY_FM float Acos(float x)
{
    return Atan2(sqrtf((1.0f - x) * (1.0f + x)), x);
}
MRAY_FM float Asin(float x)
{
    return Atan2(x, sqrtf((1.0f - x) * (1.0f + x)));
}
// atanh: the odd series a + a^3/3 + a^5/5 + ... for |a| <= 1/4 (keeps
// relative accuracy near zero), else log((1 + a) / (1 - a)) / 2
MRAY_FM float Atanh(float a)
{
    if(fabsf(a) <= 0.25f)
    {
        float z = a * a;
        float p = 0.0769230798f;                    // 1/13
        p = Fma(p, z, 0.0909090936f);               // 1/11
        p = Fma(p, z, 0.111111112f);                // 1/9
        p = Fma(p, z, 0.142857149f);                // 1/7
        p = Fma(p, z, 0.200000003f);                // 1/5
        p = Fma(p, z, 0.333333343f);                // 1/3
        return Fma(a * z, p, a);
    }
    return 0.5f * Log((1.0f + a) / (1.0f - a));
}
MRAY_FM float Cosh(float x)
{
    return 0.5f * (Exp(x) + Exp(-x));
}

} // namespace fm
} // namespace mray
