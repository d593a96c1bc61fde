// Hero-wavelength spectral support (PathTracerSpectral).
//
//   SampleWavelengths   SingleSampleSpectrumWavelength, HYPERBOLIC_PBRT mode
//                       (Tracer/SpectrumContext.cu:14-135)
//   SpectralAlbedo      Jakob2019Detail::Converter::ConvertAlbedo
//                       (Tracer/SpectrumContext.hpp:37-125): largest channel picks
//                       the table, the other two over it are (x, y), z is the
//                       twice-inverted smoothstep of the largest value; trilinear
//                       fetch of three coefficients; sigmoid of a quadratic in nm
//   SpectralRadiance    ConvertRadiance (SpectrumContext.hpp:127-150)
//   SpectrumToRGB       ConvertSpectraToRGBSingle (SpectrumContext.cu:137-171)
// Tables: the CIE 1931 observer and D65 illuminant (colour_data.cpp), the
// Rec.709 XYZ->RGB matrix, and the coefficient LUT that spectral_lut.cpp fits.
#pragma once

#include "fmath.h"

#include "common.h"

namespace mray
{

static constexpr uint32_t kCieCount = 471;      // 360 .. 830 nm, 1 nm
static constexpr float kCieStart = 360.0f;

struct SpectralView
{
    const float4* __restrict__ lut;    // [3][N][N][N] (c0, c1, c2, 0), x fastest
    const float4* __restrict__ cmf;    // [471] (xbar, ybar, zbar, 0)
    const float* __restrict__ illum;   // [471]
    float xyzToRgb[9];
    uint32_t N;
};

struct SpectralTables
{
    SpectralView view{};
    bool ready = false;
};

// Builds (once per process) and uploads the tables.
int GetSpectralTables(SpectralTables& out);

__host__ __device__ __forceinline__ void SampleWavelengths(float xi0, V4& waves, V4& pdf)
{
    const float off[4] = {-0.5f, -0.25f, 0.0f, 0.25f};
    float w[4], p[4];
    for(int i = 0; i < 4; i++)
    {
        float x = xi0 + off[i];
        if(x < 0.0f) x += 1.0f;
        if(x >= 1.0f) x -= 1.0f;
        float a = 0.85691062f - 1.82750197f * x;
        float lam = 538.0f - 138.888889f * fm::Atanh(a);
        float denom = fm::Cosh(0.0072f * (lam - 538.0f));
        denom = denom * denom;
        w[i] = lam;
        p[i] = 0.0039398042f / denom;
    }
    waves = {w[0], w[1], w[2], w[3]};
    pdf = {p[0], p[1], p[2], p[3]};
}

// Math::InvSmoothstep (Core/Math.h:1224-1232)
__device__ __forceinline__ float InvSmoothstep(float y)
{
    return 0.5f - fm::Sin(fm::Asin(1.0f - 2.0f * y) * (1.0f / 3.0f));
}

__device__ __forceinline__ V3 LutFetch(const SpectralView& sv, uint32_t ch, float x, float y, float z)
{
    // normalised coordinate u maps to texel space u * (N - 1) (texel centres),
    // trilinear with clamp-to-edge
    const int N = int(sv.N);
    float fx = ClampF(x, 0.0f, 1.0f) * float(N - 1);
    float fy = ClampF(y, 0.0f, 1.0f) * float(N - 1);
    float fz = ClampF(z, 0.0f, 1.0f) * float(N - 1);
    int x0 = min(int(fx), N - 2), y0 = min(int(fy), N - 2), z0 = min(int(fz), N - 2);
    float tx = fx - float(x0), ty = fy - float(y0), tz = fz - float(z0);
    const float4* base = sv.lut + size_t(ch) * N * N * N;
    auto at = [&](int zz, int yy, int xx) { return base[(size_t(zz) * N + yy) * N + xx]; };
    float4 c000 = at(z0, y0, x0), c001 = at(z0, y0, x0 + 1);
    float4 c010 = at(z0, y0 + 1, x0), c011 = at(z0, y0 + 1, x0 + 1);
    float4 c100 = at(z0 + 1, y0, x0), c101 = at(z0 + 1, y0, x0 + 1);
    float4 c110 = at(z0 + 1, y0 + 1, x0), c111 = at(z0 + 1, y0 + 1, x0 + 1);
    auto lerp4 = [](float4 a, float4 b, float t)
    {
        return make_float4(a.x * (1 - t) + b.x * t, a.y * (1 - t) + b.y * t,
                           a.z * (1 - t) + b.z * t, 0.0f);
    };
    float4 c00 = lerp4(c000, c001, tx), c01 = lerp4(c010, c011, tx);
    float4 c10 = lerp4(c100, c101, tx), c11 = lerp4(c110, c111, tx);
    float4 c0 = lerp4(c00, c01, ty), c1 = lerp4(c10, c11, ty);
    float4 c = lerp4(c0, c1, tz);
    return {c.x, c.y, c.z};
}

__device__ __forceinline__ float EvalSigmoidPoly(V3 c, float lambda)
{
    float t = Fma(c.x, lambda, c.y);
    float x = Fma(t, lambda, c.z);
    float denomRecip = 1.0f / sqrtf(Fma(x, x, 1.0f));
    return Fma(0.5f * x, denomRecip, 0.5f);
}

// The sigmoid-polynomial coefficients of an RGB albedo (the table lookup
// half of SpectralAlbedo); constant colours are converted once per scene.
__device__ __forceinline__ V3 AlbedoCoeffs(const SpectralView& sv, V3 rgb)
{
    float c[3] = {rgb.x, rgb.y, rgb.z};
    uint32_t maxI = 0;
    if(c[1] > c[maxI]) maxI = 1;
    if(c[2] > c[maxI]) maxI = 2;
    float maxC = c[maxI];
    float x = 0.0f, y = 0.0f;
    if(maxC > 1e-7f)
    {
        float f = 1.0f / maxC;
        x = c[(maxI + 1) % 3] * f;
        y = c[(maxI + 2) % 3] * f;
    }
    float slice = ClampF(maxC, 0.0f, 1.0f);
    float z = InvSmoothstep(InvSmoothstep(slice));
    return LutFetch(sv, maxI, x, y, z);
}
__device__ __forceinline__ V4 EvalCoeffs(V3 co, V4 waves)
{
    return {EvalSigmoidPoly(co, waves.x), EvalSigmoidPoly(co, waves.y),
            EvalSigmoidPoly(co, waves.z), EvalSigmoidPoly(co, waves.w)};
}
__device__ __forceinline__ V4 SpectralAlbedo(const SpectralView& sv, V3 rgb, V4 waves)
{
    return EvalCoeffs(AlbedoCoeffs(sv, rgb), waves);
}

__device__ __forceinline__ float Tex1D(const float* __restrict__ t, float lambda)
{
    float f = ClampF(lambda - kCieStart, 0.0f, float(kCieCount - 1));
    int i = min(int(f), int(kCieCount) - 2);
    float u = f - float(i);
    return t[i] * (1 - u) + t[i + 1] * u;
}
__device__ __forceinline__ V3 Tex1D3(const float4* __restrict__ t, float lambda)
{
    float f = ClampF(lambda - kCieStart, 0.0f, float(kCieCount - 1));
    int i = min(int(f), int(kCieCount) - 2);
    float u = f - float(i);
    float4 a = t[i], b = t[i + 1];
    return {a.x * (1 - u) + b.x * u, a.y * (1 - u) + b.y * u, a.z * (1 - u) + b.z * u};
}

// Radiance = albedo of rgb / (2 max) times the illuminant, times 2 max:
// coefficients (c0, c1, c2) and the scale, then their evaluation.
__device__ __forceinline__ float4 RadianceCoeffs(const SpectralView& sv, V3 rgb)
{
    float mx = MaxF(MaxF(rgb.x, rgb.y), rgb.z);
    float scale = mx * 2.0f;
    V3 n = (scale == 0.0f) ? V3{0, 0, 0} : V3{rgb.x / scale, rgb.y / scale, rgb.z / scale};
    V3 co = AlbedoCoeffs(sv, n);
    return make_float4(co.x, co.y, co.z, scale);
}
__device__ __forceinline__ V4 EvalRadianceCoeffs(const SpectralView& sv, float4 rc, V4 waves)
{
    V4 s = EvalCoeffs(V3{rc.x, rc.y, rc.z}, waves);
    s.x *= Tex1D(sv.illum, waves.x);
    s.y *= Tex1D(sv.illum, waves.y);
    s.z *= Tex1D(sv.illum, waves.z);
    s.w *= Tex1D(sv.illum, waves.w);
    return s * rc.w;
}
__device__ __forceinline__ V4 SpectralRadiance(const SpectralView& sv, V3 rgb, V4 waves)
{
    return EvalRadianceCoeffs(sv, RadianceCoeffs(sv, rgb), waves);
}

__device__ __forceinline__ V4 SpectrumToRGB(const SpectralView& sv, V4 value, V4 waves, V4 pdf,
                                            bool dispersed)
{
    float val[4] = {value.x, value.y, value.z, value.w};
    float wv[4] = {waves.x, waves.y, waves.z, waves.w};
    float pd[4] = {pdf.x, pdf.y, pdf.z, pdf.w};
    int count = dispersed ? 1 : 4;
    float weight = dispersed ? 1.0f : 0.25f;
    V3 xyz = {0, 0, 0};
    for(int i = 0; i < count; i++)
    {
        V3 f = Tex1D3(sv.cmf, wv[i]);
        float v = (pd[i] == 0.0f) ? 0.0f : val[i] / pd[i];
        xyz = xyz + f * v;
    }
    xyz = xyz * weight;
    const float* M = sv.xyzToRgb;
    return {M[0] * xyz.x + M[1] * xyz.y + M[2] * xyz.z,
            M[3] * xyz.x + M[4] * xyz.y + M[5] * xyz.z,
            M[6] * xyz.x + M[7] * xyz.y + M[8] * xyz.z, 0.0f};
}

} // namespace mray
