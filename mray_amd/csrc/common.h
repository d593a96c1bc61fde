// Shared host/device definitions of the MI355X (gfx950) hot path.
//
// All translation units are compiled with -ffp-contract=off: every arithmetic
// operation below is one IEEE-754 rounding, and FMAs appear only where they
// are written (__builtin_fmaf), which is where the reference calls Math::FMA
// (Core/Math.h:1586-1632).  That is what makes integer outputs (Morton codes,
// hit triangle ids) bit-identical to the CPU restatement.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/mray_amd.h"

namespace mray
{

// ---------------------------------------------------------------------------
// Error reporting (the reference throws MRayError; the C-ABI returns a code
// and keeps the message for mray_last_error()).
// ---------------------------------------------------------------------------
void SetError(const std::string& msg);

#define MRAY_HIP_CHECK(expr)                                                       \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if(e_ != hipSuccess)                                                       \
        {                                                                          \
            ::mray::SetError(std::string(#expr) + ": " + hipGetErrorString(e_));   \
            return MRAY_ERR_HIP;                                                   \
        }                                                                          \
    } while(0)

// propagate a non-OK status code
#define MRAY_TRY(expr)                  \
    do {                                \
        int r_ = (expr);                \
        if(r_ != MRAY_OK) return r_;    \
    } while(0)

static constexpr uint32_t kLeafFlag = 0x80000000u;
static constexpr uint32_t kInvalid = 0xFFFFFFFFu;
static constexpr uint32_t kTPB = 256;   // default workgroup: 4 wave64s

inline uint32_t DivUp(uint64_t a, uint64_t b) { return uint32_t((a + b - 1) / b); }

// ---------------------------------------------------------------------------
// Small vector types with the reference's arithmetic
// ---------------------------------------------------------------------------
struct V3 { float x, y, z; };
struct V4 { float x, y, z, w; };
struct Q4 { float w, x, y, z; };

#define MRAY_HD __host__ __device__ __forceinline__

MRAY_HD V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
MRAY_HD V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
MRAY_HD V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
MRAY_HD V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
MRAY_HD V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
MRAY_HD V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }

MRAY_HD V4 operator+(V4 a, V4 b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
MRAY_HD V4 operator-(V4 a, V4 b) { return {a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; }
MRAY_HD V4 operator*(V4 a, V4 b) { return {a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w}; }
MRAY_HD V4 operator*(V4 a, float s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
MRAY_HD V4 operator*(float s, V4 a) { return {s * a.x, s * a.y, s * a.z, s * a.w}; }
MRAY_HD V4 operator/(V4 a, float s) { return {a.x / s, a.y / s, a.z / s, a.w / s}; }
MRAY_HD V4 Splat4(float s) { return {s, s, s, s}; }
MRAY_HD float Sum4(V4 a) { return a.x + a.y + a.z + a.w; }

MRAY_HD float Fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// Math::Dot / Math::Cross (Core/Math.h:1586-1632): explicit FMA chains
MRAY_HD float Dot(V3 a, V3 b)
{
    float r = 0.0f;
    r = Fma(a.x, b.x, r);
    r = Fma(a.y, b.y, r);
    r = Fma(a.z, b.z, r);
    return r;
}
MRAY_HD float Dot2(float ax, float ay, float bx, float by)
{
    float r = 0.0f;
    r = Fma(ax, bx, r);
    r = Fma(ay, by, r);
    return r;
}
MRAY_HD float Dot4(V4 a, V4 b)
{
    float r = 0.0f;
    r = Fma(a.x, b.x, r);
    r = Fma(a.y, b.y, r);
    r = Fma(a.z, b.z, r);
    r = Fma(a.w, b.w, r);
    return r;
}
MRAY_HD V3 Cross(V3 a, V3 b)
{
    return {Fma(a.y, b.z, -a.z * b.y), Fma(a.z, b.x, -a.x * b.z), Fma(a.x, b.y, -a.y * b.x)};
}
MRAY_HD float LengthSqr(V3 v) { return Dot(v, v); }
MRAY_HD float Length(V3 v) { return sqrtf(Dot(v, v)); }
MRAY_HD V3 Normalize(V3 v)
{
    float r = 1.0f / Length(v);
    return v * r;
}
MRAY_HD float MaxF(float a, float b) { return fmaxf(a, b); }
MRAY_HD float MinF(float a, float b) { return fminf(a, b); }
MRAY_HD float ClampF(float v, float lo, float hi) { return MinF(MaxF(v, lo), hi); }
MRAY_HD float SqrtMax(float x) { return sqrtf(MaxF(x, 0.0f)); }

} // namespace mray
