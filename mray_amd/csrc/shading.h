// Device-side shading: surface reconstruction, Lambert / Reflect / Refract /
// Unreal BxDFs, triangle + boundary lights, pinhole camera with the Gaussian
// film filter, PCG32.  Each function restates the reference routine named in
// its comment with the same operation order (files under the reference's
// Source/ directory).
#pragma once

#include "common.h"
#include "fmath.h"
#include "scene.h"

namespace mray
{

static constexpr float kPi = 3.14159265358979323846f;
static constexpr float kInvPi = 1.0f / kPi;
static constexpr float kSqrt2 = 1.41421356237309504880f;
static constexpr float kSqrtPi = 1.77245385090551602730f;
static constexpr float kEps = 1e-5f;
static constexpr float kLargeEps = 1e-4f;

// ---------------------------------------------------------------------------
// PermutedCG32 (Tracer/Random.h:761-843), ToFloat01 (Random.h:101-120)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t PcgNext(uint32_t& st)
{
    uint32_t s = st;
    uint32_t r = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    r = (r >> 22u) ^ r;
    st = s * 747796405u + 2891336453u;
    return r;
}
__device__ __forceinline__ float ToFloat01(uint32_t v)
{
    return MinF(float(v) * 0x1p-32f, 0x1.fffffep-1f);
}
__device__ __forceinline__ float NextRN(uint32_t& st) { return ToFloat01(PcgNext(st)); }

// ---------------------------------------------------------------------------
// Quaternion (Core/Quaternion.hpp)
// ---------------------------------------------------------------------------
__device__ __forceinline__ Q4 QMul(Q4 a, Q4 b)
{
    return {a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z,
            a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
            a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
            a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w};
}
__device__ __forceinline__ float QDot(Q4 a, Q4 b)
{
    return Dot4(V4{a.w, a.x, a.y, a.z}, V4{b.w, b.x, b.y, b.z});
}
__device__ __forceinline__ Q4 QAddScaled(Q4 a, float sa, Q4 b, float sb)
{
    Q4 x = {a.w * sa, a.x * sa, a.y * sa, a.z * sa};
    Q4 y = {b.w * sb, b.x * sb, b.y * sb, b.z * sb};
    return {x.w + y.w, x.x + y.x, x.y + y.y, x.z + y.z};
}
__device__ __forceinline__ Q4 QNormalize(Q4 q)
{
    V4 v = {q.w, q.x, q.y, q.z};
    float r = 1.0f / sqrtf(Dot4(v, v));
    return {q.w * r, q.x * r, q.y * r, q.z * r};
}
__device__ __forceinline__ V3 QRotate(Q4 q, V3 v)
{
    V3 u = {q.x, q.y, q.z};
    V3 t = Cross(u, v) * 2.0f;
    return v + q.w * t + Cross(u, t);
}
__device__ __forceinline__ V3 QInvRotate(Q4 q, V3 v) { return QRotate(Q4{q.w, -q.x, -q.y, -q.z}, v); }
__device__ __forceinline__ V3 QBasisZ(Q4 q)
{
    float v00 = q.w * q.w, v01 = q.w * q.x, v02 = q.w * q.y;
    float v11 = q.x * q.x, v13 = q.x * q.z, v22 = q.y * q.y;
    float v23 = q.y * q.z, v33 = q.z * q.z;
    return {v13 - v02 + v13 - v02, v23 + v01 + v23 + v01, v00 - v11 - v22 + v33};
}
__device__ __forceinline__ Q4 QSLerp(Q4 s, Q4 e, float t)
{
    float cosTheta = QDot(s, e);
    float cosFlipped = (cosTheta >= 0) ? cosTheta : -cosTheta;
    float s0, s1;
    if(cosFlipped < (1 - kEps))
    {
        float angle = fm::Acos(cosFlipped);
        float sinAngleRecip = 1.0f / fm::Sin(angle);
        s0 = fm::Sin(angle * (1 - t)) * sinAngleRecip;
        s1 = fm::Sin(angle * t) * sinAngleRecip;
    }
    else
    {
        s0 = (1 - t);
        s1 = t;
    }
    s1 = (cosTheta >= 0) ? s1 : -s1;
    return QAddScaled(s, s0, e, s1);
}
__device__ __forceinline__ Q4 QBarySLerp(Q4 q0, Q4 q1, Q4 q2, float a, float b)
{
    float c = (1 - a - b);
    if(fabsf(a + b) < kEps) return q2;
    float ab = a / (a + b);
    Q4 qAB = QSLerp(q1, q0, ab);
    return QSLerp(qAB, q2, c);
}

// ---------------------------------------------------------------------------
// Ray::Nudge (Core/Ray.hpp:258-300)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float NudgeC(float p, float n)
{
    const float ORIGIN = 1.0f / 32.0f;
    const float FLOAT_SCALE = 1.0f / 65536.0f;
    const float INT_SCALE = 256.0f;
    int32_t ofi = int32_t(INT_SCALE * n);
    int32_t pi = __float_as_int(p);
    pi += (p < 0.0f) ? -ofi : ofi;
    float pf = __int_as_float(pi);
    return (fabsf(p) < ORIGIN) ? (p + FLOAT_SCALE * n) : pf;
}
__device__ __forceinline__ V3 Nudge(V3 p, V3 n) { return {NudgeC(p.x, n.x), NudgeC(p.y, n.y), NudgeC(p.z, n.z)}; }

// ---------------------------------------------------------------------------
// Distribution helpers (Tracer/DistributionFunctions.h, Core/Math.h)
// ---------------------------------------------------------------------------
// Math::InvErrFunc host polynomial (Math.h:787-833)
__device__ __forceinline__ float InvErf(float x)
{
    float t = fm::Log(Fma(x, 0.0f - x, 1.0f));
    float p;
    if(fabsf(t) > 6.125f)
    {
        p = 3.03697567e-10f;
        p = Fma(p, t, 2.93243101e-8f);
        p = Fma(p, t, 1.22150334e-6f);
        p = Fma(p, t, 2.84108955e-5f);
        p = Fma(p, t, 3.93552968e-4f);
        p = Fma(p, t, 3.02698812e-3f);
        p = Fma(p, t, 4.83185798e-3f);
        p = Fma(p, t, -2.64646143e-1f);
        p = Fma(p, t, 8.40016484e-1f);
    }
    else
    {
        p = 5.438778320e-9f;
        p = Fma(p, t, 1.43285448e-7f);
        p = Fma(p, t, 1.22774793e-6f);
        p = Fma(p, t, 1.12963626e-7f);
        p = Fma(p, t, -5.61530760e-5f);
        p = Fma(p, t, -1.47697632e-4f);
        p = Fma(p, t, 2.31468678e-3f);
        p = Fma(p, t, 1.15392581e-2f);
        p = Fma(p, t, -2.32015476e-1f);
        p = Fma(p, t, 8.86226892e-1f);
    }
    return x * p;
}
__device__ __forceinline__ float Gaussian(float x, float sigma)
{
    const float invSqrt2Pi = (1.0f / kSqrt2) * (1.0f / kSqrtPi);
    float sigmaInv = 1.0f / sigma;
    float result = invSqrt2Pi * sigmaInv;
    float p = (x - 0.0f) * sigmaInv;
    result *= fm::Exp(-0.5f * p * p);
    return result;
}
__device__ __forceinline__ float SampleGaussian(float xi, float sigma)
{
    float x = kSqrt2 * sigma;
    float e = InvErf(2.0f * xi - 1.0f);
    x = x * e + 0.0f;
    if(isinf(e))
    {
        float mm = 3.5f * sigma;
        x = ClampF(x, -mm, mm);
    }
    return x;
}
__device__ __forceinline__ V3 SampleCosDir(float xi0, float xi1, float& pdf)
{
    float ang = 2.0f * kPi * xi1;
    float r = sqrtf(xi0);
    float s, c;
    fm::SinCos(ang, s, c);
    V3 d;
    d.x = r * c;
    d.y = r * s;
    d.z = SqrtMax(1.0f - Dot2(d.x, d.y, d.x, d.y));
    pdf = d.z * kInvPi;
    return d;
}
__device__ __forceinline__ float PdfCosDir(V3 v)
{
    float pdf = v.z * kInvPi;
    return (pdf <= kEps) ? 0.0f : pdf;
}
__device__ __forceinline__ V3 Reflect(V3 n, V3 v) { return 2.0f * Dot(v, n) * n - v; }
__device__ __forceinline__ bool Refract(V3& out, V3 n, V3 v, float etaFrom, float etaTo)
{
    float etaRatio = etaFrom / etaTo;
    float cosIn = Dot(n, v);
    float sinInSqr = MaxF(0.0f, 1.0f - cosIn * cosIn);
    float sinOutSqr = etaRatio * etaRatio * sinInSqr;
    float cosOut = SqrtMax(1.0f - sinOutSqr);
    if(sinOutSqr >= 1.0f) return false;
    out = etaRatio * (-v) + (etaRatio * cosIn - cosOut) * n;
    return true;
}
__device__ __forceinline__ float FresnelDielectric(float cosF, float etaF, float etaB)
{
    float sinF = SqrtMax(1.0f - cosF * cosF);
    float sinB = etaF / etaB * sinF;
    if(sinF >= 1.0f) return 1.0f;
    float cosB = SqrtMax(1.0f - sinB * sinB);
    float par = ((etaB * cosF - etaF * cosB) / (etaB * cosF + etaF * cosB));
    par = par * par;
    float perp = ((etaF * cosF - etaB * cosB) / (etaF * cosF + etaB * cosB));
    perp = perp * perp;
    return (par + perp) * 0.5f;
}
__device__ __forceinline__ float DGGX(float NdH, float alpha)
{
    float alpha2 = alpha * alpha;
    float denom = NdH * NdH * (alpha2 - 1.0f) + 1.0f;
    denom = denom * denom;
    denom *= kPi;
    return alpha2 / denom;
}
__device__ __forceinline__ float LambdaSmith(V3 v, float alpha)
{
    V3 vs = v * v;
    float alpha2 = alpha * alpha;
    float inner = alpha2 * (vs.x + vs.y) / vs.z;
    float lambda = sqrtf(1.0f + inner) - 1.0f;
    return lambda * 0.5f;
}
__device__ __forceinline__ float GSmithSingle(V3 v, float alpha) { return 1.0f / (1.0f + LambdaSmith(v, alpha)); }
__device__ __forceinline__ float GSmithCorrelated(V3 o, V3 i, float alpha)
{
    return 1.0f / (LambdaSmith(o, alpha) + LambdaSmith(i, alpha) + 1.0f);
}
__device__ __forceinline__ V4 FSchlick(float VdH, V4 f0)
{
    float pw = 1.0f - VdH;
    float pw2 = pw * pw;
    float pw5 = pw2 * pw2 * pw;
    V4 r = (Splat4(1.0f) - f0) * pw5;
    return r + f0;
}
__device__ __forceinline__ float VNDFPdf(V3 V, V3 H, float alpha)
{
    float VdH = MaxF(0.0f, Dot(H, V));
    float NdH = MaxF(0.0f, H.z);
    float NdV = MaxF(0.0f, V.z);
    float D = DGGX(NdH, alpha);
    float G = GSmithSingle(V, alpha);
    if(NdV == 0.0f) return 0.0f;
    return VdH * D * G / NdV;
}
__device__ __forceinline__ V3 VNDFSample(V3 V, float a, float xi0, float xi1, float& pdf)
{
    V3 VH = Normalize(V3{a * V.x, a * V.y, V.z});
    float len2 = Dot2(VH.x, VH.y, VH.x, VH.y);
    V3 T1 = (len2 > 0) ? V3{-VH.y, VH.x, 0.0f} * (1.0f / sqrtf(len2)) : V3{1, 0, 0};
    V3 T2 = Cross(VH, T1);
    float r = sqrtf(xi0);
    float phi = 2.0f * kPi * xi1;
    float sinPhi, cosPhi;
    fm::SinCos(phi, sinPhi, cosPhi);
    float t1 = r * cosPhi;
    float t2 = r * sinPhi;
    float s = 0.5f * (1.0f + VH.z);
    t2 = (1.0f - s) * sqrtf(1.0f - t1 * t1) + s * t2;
    float val = 1.0f - t1 * t1 - t2 * t2;
    V3 NH = (t1 * T1 + t2 * T2) + SqrtMax(val) * VH;
    V3 N = {a * NH.x, a * NH.y, SqrtMax(NH.z)};
    float nLen2 = LengthSqr(N);
    if(nLen2 < kEps) N = {0, 0, 1};
    else N = N * (1.0f / sqrtf(nLen2));
    pdf = VNDFPdf(V, N, a);
    return N;
}
__device__ __forceinline__ float BurleyCorr(float NdL, float NdV, float LdH, float roughness)
{
    float Fd90 = 0.5f + 2.0f * roughness * LdH * LdH;
    float pl = 1.0f - NdL, pl2 = pl * pl, pl5 = pl2 * pl2 * pl;
    float pv = 1.0f - NdV, pv2 = pv * pv, pv5 = pv2 * pv2 * pv;
    float fl = 1.0f + (Fd90 - 1.0f) * pl5;
    float fv = 1.0f + (Fd90 - 1.0f) * pv5;
    return fl * fv;
}
__device__ __forceinline__ V4 DivideByPDF(V4 t, float pdf)
{
    if(pdf == 0.0f) return Splat4(0.0f);
    float r = 1.0f / pdf;
    return t * r;
}
__device__ __forceinline__ float Balance2(float p0, float w0, float p1, float w1)
{
    float r = 0.0f;
    r += p0 * w0;
    r += p1 * w1;
    return r;
}
__device__ __forceinline__ float CauchyIoR(float wavelength, const float* c)
{
    float w = wavelength * 1e-3f;
    float w2 = w * w;
    float w4 = w2 * w2;
    return c[0] + c[1] / w2 + c[2] / w4;
}

// ---------------------------------------------------------------------------
// Triangle helpers (Core/ShapeFunctions.h, PrimitiveDefaultTriangle.hpp)
// ---------------------------------------------------------------------------
// Light triangle li's world-space vertices.
__device__ __forceinline__ void LoadLightTri(const SceneView& s, uint32_t li, V3& p0, V3& p1, V3& p2)
{
    const float* v = s.lightPos + size_t(li) * 9u;
    p0 = {v[0], v[1], v[2]};
    p1 = {v[3], v[4], v[5]};
    p2 = {v[6], v[7], v[8]};
}

__device__ __forceinline__ void LoadTri(const SceneView& s, uint32_t t, V3& p0, V3& p1, V3& p2)
{
    uint32_t i0 = s.idx[3 * t], i1 = s.idx[3 * t + 1], i2 = s.idx[3 * t + 2];
    p0 = {s.pos[3 * i0], s.pos[3 * i0 + 1], s.pos[3 * i0 + 2]};
    p1 = {s.pos[3 * i1], s.pos[3 * i1 + 1], s.pos[3 * i1 + 2]};
    p2 = {s.pos[3 * i2], s.pos[3 * i2 + 1], s.pos[3 * i2 + 2]};
}
__device__ __forceinline__ V3 TriNormal(V3 p0, V3 p1, V3 p2) { return Normalize(Cross(p1 - p0, p2 - p0)); }
__device__ __forceinline__ float TriArea(V3 p0, V3 p1, V3 p2) { return Length(Cross(p1 - p0, p2 - p0)) * 0.5f; }
__device__ __forceinline__ bool BaryInRange(float a, float b)
{
    float c = 1.0f - a - b;
    return !(a < 0 || a > 1 || b < 0 || b > 1 || c < 0 || c > 1);
}
__device__ __forceinline__ void ProjectedHit(V3 p0, V3 p1, V3 p2, V3 point, float& ha, float& hb)
{
    V3 n = TriNormal(p0, p1, p2);
    V3 dir = point - p0;
    n = (Dot(n, dir) >= 0.0f) ? n : -n;
    V3 proj = point - Dot(dir, n) * n;
    V3 e0 = p1 - p0, e1 = p2 - p0, v = proj - p0;
    float d00 = Dot(e0, e0), d01 = Dot(e0, e1), d11 = Dot(e1, e1);
    float d20 = Dot(v, e0), d21 = Dot(v, e1);
    float denom = 1.0f / (d00 * d11 - d01 * d01);
    ha = (d11 * d20 - d01 * d21) * denom;
    hb = (d00 * d21 - d01 * d20) * denom;
}

struct Surface
{
    V3 position;
    V3 geoNormal;
    Q4 tbn;
    bool backSide;
};

// Triangle::GenerateSurface(DefaultSurface) (PrimitiveDefaultTriangle.hpp:443-630),
// without the ray-cone / texture-derivative terms (no texture is sampled):
// position, geometric normal and interpolated TBN in the mesh's own space.
__device__ __forceinline__ Surface GenSurfaceRaw(const SceneView& s, uint32_t tri, float a, float b)
{
    uint32_t i0 = s.idx[3 * tri], i1 = s.idx[3 * tri + 1], i2 = s.idx[3 * tri + 2];
    V3 p0 = {s.pos[3 * i0], s.pos[3 * i0 + 1], s.pos[3 * i0 + 2]};
    V3 p1 = {s.pos[3 * i1], s.pos[3 * i1 + 1], s.pos[3 * i1 + 2]};
    V3 p2 = {s.pos[3 * i2], s.pos[3 * i2 + 1], s.pos[3 * i2 + 2]};
    float c = 1.0f - a - b;
    Surface r;
    r.position = (p0 * a + p1 * b) + p2 * c;
    r.geoNormal = TriNormal(p0, p1, p2);
    float4 f0 = s.tbn[i0], f1 = s.tbn[i1], f2 = s.tbn[i2];
    Q4 q0 = {f0.x, f0.y, f0.z, f0.w}, q1 = {f1.x, f1.y, f1.z, f1.w}, q2 = {f2.x, f2.y, f2.z, f2.w};
    r.tbn = QNormalize(QBarySLerp(q0, q1, q2, a, b));
    r.backSide = false;
    return r;
}

// Back-side hits flip the geometric normal and the tangent frame.
__device__ __forceinline__ Surface OrientSurface(Surface r, V3 rayDir)
{
    float dDotN = Dot(r.geoNormal, Normalize(rayDir));
    bool backSide = dDotN > 0.0f;
    if(backSide)
    {
        r.geoNormal = -r.geoNormal;
        r.tbn = QMul(Q4{0, 1, 0, 0}, r.tbn);
    }
    r.backSide = backSide;
    return r;
}

__device__ __forceinline__ Surface GenSurface(const SceneView& s, uint32_t tri, float a, float b,
                                              V3 rayDir)
{
    return OrientSurface(GenSurfaceRaw(s, tri, a, b), rayDir);
}

// The surface of an instanced mesh triangle in world space: position by
// the TRS transform, the geometric normal by its rotation q (and, with a
// non-uniform scale, by the inverse scale first, as the normal of the
// transformed vertices); the
// shading frame stays the mesh's, as the reference leaves it
// (PrimitiveDefaultTriangle.hpp:611-614), unless the instance rotates its
// frames (world -> tangent = mesh tangent frame after q^-1).
__device__ __forceinline__ Surface GenSurfaceInst(const SceneView& s, uint32_t tri, float a, float b,
                                                  V3 rayDir, const InstGPU& I)
{
    Surface r = GenSurfaceRaw(s, tri, a, b);
    const Q4 q = {I.rot.x, I.rot.y, I.rot.z, I.rot.w};
    const V3 t = {I.transScale.x, I.transScale.y, I.transScale.z};
    const V3 sc = {I.scale.x, I.scale.y, I.scale.z};
    r.position = QRotate(q, V3{r.position.x * sc.x, r.position.y * sc.y, r.position.z * sc.z}) + t;
    if(I.transScale.w != 0.0f) r.geoNormal = QRotate(q, r.geoNormal);
    else
        r.geoNormal = Normalize(QRotate(q, V3{r.geoNormal.x / sc.x, r.geoNormal.y / sc.y,
                                              r.geoNormal.z / sc.z}));
    if(I.flags & MRAY_INST_ROTATE_FRAMES) r.tbn = QMul(r.tbn, Q4{q.w, -q.x, -q.y, -q.z});
    return OrientSurface(r, rayDir);
}

// ---------------------------------------------------------------------------
// Materials (Tracer/MaterialsDefault.hpp)
// ---------------------------------------------------------------------------
struct MatInst
{
    uint32_t type;
    V4 albedo;
    float roughness, metallic, specular;
    float frontIoR, backIoR;
};

__device__ __forceinline__ float UnrealMISRatio(const MatInst& m, float avgAlbedo)
{
    float integralDiffuse = 2.0f * kPi * avgAlbedo;
    integralDiffuse *= (1.0f - m.metallic);
    float specularRatio = m.specular * (1.0f - m.metallic) + avgAlbedo * m.metallic;
    float total = specularRatio + integralDiffuse;
    return (total == 0.0f) ? 0.0f : integralDiffuse / total;
}
__device__ __forceinline__ V4 UnrealF0(const MatInst& m)
{
    float specOut = m.specular * 0.08f;
    return Splat4(specOut) * (1.0f - m.metallic) + m.albedo * m.metallic;
}
__device__ __forceinline__ float Specularity(const MatInst& m)
{
    if(m.type == MRAY_MAT_LAMBERT) return 0.0f;
    if(m.type == MRAY_MAT_REFLECT || m.type == MRAY_MAT_REFRACT) return 1.0f;
    float avg = Sum4(m.albedo) * 0.3333f;
    return 1.0f - UnrealMISRatio(m, avg);
}
__device__ __forceinline__ bool IsSpecular(float s) { return s >= 0.95f; }
__device__ __forceinline__ uint32_t MatRNCount(uint32_t type)
{
    return (type == MRAY_MAT_LAMBERT) ? 2u : (type == MRAY_MAT_UNREAL) ? 3u : 0u;
}

struct BxDFSample
{
    V3 wIPos, wIDir;
    float pdf;
    V4 refl;
    bool passedThrough, dispersed;
};

__device__ __forceinline__ BxDFSample SampleBxDF(const MatInst& m, const Surface& sf, V3 wO,
                                                 const float* rn)
{
    BxDFSample out;
    out.passedThrough = false;
    out.dispersed = false;
    out.wIPos = sf.position;
    if(m.type == MRAY_MAT_LAMBERT)
    {
        float pdf;
        V3 d = SampleCosDir(rn[0], rn[1], pdf);
        float nDotL = MaxF(d.z, 0.0f);
        out.refl = m.albedo * nDotL * kInvPi;
        out.wIDir = QInvRotate(sf.tbn, d);
        out.pdf = pdf;
        return out;
    }
    if(m.type == MRAY_MAT_REFLECT)
    {
        V3 n = QBasisZ(sf.tbn);
        out.wIDir = Normalize(Reflect(n, wO));
        out.pdf = 1.0f;
        out.refl = Splat4(1.0f);
        return out;
    }
    if(m.type == MRAY_MAT_REFRACT)
    {
        float fromEta = m.frontIoR, toEta = m.backIoR;
        if(sf.backSide) { float t = fromEta; fromEta = toEta; toEta = t; }
        V3 nL = QBasisZ(sf.tbn);
        float cosTheta = fabsf(Dot(wO, nL));
        float f = FresnelDielectric(cosTheta, fromEta, toEta);
        bool doRefl = (rn[0] < f);
        V3 wI;
        if(doRefl) wI = Reflect(nL, wO);
        else Refract(wI, nL, wO, fromEta, toEta);
        float pdf = doRefl ? f : (1.0f - f);
        out.wIDir = wI;
        out.wIPos = Nudge(sf.position, nL);
        out.pdf = pdf;
        out.refl = Splat4(pdf);
        out.passedThrough = !doRefl;
        out.dispersed = !doRefl;
        return out;
    }
    // Unreal
    float alpha = m.roughness * m.roughness;
    float avg = Sum4(m.albedo) * 0.3333f;
    float sXi = rn[0];
    float misRatio = UnrealMISRatio(m, avg);
    float w0 = misRatio, w1 = 1.0f - misRatio;
    bool doDiffuse = (sXi < misRatio);
    V3 V = QRotate(sf.tbn, wO);
    V3 L, H;
    float pdfD, pdfS;
    if(doDiffuse)
    {
        L = SampleCosDir(rn[1], rn[2], pdfD);
        H = Normalize(L + V);
        float VdH = MaxF(0.0f, Dot(V, H));
        pdfS = VNDFPdf(V, H, alpha);
        pdfS = (VdH == 0.0f) ? 0.0f : (pdfS / (4.0f * VdH));
    }
    else
    {
        float sp;
        H = VNDFSample(V, alpha, rn[1], rn[2], sp);
        L = Reflect(H, V);
        float VdH = MaxF(0.0f, Dot(V, H));
        pdfS = (VdH == 0.0f) ? 0.0f : (sp / (4.0f * VdH));
        pdfD = PdfCosDir(L);
    }
    float VdH = MaxF(0.0f, Dot(V, H));
    float LdH = MaxF(0.0f, Dot(L, H));
    float NdH = MaxF(0.0f, H.z);
    float NdV = MaxF(0.0f, V.z);
    float D = DGGX(NdH, alpha);
    float G = GSmithCorrelated(V, L, alpha);
    G = (LdH == 0.0f) ? 0.0f : G;
    G = (VdH == 0.0f) ? 0.0f : G;
    V4 F = FSchlick(VdH, UnrealF0(m));
    V4 spec = D * F * G * 0.25f / NdV;
    spec = (NdV == 0.0f) ? Splat4(0.0f) : spec;
    if(isinf(D) || isnan(D))
    {
        spec = G * F / GSmithSingle(V, alpha);
        pdfS = 1.0f;
    }
    float NdL = MaxF(0.0f, L.z);
    V4 diffAlb = (1.0f - m.metallic) * m.albedo;
    V4 diff = NdL * diffAlb * kInvPi;
    diff = diff * BurleyCorr(NdL, NdV, LdH, m.roughness);
    out.pdf = Balance2(pdfD, w0, pdfS, w1);
    out.wIDir = QInvRotate(sf.tbn, L);
    out.refl = diff + spec;
    return out;
}

__device__ __forceinline__ V4 EvaluateBxDF(const MatInst& m, const Surface& sf, V3 wI, V3 wO,
                                           bool& passedThrough, bool& dispersed)
{
    passedThrough = false;
    dispersed = false;
    if(m.type == MRAY_MAT_LAMBERT)
    {
        V3 wl = QRotate(sf.tbn, wI);
        float nDotL = MaxF(wl.z, 0.0f);
        return nDotL * m.albedo * kInvPi;
    }
    if(m.type == MRAY_MAT_REFLECT) return Splat4(1.0f);
    if(m.type == MRAY_MAT_REFRACT)
    {
        bool isFront = Dot(sf.geoNormal, wI) > 0.0f;
        passedThrough = !isFront;
        dispersed = !isFront;
        return Splat4(1.0f);
    }
    float alpha = m.roughness * m.roughness;
    V3 V = QRotate(sf.tbn, wO);
    V3 L = QRotate(sf.tbn, wI);
    V3 H = Normalize(L + V);
    float LdH = MaxF(0.0f, Dot(L, H));
    float VdH = MaxF(0.0f, Dot(V, H));
    float NdH = MaxF(0.0f, H.z);
    float NdV = MaxF(0.0f, V.z);
    float D = DGGX(NdH, alpha);
    D = (isnan(D) || isinf(D)) ? 0.0f : D;
    float G = GSmithCorrelated(V, L, alpha);
    G = (LdH == 0.0f) ? 0.0f : G;
    G = (VdH == 0.0f) ? 0.0f : G;
    V4 F = FSchlick(VdH, UnrealF0(m));
    V4 spec = D * F * G * 0.25f / NdV;
    spec = (NdV == 0.0f) ? Splat4(0.0f) : spec;
    float NdL = MaxF(0.0f, L.z);
    V4 diffAlb = (1.0f - m.metallic) * m.albedo;
    V4 diff = NdL * diffAlb * kInvPi;
    diff = diff * BurleyCorr(NdL, NdV, LdH, m.roughness);
    return diff + spec;
}

__device__ __forceinline__ float PdfBxDF(const MatInst& m, const Surface& sf, V3 wI, V3 wO)
{
    if(m.type == MRAY_MAT_LAMBERT)
    {
        V3 wl = QRotate(sf.tbn, wI);
        return MaxF(PdfCosDir(wl), 0.0f);
    }
    if(m.type != MRAY_MAT_UNREAL) return 0.0f;
    float alpha = m.roughness * m.roughness;
    float avg = Sum4(m.albedo) * 0.3333f;
    V3 V = QRotate(sf.tbn, wO);
    V3 L = QRotate(sf.tbn, wI);
    V3 H = Normalize(L + V);
    float misRatio = UnrealMISRatio(m, avg);
    float w0 = misRatio, w1 = 1.0f - misRatio;
    float p0 = PdfCosDir(L);
    float NdH = MaxF(0.0f, H.z);
    float VdH = MaxF(0.0f, Dot(V, H));
    float D = DGGX(NdH, alpha);
    float p1 = VNDFPdf(V, H, alpha);
    p1 = (isnan(D) || isinf(D)) ? 0.0f : p1;
    p1 = (VdH == 0.0f) ? 0.0f : p1 / (4.0f * VdH);
    return Balance2(p0, w0, p1, w1);
}

// ---------------------------------------------------------------------------
// Camera (CamerasDefault.hpp) -- basis built on the host, same formulas
// ---------------------------------------------------------------------------
struct CameraGPU
{
    float position[3], right[3], up[3], bottomLeft[3];
    float planeSize[2];
    float nearFar[2];
    float sigma;
    uint32_t width, height;
};

} // namespace mray
