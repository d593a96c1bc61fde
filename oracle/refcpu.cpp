// refcpu -- CPU restatement of yalcinerbora/mray's wavefront path-tracing hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity oracle and the CPU baseline
// ("CPU restatement of reference TracerCPU semantics").  It is loaded by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the product.
//
// Every function names the reference file:line it restates (paths relative to the
// reference's Source/ directory).  Floating-point policy: built with
// -ffp-contract=off; FMAs appear only where the reference calls Math::FMA
// explicitly (Dot, Cross: Core/Math.h:1586-1632), so every operation is a single
// IEEE-754 rounding that the gfx950 kernels reproduce.
//
// Parity pinning (what checks each part):
//   Morton 63/30       reference golden words T_GraphicsFunctions.cpp:147-305,
//                      brute-force bit interleave (tests/test_oracle_morton.py)
//   LBVH topology      Karras invariants, and node for node (children,
//                      parents, numbering) an independent top-down Karras
//                      construction (tests/test_oracle_lbvh.py); no
//                      reference vector exists
//   closest / any hit  brute force over all triangles (same Moller-Trumbore)
//   PCG32 / seeding    std::mt19937 standard values + closed form
//   camera / filter    closed forms (tests/test_oracle_sampling.py)
//   materials          furnace, reciprocity, pdf integrals (tests/test_oracle_bsdf.py)
//   path tracer        closed-form scenes: white furnace, directly seen emitter
//                      (tests/test_oracle_pt.py)
//   colour data        CIE 1931 observer and D65 values of the standard at
//                      5-nm knots, D65 chromaticity, trapezoid integrals
//                      (tests/test_spectral.py)
//   spectral LUT fit   round trip of T_Spectrum.cu:54-236; the coefficient
//                      values are parity unpinned (the reference generates
//                      its table at build time and ships none)

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "../include/mray_amd.h"
#include "../mray_amd/csrc/colour_data.h"
#include "../mray_amd/csrc/fmath.h"

namespace rc
{
namespace fm = mray::fm;   // the elementary functions the GPU kernels also use

// ------------------------------------------------------------------------
// Vector math (Core/Vector.hpp, Core/Math.h)
// ------------------------------------------------------------------------
struct V2 { float x, y; };
struct V3 { float x, y, z; };
struct V4 { float x, y, z, w; };

static inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
static inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
static inline V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
static inline V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
static inline V3 operator/(V3 a, V3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }

static inline V4 operator+(V4 a, V4 b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
static inline V4 operator-(V4 a, V4 b) { return {a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; }
static inline V4 operator*(V4 a, V4 b) { return {a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w}; }
static inline V4 operator*(V4 a, float s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
static inline V4 operator*(float s, V4 a) { return {s * a.x, s * a.y, s * a.z, s * a.w}; }
static inline V4 operator/(V4 a, float s) { return {a.x / s, a.y / s, a.z / s, a.w / s}; }
static inline V4 Splat4(float s) { return {s, s, s, s}; }
static inline float Sum4(V4 a) { return a.x + a.y + a.z + a.w; }

// Math::Dot (Math.h:1586-1596): r = FMA(v0[i], v1[i], r) from r = 0
static inline float Dot(V3 a, V3 b)
{
    float r = 0.0f;
    r = std::fma(a.x, b.x, r);
    r = std::fma(a.y, b.y, r);
    r = std::fma(a.z, b.z, r);
    return r;
}
static inline float Dot2(V2 a, V2 b)
{
    float r = 0.0f;
    r = std::fma(a.x, b.x, r);
    r = std::fma(a.y, b.y, r);
    return r;
}
static inline float Dot4(V4 a, V4 b)
{
    float r = 0.0f;
    r = std::fma(a.x, b.x, r);
    r = std::fma(a.y, b.y, r);
    r = std::fma(a.z, b.z, r);
    r = std::fma(a.w, b.w, r);
    return r;
}
// Math::Cross (Math.h:1627-1632)
static inline V3 Cross(V3 a, V3 b)
{
    return {std::fma(a.y, b.z, -a.z * b.y),
            std::fma(a.z, b.x, -a.x * b.z),
            std::fma(a.x, b.y, -a.y * b.x)};
}
// Math::Length is Sqrt(Dot(v, v)) on the host path (Math.h:1598-1611)
static inline float Length(V3 v) { return std::sqrt(Dot(v, v)); }
static inline float LengthSqr(V3 v) { return Dot(v, v); }
// Math::Normalize (Math.h:1618-1624)
static inline V3 Normalize(V3 v)
{
    float lRecip = 1.0f / Length(v);
    return v * lRecip;
}
static inline V4 Normalize4(V4 v)
{
    float lRecip = 1.0f / std::sqrt(Dot4(v, v));
    return v * lRecip;
}
static inline float MaxF(float a, float b) { return std::fmax(a, b); }
static inline float MinF(float a, float b) { return std::fmin(a, b); }
static inline float Clamp(float v, float lo, float hi) { return MinF(MaxF(v, lo), hi); }
static inline float SqrtMax(float x) { return std::sqrt(MaxF(x, 0.0f)); }
static inline float RSqrt(float x) { return 1.0f / std::sqrt(x); }
static inline float SignPM1(float x) { return std::signbit(x) ? -1.0f : 1.0f; }

// MathConstants (Core/MathConstants.h)
static constexpr float kPi = 3.14159265358979323846f;
static constexpr float kInvPi = 1.0f / kPi;
static constexpr float kSqrt2 = 1.41421356237309504880f;
static constexpr float kSqrtPi = 1.77245385090551602730f;
static constexpr float kSmallEps = 1e-7f;
static constexpr float kEps = 1e-5f;
static constexpr float kLargeEps = 1e-4f;

// ------------------------------------------------------------------------
// Quaternion (Core/Quaternion.hpp)
// ------------------------------------------------------------------------
struct Quat { float w, x, y, z; };

// Quat::operator* (Quaternion.hpp:68-75)
static inline Quat QMul(Quat a, Quat b)
{
    return {a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z,
            a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
            a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
            a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w};
}
static inline float QDot(Quat a, Quat b)
{
    return Dot4(V4{a.w, a.x, a.y, a.z}, V4{b.w, b.x, b.y, b.z});
}
static inline Quat QScale(Quat a, float s) { return {a.w * s, a.x * s, a.y * s, a.z * s}; }
static inline Quat QAdd(Quat a, Quat b) { return {a.w + b.w, a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline Quat QNormalize(Quat q)
{
    V4 v = Normalize4(V4{q.w, q.x, q.y, q.z});
    return {v.x, v.y, v.z, v.w};
}
static inline Quat QConj(Quat q) { return {q.w, -q.x, -q.y, -q.z}; }
// ApplyRotation (Quaternion.hpp:181-205): t = 2 cross(u, v); r = v + w t + cross(u, t)
static inline V3 QRotate(Quat q, V3 v)
{
    V3 u = {q.x, q.y, q.z};
    V3 t = Cross(u, v) * 2.0f;
    V3 r = v + q.w * t + Cross(u, t);
    return r;
}
static inline V3 QInvRotate(Quat q, V3 v) { return QRotate(QConj(q), v); }
// OrthoBasisZ (Quaternion.hpp:256-271)
static inline V3 QBasisZ(Quat q)
{
    float v00 = q.w * q.w, v01 = q.w * q.x, v02 = q.w * q.y;
    float v11 = q.x * q.x, v13 = q.x * q.z, v22 = q.y * q.y;
    float v23 = q.y * q.z, v33 = q.z * q.z;
    float X = v13 - v02 + v13 - v02;
    float Y = v23 + v01 + v23 + v01;
    float Z = v00 - v11 - v22 + v33;
    return {X, Y, Z};
}
// SLerp (Quaternion.hpp:289-316)
static inline Quat QSLerp(Quat s, Quat e, float t)
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
    return QAdd(QScale(s, s0), QScale(e, s1));
}
// BarySLerp (Quaternion.hpp:319-352)
static inline Quat QBarySLerp(Quat q0, Quat q1, Quat q2, float a, float b)
{
    float c = (1 - a - b);
    if(std::fabs(a + b) < kEps) return q2;
    float ab = a / (a + b);
    Quat qAB = QSLerp(q1, q0, ab);
    return QSLerp(qAB, q2, c);
}

// ------------------------------------------------------------------------
// Morton codes (Core/GraphicsFunctions.h:573-625)
// ------------------------------------------------------------------------
static inline uint32_t Expand3D_32(uint32_t x)
{
    x &= 0x000003ff;
    x = (x ^ (x << 16)) & 0xff0000ff;
    x = (x ^ (x << 8)) & 0x0300f00f;
    x = (x ^ (x << 4)) & 0x030c30c3;
    x = (x ^ (x << 2)) & 0x09249249;
    return x;
}
static inline uint64_t Expand3D_64(uint32_t v)
{
    uint64_t x = v;
    x &= 0x1fffff;
    x = (x | x << 32) & 0x001f00000000ffffull;
    x = (x | x << 16) & 0x001f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}
static inline uint32_t Shrink3D_64(uint64_t x)
{
    x &= 0x1249249249249249ull;
    x = (x ^ (x >> 2)) & 0x30c30c30c30c30c3ull;
    x = (x ^ (x >> 4)) & 0xf00f00f00f00f00full;
    x = (x ^ (x >> 8)) & 0x00ff0000ff0000ffull;
    x = (x ^ (x >> 16)) & 0x00ff00000000ffffull;
    x = (x ^ (x >> 32)) & 0x00000000001fffffull;
    return uint32_t(x);
}
static inline uint32_t Shrink3D_32(uint32_t x)
{
    x &= 0x09249249;
    x = (x ^ (x >> 2)) & 0x030c30c3;
    x = (x ^ (x >> 4)) & 0x0300f00f;
    x = (x ^ (x >> 8)) & 0xff0000ff;
    x = (x ^ (x >> 16)) & 0x000003ff;
    return x;
}
static inline uint64_t Compose3D_64(uint32_t x, uint32_t y, uint32_t z)
{
    return Expand3D_64(x) | (Expand3D_64(y) << 1) | (Expand3D_64(z) << 2);
}
static inline uint32_t Compose3D_32(uint32_t x, uint32_t y, uint32_t z)
{
    return Expand3D_32(x) | (Expand3D_32(y) << 1) | (Expand3D_32(z) << 2);
}

// Triangle centroid: Triangle::GetCenter (PrimitiveDefaultTriangle.hpp:107-115)
static inline V3 TriCenter(V3 p0, V3 p1, V3 p2)
{
    const float k = 0.333333333f;
    return (p0 * k + p1 * k) + p2 * k;
}

// KCGenMortonCode body (AcceleratorLBVH.cu:135-165).  float diff, double
// product with 2^21/maxSide, back to float, lroundf, clamp to [0, 2^21-1].
static inline uint64_t Morton63Of(V3 center, V3 aabbMin, V3 aabbMax)
{
    V3 span = aabbMax - aabbMin;
    // AABB3::GeomSpan().Maximum(): index of the first maximal component
    int maxI = 0;
    float maxV = span.x;
    if(span.y > maxV) { maxV = span.y; maxI = 1; }
    if(span.z > maxV) { maxV = span.z; maxI = 2; }
    (void)maxI;
    const double sliceCount = double(1ull << 21);
    double deltaRecip = sliceCount / double(maxV);
    V3 diff = center - aabbMin;
    diff = {MaxF(diff.x, 0.0f), MaxF(diff.y, 0.0f), MaxF(diff.z, 0.0f)};
    float qx = float(double(diff.x) * deltaRecip);
    float qy = float(double(diff.y) * deltaRecip);
    float qz = float(double(diff.z) * deltaRecip);
    // Math::RoundInt -> int32_t(lroundf); Vector3ui(int) then Clamp(0, last)
    const uint32_t last = (1u << 21) - 1u;
    uint32_t ix = uint32_t(int32_t(std::lround(qx)));
    uint32_t iy = uint32_t(int32_t(std::lround(qy)));
    uint32_t iz = uint32_t(int32_t(std::lround(qz)));
    ix = std::min(std::max(ix, 0u), last);
    iy = std::min(std::max(iy, 0u), last);
    iz = std::min(std::max(iz, 0u), last);
    return Compose3D_64(ix, iy, iz);
}

// The 30-bit variant: same quantisation into 2^10 slices, Compose3D<uint32_t>.
static inline uint32_t Morton30Of(V3 center, V3 aabbMin, V3 aabbMax)
{
    V3 span = aabbMax - aabbMin;
    float maxV = span.x;
    if(span.y > maxV) maxV = span.y;
    if(span.z > maxV) maxV = span.z;
    const double sliceCount = double(1u << 10);
    double deltaRecip = sliceCount / double(maxV);
    V3 diff = center - aabbMin;
    diff = {MaxF(diff.x, 0.0f), MaxF(diff.y, 0.0f), MaxF(diff.z, 0.0f)};
    float qx = float(double(diff.x) * deltaRecip);
    float qy = float(double(diff.y) * deltaRecip);
    float qz = float(double(diff.z) * deltaRecip);
    const uint32_t last = (1u << 10) - 1u;
    uint32_t ix = std::min(uint32_t(int32_t(std::lround(qx))), last);
    uint32_t iy = std::min(uint32_t(int32_t(std::lround(qy))), last);
    uint32_t iz = std::min(uint32_t(int32_t(std::lround(qz))), last);
    return Compose3D_32(ix, iy, iz);
}

// ------------------------------------------------------------------------
// Karras LBVH (AcceleratorLBVH.cu:69-94 Delta, 170-299 internal nodes,
// 301-435 bottom-up refit).  Node child words: bit 31 = leaf flag
// (ChildIndex IS_LEAF, AcceleratorLBVH.h:48-69), low 31 bits = index.
// ------------------------------------------------------------------------
static constexpr uint32_t kLeafFlag = 0x80000000u;
static constexpr uint32_t kInvalid = 0xFFFFFFFFu;

// strict = 0 is the reference's equal-code fallback (clz of the index XOR);
// strict = 1 ranks equal codes as the (code, index) key does: 64 + clz32(i^j).
// The reference form can break the Karras prefix order when a run of equal
// codes neighbours codes that differ only in low bits; builds that produce a
// malformed tree are redone with strict = 1 (ValidateTree).
static inline int32_t Delta(const uint64_t* codes, int32_t n, int32_t i, int32_t j, int strict)
{
    if(j < 0 || j >= n) return -1;
    uint64_t l = codes[i], r = codes[j];
    if(l == r)
    {
        if(strict) return 64 + int32_t(__builtin_clz(uint32_t(i ^ j)));
        l = uint64_t(i);
        r = uint64_t(j);
    }
    uint64_t d = l ^ r;
    return (d == 0) ? 64 : int32_t(__builtin_clzll(d));
}

static inline int32_t DivideUpI(int32_t v, int32_t d) { return (v + d - 1) / d; }

static void KarrasBuild(const uint64_t* codes, const uint32_t* leafOrder, int32_t n,
                        uint32_t* nodes /*[n-1][3]*/, uint32_t* leafParent /*[n]*/,
                        uint32_t* range /*[n-1][2] or null*/, int strict)
{
    if(n == 1)
    {
        nodes[0] = kLeafFlag | 0u;
        nodes[1] = kInvalid;
        nodes[2] = kInvalid;
        leafParent[leafOrder[0]] = 0;
        if(range) { range[0] = 0; range[1] = 0; }
        return;
    }
    int32_t nodeCount = n - 1;
    for(int32_t i = 0; i < nodeCount; i++)
    {
        int32_t diff = Delta(codes, n, i, i + 1, strict) - Delta(codes, n, i, i - 1, strict);
        int32_t d = (diff < 0) ? -1 : 1;
        int32_t deltaMin = Delta(codes, n, i, i - d, strict);
        int32_t lMax = 2;
        while(Delta(codes, n, i, i + lMax * d, strict) > deltaMin) lMax <<= 1;
        int32_t l = 0;
        for(int32_t t = lMax >> 1; t != 0; t >>= 1)
            if(Delta(codes, n, i, i + (l + t) * d, strict) > deltaMin) l += t;
        int32_t j = i + l * d;
        int32_t s = 0;
        int32_t deltaNode = Delta(codes, n, i, j, strict);
        auto AdvanceCeil = [](int32_t x) { return (x == 1) ? 0 : DivideUpI(x, 2); };
        for(int32_t t = DivideUpI(l, 2); t != 0; t = AdvanceCeil(t))
            if(Delta(codes, n, i, i + (s + t) * d, strict) > deltaNode) s += t;
        int32_t gamma = i + s * d + std::min(d, 0);

        uint32_t* me = nodes + size_t(i) * 3;
        if(std::min(i, j) == gamma)
        {
            uint32_t leaf = leafOrder[gamma];
            me[0] = kLeafFlag | leaf;
            leafParent[leaf] = uint32_t(i);
        }
        else
        {
            me[0] = uint32_t(gamma);
            nodes[size_t(gamma) * 3 + 2] = uint32_t(i);
        }
        if(std::max(i, j) == gamma + 1)
        {
            uint32_t leaf = leafOrder[gamma + 1];
            me[1] = kLeafFlag | leaf;
            leafParent[leaf] = uint32_t(i);
        }
        else
        {
            me[1] = uint32_t(gamma + 1);
            nodes[size_t(gamma + 1) * 3 + 2] = uint32_t(i);
        }
        if(i == 0) me[2] = kInvalid;
        if(range)
        {
            range[2 * i] = uint32_t(std::min(i, j));
            range[2 * i + 1] = uint32_t(std::max(i, j));
        }
    }
}

// Every internal node and leaf reached once from the root; each node's
// children split its leaf range at one point.
static bool ValidateTree(const uint32_t* nodes, const uint32_t* range, uint32_t n)
{
    if(n <= 1) return true;
    uint32_t nodeCount = n - 1;
    if(range[0] != 0 || range[1] != n - 1) return false;
    std::vector<uint8_t> seenN(nodeCount, 0), seenL(n, 0);
    std::vector<uint32_t> st{0u};
    uint32_t vn = 0, vl = 0;
    while(!st.empty())
    {
        uint32_t v = st.back(); st.pop_back();
        if(v >= nodeCount || seenN[v]) return false;
        seenN[v] = 1; vn++;
        uint32_t lo = range[2 * v], hi = range[2 * v + 1];
        if(lo >= hi || hi >= n) return false;
        uint32_t l = nodes[3 * size_t(v)], r = nodes[3 * size_t(v) + 1], mid;
        if(l & kLeafFlag)
        {
            uint32_t t = l & ~kLeafFlag;
            if(t >= n || seenL[t]) return false;
            seenL[t] = 1; vl++; mid = lo;
        }
        else
        {
            if(l >= nodeCount || range[2 * l] != lo) return false;
            mid = range[2 * l + 1];
            st.push_back(l);
        }
        if(r & kLeafFlag)
        {
            uint32_t t = r & ~kLeafFlag;
            if(t >= n || seenL[t] || hi != mid + 1) return false;
            seenL[t] = 1; vl++;
        }
        else
        {
            if(r >= nodeCount || range[2 * r] != mid + 1 || range[2 * r + 1] != hi) return false;
            st.push_back(r);
        }
    }
    return vn == nodeCount && vl == n;
}

static inline void UnionBox(float* dst, const float* a, const float* b)
{
    dst[0] = MinF(a[0], b[0]); dst[1] = MinF(a[1], b[1]); dst[2] = MinF(a[2], b[2]);
    dst[3] = MaxF(a[3], b[3]); dst[4] = MaxF(a[4], b[4]); dst[5] = MaxF(a[5], b[5]);
}

// Bottom-up refit: a node's box is the union of its children's boxes
// (the atomic second-arrival scheme of AcceleratorLBVH.cu:301-435 computes
// the same union; order is irrelevant for min/max).
static void RefitBoxes(const uint32_t* nodes, int32_t nodeCount, const float* leafAabbs,
                       float* boxes)
{
    // Post-order via explicit stack from the root
    std::vector<int32_t> order;
    order.reserve(size_t(nodeCount));
    std::vector<int32_t> st;
    st.push_back(0);
    while(!st.empty())
    {
        int32_t v = st.back(); st.pop_back();
        order.push_back(v);
        const uint32_t* nd = nodes + size_t(v) * 3;
        for(int c = 0; c < 2; c++)
            if(nd[c] != kInvalid && !(nd[c] & kLeafFlag)) st.push_back(int32_t(nd[c]));
    }
    for(auto it = order.rbegin(); it != order.rend(); ++it)
    {
        int32_t v = *it;
        const uint32_t* nd = nodes + size_t(v) * 3;
        const float* cb[2];
        for(int c = 0; c < 2; c++)
        {
            if(nd[c] == kInvalid) cb[c] = nullptr;
            else if(nd[c] & kLeafFlag) cb[c] = leafAabbs + size_t(nd[c] & ~kLeafFlag) * 6;
            else cb[c] = boxes + size_t(nd[c]) * 6;
        }
        float* dst = boxes + size_t(v) * 6;
        if(cb[1] == nullptr) std::memcpy(dst, cb[0], 6 * sizeof(float));
        else UnionBox(dst, cb[0], cb[1]);
    }
}

// ------------------------------------------------------------------------
// Ray / triangle (Core/Ray.hpp:121-167) and slab test (Ray.hpp:192-219)
// ------------------------------------------------------------------------
struct TriHit { bool hit; bool backFace; float t, a, b; };

static inline TriHit IntersectTriangle(V3 pos, V3 dir, V3 t0, V3 t1, V3 t2, bool cullFace)
{
    TriHit r{false, false, 0, 0, 0};
    V3 e0 = t1 - t0;
    V3 e1 = t2 - t0;
    V3 p = Cross(dir, e1);
    float det = Dot(e0, p);
    bool backfaceHit = (det < kSmallEps);
    bool nearlyParallel = (std::fabs(det) < kSmallEps);
    if((cullFace && backfaceHit) || nearlyParallel) return r;
    float invDet = 1.0f / det;
    V3 tVec = pos - t0;
    float u = Dot(tVec, p) * invDet;
    if(u < 0 || u > 1) return r;
    V3 qVec = Cross(tVec, e0);
    float v = Dot(dir, qVec) * invDet;
    if((v < 0) || (v + u) > 1) return r;
    float t = Dot(e1, qVec) * invDet;
    if(t <= kSmallEps) return r;
    float c = 1 - u - v;
    // baryCoords reordered to (c, u, v); the hit keeps (c, u)
    r.hit = true;
    r.backFace = backfaceHit;
    r.t = t;
    r.a = c;
    r.b = u;
    return r;
}

// Ray::IntersectsAABB (Ray.hpp:192-219): slab test with invD = 1/dir and a
// swap on negative directions.  Restated with two changes that only affect
// pruning, never which triangle is hit: the near/far plane is chosen by the
// sign bit of invD before the subtraction, so a ray lying exactly in a slab
// plane (0 * inf = NaN) leaves that slab unbounded (std::fmax / std::fmin drop
// NaN) instead of being rejected; and the far bound is widened by 1 + 2*gamma(4)
// (Ize 2013) so rounding never prunes a box the ray touches.  Closest hits
// therefore equal the brute-force ones.
static inline bool IntersectAABB(V3 pos, V3 dir, const float* box, float tMin, float tMax)
{
    float o[3] = {pos.x, pos.y, pos.z};
    float id[3] = {1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z};
    float t0v = tMin, t1v = tMax;
    for(int i = 0; i < 3; i++)
    {
        bool neg = std::signbit(id[i]);
        float nearP = neg ? box[3 + i] : box[i];
        float farP = neg ? box[i] : box[3 + i];
        t0v = MaxF(t0v, (nearP - o[i]) * id[i]);
        t1v = MinF(t1v, (farP - o[i]) * id[i]);
    }
    const float grow = 1.0f + 2.0f * 5.96e-8f * 4.0f;
    return t1v * grow >= t0v;
}

// Ray::Nudge (Ray.hpp:258-300), Ray Tracing Gems I ch. 6
static inline V3 Nudge(V3 p, V3 n)
{
    const float ORIGIN = 1.0f / 32.0f;
    const float FLOAT_SCALE = 1.0f / 65536.0f;
    const float INT_SCALE = 256.0f;
    int32_t ofi[3] = {int32_t(INT_SCALE * n.x), int32_t(INT_SCALE * n.y), int32_t(INT_SCALE * n.z)};
    float pp[3] = {p.x, p.y, p.z};
    float nn[3] = {n.x, n.y, n.z};
    float out[3];
    for(int i = 0; i < 3; i++)
    {
        int32_t pi;
        std::memcpy(&pi, &pp[i], 4);
        pi += (pp[i] < 0.0f) ? -ofi[i] : ofi[i];
        float pf;
        std::memcpy(&pf, &pi, 4);
        out[i] = (std::fabs(pp[i]) < ORIGIN) ? (pp[i] + FLOAT_SCALE * nn[i]) : pf;
    }
    return {out[0], out[1], out[2]};
}

// ------------------------------------------------------------------------
// Scene
// ------------------------------------------------------------------------
struct Material
{
    uint32_t type;
    V3 albedo;
    float roughness, metallic, specular;
    V3 cauchyFront, cauchyBack;
};

struct Scene
{
    uint32_t vCount = 0, tCount = 0;
    std::vector<V3> pos;
    std::vector<Quat> tbn;
    std::vector<uint32_t> idx;
    std::vector<uint32_t> surfKey;
    std::vector<uint8_t> cull;
    std::vector<Material> mats;
    std::vector<uint32_t> lightTri;
    std::vector<V3> lightRad;
    std::vector<uint8_t> lightTwoSided;
    uint32_t boundaryType = 0;
    V3 boundaryRad{0, 0, 0};
    float sceneDiameter = 0;
    MRayCamera cam{};
    // textured Skysphere: texels [envH][envW], luminance CDFs
    uint32_t envW = 0, envH = 0;
    std::vector<V3> envTex;
    std::vector<float> envCdfX;      // [envH][envW], rows normalized
    std::vector<float> envCdfY;      // [envH]
    // NEE light selection by power (used with MRAY_RENDER_LIGHT_POWER)
    std::vector<float> lightProb;
    std::vector<uint32_t> lightAlias;
    std::vector<float> lightPmf;
    // LBVH products
    std::vector<uint64_t> codes;      // triangle order
    std::vector<uint32_t> leafOrder;  // sorted position -> triangle
    std::vector<uint32_t> nodes;      // [n-1][3]
    std::vector<float> boxes;         // [n-1][6]
    std::vector<float> triAabbs;      // [n][6]
    V3 sceneMin{0, 0, 0}, sceneMax{0, 0, 0};
    bool strictTie = false;

    void Tri(uint32_t t, V3& p0, V3& p1, V3& p2) const
    {
        p0 = pos[idx[3 * t + 0]];
        p1 = pos[idx[3 * t + 1]];
        p2 = pos[idx[3 * t + 2]];
    }
};

static void BuildLBVH(Scene& s)
{
    uint32_t n = s.tCount;
    s.triAabbs.resize(size_t(n) * 6);
    std::vector<V3> centers(n);
    float smin[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, smax[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for(uint32_t t = 0; t < n; t++)
    {
        V3 p0, p1, p2;
        s.Tri(t, p0, p1, p2);
        centers[t] = TriCenter(p0, p1, p2);
        // Shape::Triangle::BoundingBox (ShapeFunctions.h:49-57)
        float* b = &s.triAabbs[size_t(t) * 6];
        b[0] = MinF(MinF(p0.x, p1.x), p2.x);
        b[1] = MinF(MinF(p0.y, p1.y), p2.y);
        b[2] = MinF(MinF(p0.z, p1.z), p2.z);
        b[3] = MaxF(MaxF(p0.x, p1.x), p2.x);
        b[4] = MaxF(MaxF(p0.y, p1.y), p2.y);
        b[5] = MaxF(MaxF(p0.z, p1.z), p2.z);
        for(int k = 0; k < 3; k++)
        {
            smin[k] = MinF(smin[k], b[k]);
            smax[k] = MaxF(smax[k], b[3 + k]);
        }
    }
    s.sceneMin = {smin[0], smin[1], smin[2]};
    s.sceneMax = {smax[0], smax[1], smax[2]};
    s.codes.resize(n);
    for(uint32_t t = 0; t < n; t++) s.codes[t] = Morton63Of(centers[t], s.sceneMin, s.sceneMax);
    s.leafOrder.resize(n);
    for(uint32_t t = 0; t < n; t++) s.leafOrder[t] = t;
    std::stable_sort(s.leafOrder.begin(), s.leafOrder.end(),
                     [&](uint32_t a, uint32_t b) { return s.codes[a] < s.codes[b]; });
    std::vector<uint64_t> sorted(n);
    for(uint32_t i = 0; i < n; i++) sorted[i] = s.codes[s.leafOrder[i]];
    uint32_t nodeCount = (n > 1) ? n - 1 : 1;
    s.nodes.assign(size_t(nodeCount) * 3, kInvalid);
    s.boxes.assign(size_t(nodeCount) * 6, 0.0f);
    std::vector<uint32_t> leafParent(n, kInvalid);
    std::vector<uint32_t> range(size_t(nodeCount) * 2, 0u);
    KarrasBuild(sorted.data(), s.leafOrder.data(), int32_t(n), s.nodes.data(), leafParent.data(),
                range.data(), 0);
    s.strictTie = false;
    if(!ValidateTree(s.nodes.data(), range.data(), n))
    {
        s.nodes.assign(size_t(nodeCount) * 3, kInvalid);
        KarrasBuild(sorted.data(), s.leafOrder.data(), int32_t(n), s.nodes.data(),
                    leafParent.data(), range.data(), 1);
        s.strictTie = true;
    }
    RefitBoxes(s.nodes.data(), int32_t(nodeCount), s.triAabbs.data(), s.boxes.data());
}

// Closest hit tie rule: smaller t wins; equal t goes to the smaller triangle
// index.  The reference keeps the first leaf found in left-first order
// (AcceleratorLBVH.hpp:359-370); an order-free rule is needed because the
// GPU's wide BVH visits leaves in another order.
struct Hit { uint32_t prim; float t, a, b; };

struct TravCounters { uint64_t nodes = 0, tris = 0; };

static inline bool Better(float t, uint32_t p, const Hit& h)
{
    return (t < h.t) || (t == h.t && p < h.prim);
}

static inline void TestLeaf(const Scene& s, V3 pos, V3 dir, float tMin, float tMax, uint32_t tri,
                            Hit& best)
{
    V3 p0, p1, p2;
    s.Tri(tri, p0, p1, p2);
    TriHit th = IntersectTriangle(pos, dir, p0, p1, p2, s.cull[tri] != 0);
    if(!th.hit) return;
    // IntersectionCheck IsInRange (AcceleratorLBVH.hpp:233-236) against the
    // original [tMin, tMax)
    if(!(th.t >= tMin && th.t < tMax)) return;
    if(Better(th.t, tri, best)) best = {tri, th.t, th.a, th.b};
}

// TraverseLBVHStack (AcceleratorLBVH.hpp:113-166): stack, left then right
static Hit ClosestHitLBVH(const Scene& s, V3 pos, V3 dir, float tMin, float tMax,
                          TravCounters* ctr)
{
    Hit best{kInvalid, FLT_MAX, 0, 0};
    if(s.tCount == 0) return best;
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while(sp > 0)
    {
        uint32_t ni = stack[--sp];
        if(ni == kInvalid) continue;
        if(ni & kLeafFlag)
        {
            if(ctr) ctr->tris++;
            TestLeaf(s, pos, dir, tMin, tMax, ni & ~kLeafFlag, best);
            continue;
        }
        if(ctr) ctr->nodes++;
        float curMax = (best.prim == kInvalid) ? tMax : MinF(tMax, best.t);
        // boxes are tested with the closed interval so that an equal-t
        // candidate with a smaller index can still be reached
        if(IntersectAABB(pos, dir, &s.boxes[size_t(ni) * 6], tMin, curMax))
        {
            const uint32_t* nd = &s.nodes[size_t(ni) * 3];
            stack[sp++] = nd[1];
            stack[sp++] = nd[0];
        }
    }
    return best;
}

static Hit ClosestHitBrute(const Scene& s, V3 pos, V3 dir, float tMin, float tMax)
{
    Hit best{kInvalid, FLT_MAX, 0, 0};
    for(uint32_t t = 0; t < s.tCount; t++) TestLeaf(s, pos, dir, tMin, tMax, t, best);
    return best;
}

// FirstHit (AcceleratorLBVH.hpp:376-410): any accepted leaf ends traversal
static bool AnyHitLBVH(const Scene& s, V3 pos, V3 dir, float tMin, float tMax)
{
    if(s.tCount == 0) return false;
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while(sp > 0)
    {
        uint32_t ni = stack[--sp];
        if(ni == kInvalid) continue;
        if(ni & kLeafFlag)
        {
            uint32_t tri = ni & ~kLeafFlag;
            V3 p0, p1, p2;
            s.Tri(tri, p0, p1, p2);
            TriHit th = IntersectTriangle(pos, dir, p0, p1, p2, s.cull[tri] != 0);
            if(th.hit && th.t >= tMin && th.t < tMax) return true;
            continue;
        }
        if(IntersectAABB(pos, dir, &s.boxes[size_t(ni) * 6], tMin, tMax))
        {
            const uint32_t* nd = &s.nodes[size_t(ni) * 3];
            stack[sp++] = nd[1];
            stack[sp++] = nd[0];
        }
    }
    return false;
}

// ------------------------------------------------------------------------
// RNG: PermutedCG32 (Tracer/Random.h:761-843), ToFloat01 (Random.h:101-120),
// host seeding (Random.cu:661-703)
// ------------------------------------------------------------------------
static inline uint32_t PcgStep(uint32_t r) { return r * 747796405u + 2891336453u; }
static inline uint32_t PcgPermute(uint32_t s)
{
    uint32_t result = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (result >> 22u) ^ result;
}
static inline uint32_t PcgGenState(uint32_t seed)
{
    uint32_t r = PcgStep(0);
    r += seed;
    return PcgStep(r);
}
static inline uint32_t PcgNext(uint32_t& st)
{
    uint32_t r = PcgPermute(st);
    st = PcgStep(st);
    return r;
}
static inline float ToFloat01(uint32_t v)
{
    const float prevOne = 0x1.fffffep-1f;
    return MinF(float(v) * 0x1p-32f, prevOne);
}

static void SeedStates(uint64_t seed, uint32_t count, uint32_t* out)
{
    uint32_t seed32 = uint32_t((seed >> 32) ^ (seed & 0xFFFFFFFFull));
    std::mt19937 rng(seed32);
    for(uint32_t i = 0; i < count; i++) out[i] = PcgGenState(uint32_t(rng()));
}

// ------------------------------------------------------------------------
// Sampling helpers (Tracer/DistributionFunctions.h, Core/Math.h)
// ------------------------------------------------------------------------
// Math::InvErrFunc host path (Math.h:787-833)
static inline float InvErf(float x)
{
    float t = fm::Log(std::fma(x, 0.0f - x, 1.0f));
    float p;
    if(std::fabs(t) > 6.125f)
    {
        p = 3.03697567e-10f;
        p = std::fma(p, t, 2.93243101e-8f);
        p = std::fma(p, t, 1.22150334e-6f);
        p = std::fma(p, t, 2.84108955e-5f);
        p = std::fma(p, t, 3.93552968e-4f);
        p = std::fma(p, t, 3.02698812e-3f);
        p = std::fma(p, t, 4.83185798e-3f);
        p = std::fma(p, t, -2.64646143e-1f);
        p = std::fma(p, t, 8.40016484e-1f);
    }
    else
    {
        p = 5.438778320e-9f;
        p = std::fma(p, t, 1.43285448e-7f);
        p = std::fma(p, t, 1.22774793e-6f);
        p = std::fma(p, t, 1.12963626e-7f);
        p = std::fma(p, t, -5.61530760e-5f);
        p = std::fma(p, t, -1.47697632e-4f);
        p = std::fma(p, t, 2.31468678e-3f);
        p = std::fma(p, t, 1.15392581e-2f);
        p = std::fma(p, t, -2.32015476e-1f);
        p = std::fma(p, t, 8.86226892e-1f);
    }
    return x * p;
}

// Math::Gaussian (Math.h:1031-1042)
static inline float Gaussian(float x, float sigma)
{
    const float invSqrt2Pi = (1.0f / kSqrt2) * (1.0f / kSqrtPi);
    float sigmaInv = 1.0f / sigma;
    float result = invSqrt2Pi * sigmaInv;
    float p = (x - 0.0f) * sigmaInv;
    result *= fm::Exp(-0.5f * p * p);
    return result;
}

struct Sample1 { float value, pdf; };
// Common::SampleGaussian (DistributionFunctions.h:686-705)
static inline Sample1 SampleGaussian(float xi, float sigma)
{
    float x = kSqrt2 * sigma;
    float e = InvErf(2.0f * xi - 1.0f);
    x = x * e + 0.0f;
    if(std::isinf(e))
    {
        float minMax = 3.5f * sigma;
        x = Clamp(x, -minMax, minMax);
    }
    return {x, Gaussian(x, sigma)};
}

struct DirSample { V3 value; float pdf; };
// Common::SampleCosDirection (DistributionFunctions.h:846-871)
static inline DirSample SampleCosDirection(float xi0, float xi1)
{
    float xi1Angle = 2.0f * kPi * xi1;
    float xi0Sqrt = std::sqrt(xi0);
    float s, c;
    fm::SinCos(xi1Angle, s, c);
    V3 dir;
    dir.x = xi0Sqrt * c;
    dir.y = xi0Sqrt * s;
    dir.z = SqrtMax(1.0f - Dot2(V2{dir.x, dir.y}, V2{dir.x, dir.y}));
    return {dir, dir.z * kInvPi};
}
// Common::PDFCosDirection (DistributionFunctions.h:882-889)
static inline float PDFCosDirection(V3 v)
{
    float pdf = v.z * kInvPi;
    return (pdf <= kEps) ? 0.0f : pdf;
}
// Graphics::Reflect (GraphicsFunctions.h:157-171)
static inline V3 Reflect(V3 n, V3 v) { return 2.0f * Dot(v, n) * n - v; }
// Graphics::Refract (GraphicsFunctions.h:173-203)
static inline bool Refract(V3& out, V3 n, V3 v, float etaFrom, float etaTo)
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
// BxDF::FresnelDielectric (DistributionFunctions.h:381-403)
static inline float FresnelDielectric(float cosFront, float etaFront, float etaBack)
{
    float sinFront = SqrtMax(1.0f - cosFront * cosFront);
    float sinBack = etaFront / etaBack * sinFront;
    if(sinFront >= 1.0f) return 1.0f;
    float cosBack = SqrtMax(1.0f - sinBack * sinBack);
    float par = ((etaBack * cosFront - etaFront * cosBack) /
                 (etaBack * cosFront + etaFront * cosBack));
    par = par * par;
    float perp = ((etaFront * cosFront - etaBack * cosBack) /
                  (etaFront * cosFront + etaBack * cosBack));
    perp = perp * perp;
    return (par + perp) * 0.5f;
}
// BxDF::DGGX (:433-442)
static inline float DGGX(float NdH, float alpha)
{
    float alpha2 = alpha * alpha;
    float denom = NdH * NdH * (alpha2 - 1.0f) + 1.0f;
    denom = denom * denom;
    denom *= kPi;
    return alpha2 / denom;
}
// BxDF::LambdaSmith / GSmithSingle / GSmithCorrelated (:444-474)
static inline float LambdaSmith(V3 v, float alpha)
{
    V3 vs = v * v;
    float alpha2 = alpha * alpha;
    float inner = alpha2 * (vs.x + vs.y) / vs.z;
    float lambda = std::sqrt(1.0f + inner) - 1.0f;
    return lambda * 0.5f;
}
static inline float GSmithSingle(V3 v, float alpha) { return 1.0f / (1.0f + LambdaSmith(v, alpha)); }
static inline float GSmithCorrelated(V3 wO, V3 wI, float alpha)
{
    return 1.0f / (LambdaSmith(wO, alpha) + LambdaSmith(wI, alpha) + 1.0f);
}
// BxDF::FSchlick (:501-512)
static inline V4 FSchlick(float VdH, V4 f0)
{
    float pw = 1.0f - VdH;
    float pw2 = pw * pw;
    float pw5 = pw2 * pw2 * pw;
    V4 r = (Splat4(1.0f) - f0) * pw5;
    return r + f0;
}
// BxDF::VNDFGGXSmithPDF (:514-526)
static inline float VNDFGGXSmithPDF(V3 V, V3 H, float alpha)
{
    float VdH = MaxF(0.0f, Dot(H, V));
    float NdH = MaxF(0.0f, H.z);
    float NdV = MaxF(0.0f, V.z);
    float D = DGGX(NdH, alpha);
    float GSingle = GSmithSingle(V, alpha);
    if(NdV == 0.0f) return 0.0f;
    return VdH * D * GSingle / NdV;
}
// BxDF::VNDFGGXSmithSample (:528-582), Heitz 2018
static inline DirSample VNDFGGXSmithSample(V3 V, float a, float xi0, float xi1)
{
    V3 VHemi = Normalize(V3{a * V.x, a * V.y, V.z});
    float len2 = Dot2(V2{VHemi.x, VHemi.y}, V2{VHemi.x, VHemi.y});
    V3 T1 = (len2 > 0) ? V3{-VHemi.y, VHemi.x, 0.0f} * RSqrt(len2) : V3{1, 0, 0};
    V3 T2 = Cross(VHemi, T1);
    float r = std::sqrt(xi0);
    float phi = 2.0f * kPi * xi1;
    float sinPhi, cosPhi;
    fm::SinCos(phi, sinPhi, cosPhi);
    float t1 = r * cosPhi;
    float t2 = r * sinPhi;
    float s = 0.5f * (1.0f + VHemi.z);
    t2 = (1.0f - s) * std::sqrt(1.0f - t1 * t1) + s * t2;
    float val = 1.0f - t1 * t1 - t2 * t2;
    V3 NHemi = (t1 * T1 + t2 * T2) + SqrtMax(val) * VHemi;
    V3 N = {a * NHemi.x, a * NHemi.y, SqrtMax(NHemi.z)};
    float nLen2 = LengthSqr(N);
    if(nLen2 < kEps) N = {0, 0, 1};
    else N = N * RSqrt(nLen2);
    return {N, VNDFGGXSmithPDF(V, N, a)};
}
// BxDF::BurleyDiffuseCorrection (:584-597)
static inline float BurleyDiffuseCorrection(float NdL, float NdV, float LdH, float roughness)
{
    float Fd90 = 0.5f + 2.0f * roughness * LdH * LdH;
    auto F = [Fd90](float d)
    {
        float pw = 1.0f - d;
        float pw2 = pw * pw;
        float pw5 = pw2 * pw2 * pw;
        return 1.0f + (Fd90 - 1.0f) * pw5;
    };
    return F(NdL) * F(NdV);
}
// Common::DivideByPDF (:599-612)
static inline V4 DivideByPDF(V4 t, float pdf)
{
    if(pdf == 0.0f) return Splat4(0.0f);
    float rcp = 1.0f / pdf;
    return t * rcp;
}
// MIS::BalanceCancelled (:955-965) with unit weights
static inline float Balance2(float p0, float w0, float p1, float w1)
{
    float r = 0.0f;
    r += p0 * w0;
    r += p1 * w1;
    return r;
}
// Medium::WavelengthToIoRCauchy (:977-986)
static inline float CauchyIoR(float wavelength, V3 c)
{
    float w = wavelength * 1e-3f;
    float w2 = w * w;
    float w4 = w2 * w2;
    return c.x + c.y / w2 + c.z / w4;
}

// ------------------------------------------------------------------------
// Triangle primitive (Tracer/PrimitiveDefaultTriangle.hpp,
// Core/ShapeFunctions.h)
// ------------------------------------------------------------------------
static inline V3 TriNormal(V3 p0, V3 p1, V3 p2) { return Normalize(Cross(p1 - p0, p2 - p0)); }
static inline float TriArea(V3 p0, V3 p1, V3 p2) { return Length(Cross(p1 - p0, p2 - p0)) * 0.5f; }

// Triangle::SurfaceFromHit validity (PrimitiveDefaultTriangle.hpp:318-330)
static inline bool BaryInRange(float a, float b)
{
    float c = 1.0f - a - b;
    return !(a < 0 || a > 1 || b < 0 || b > 1 || c < 0 || c > 1);
}

// Triangle::ProjectedHit -> Shape::Triangle::Project + PointToBarycentrics
// (ShapeFunctions.h:113-139); returns (a, b) as Hit
static inline void ProjectedHit(V3 p0, V3 p1, V3 p2, V3 point, float& ha, float& hb)
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
    Quat tbn;
    bool backSide;
};

// Triangle::GenerateSurface(DefaultSurface) (PrimitiveDefaultTriangle.hpp:443-630);
// ray-cone and texture-derivative terms are omitted: no texture is sampled.
static inline Surface GenSurface(const Scene& s, uint32_t tri, float a, float b, V3 rayDir)
{
    V3 p0, p1, p2;
    s.Tri(tri, p0, p1, p2);
    float c = 1.0f - a - b;
    Surface r;
    r.position = (p0 * a + p1 * b) + p2 * c;
    V3 geoN = TriNormal(p0, p1, p2);
    Quat q0 = s.tbn[s.idx[3 * tri + 0]];
    Quat q1 = s.tbn[s.idx[3 * tri + 1]];
    Quat q2 = s.tbn[s.idx[3 * tri + 2]];
    Quat tbn = QNormalize(QBarySLerp(q0, q1, q2, a, b));
    float dDotN = Dot(geoN, Normalize(rayDir));
    bool backSide = dDotN > 0.0f;
    if(backSide)
    {
        geoN = -geoN;
        tbn = QMul(Quat{0, 1, 0, 0}, tbn);
    }
    r.geoNormal = geoN;
    r.tbn = tbn;
    r.backSide = backSide;
    return r;
}

// ------------------------------------------------------------------------
// Spectrum handling.  RGB mode: SpectrumConverterIdentity (SpectrumC.h:241-250)
// converts colours to Spectrum(rgb, 0).  Spectral mode is set per path.
// ------------------------------------------------------------------------
struct SpecCtx
{
    bool spectral = false;
    V4 waves{0, 0, 0, 0};
};

// Spectral conversion hooks (filled by the spectral section below)
static V4 ConvertAlbedo(const SpecCtx& sc, V3 rgb);
static V4 ConvertRadiance(const SpecCtx& sc, V3 rgb);

// ------------------------------------------------------------------------
// Materials (Tracer/MaterialsDefault.hpp)
// ------------------------------------------------------------------------
struct BxDFSample
{
    V3 wIPos, wIDir;   // wI ray
    float pdf;
    V4 reflectance;
    bool passedThrough, dispersed;
};
struct BxDFEval { V4 reflectance; bool passedThrough, dispersed; };

struct MatInst
{
    uint32_t type;
    V4 albedo;
    float roughness, metallic, specular;
    float frontIoR, backIoR;
    Surface surf;
};

static inline MatInst MakeMaterial(const Scene& s, uint32_t mi, const Surface& surf, const SpecCtx& sc)
{
    const Material& m = s.mats[mi];
    MatInst r{};
    r.type = m.type;
    r.surf = surf;
    r.albedo = ConvertAlbedo(sc, m.albedo);
    r.roughness = m.roughness;
    r.metallic = m.metallic;
    r.specular = m.specular;
    if(m.type == MRAY_MAT_REFRACT)
    {
        // RefractMaterial ctor (MaterialsDefault.hpp:243-266)
        if(sc.spectral)
        {
            r.frontIoR = CauchyIoR(sc.waves.x, m.cauchyFront);
            r.backIoR = CauchyIoR(sc.waves.x, m.cauchyBack);
        }
        else
        {
            r.frontIoR = m.cauchyFront.x;
            r.backIoR = m.cauchyBack.x;
        }
    }
    return r;
}

// UnrealMaterial::MISRatio (:480-490)
static inline float UnrealMISRatio(const MatInst& m, float avgAlbedo)
{
    float integralDiffuse = 2.0f * kPi * avgAlbedo;
    integralDiffuse *= (1.0f - m.metallic);
    float specularRatio = m.specular * (1.0f - m.metallic) + avgAlbedo * m.metallic;
    float total = specularRatio + integralDiffuse;
    return (total == 0.0f) ? 0.0f : integralDiffuse / total;
}
static inline V4 UnrealF0(const MatInst& m)
{
    float specOut = m.specular * 0.08f;
    return Splat4(specOut) * (1.0f - m.metallic) + m.albedo * m.metallic;
}

// Material::Specularity
static inline float Specularity(const MatInst& m)
{
    switch(m.type)
    {
        case MRAY_MAT_LAMBERT: return 0.0f;
        case MRAY_MAT_REFLECT: return 1.0f;
        case MRAY_MAT_REFRACT: return 1.0f;
        default:
        {
            float avgAlbedo = Sum4(m.albedo) * 0.3333f;
            return 1.0f - UnrealMISRatio(m, avgAlbedo);
        }
    }
}
static inline bool IsSpecular(float s) { return s >= 0.95f; }   // MaterialC.h:11,128-132

// Material::SampleRNList sizes (MaterialsDefault.h:25,70,115,170)
static inline uint32_t MatRNCount(uint32_t type)
{
    switch(type)
    {
        case MRAY_MAT_LAMBERT: return 2;
        case MRAY_MAT_REFLECT: return 0;
        case MRAY_MAT_REFRACT: return 0;
        default: return 3;
    }
}

static BxDFSample SampleBxDF(const MatInst& m, V3 wO, const float* rn)
{
    BxDFSample out{};
    const Surface& sf = m.surf;
    switch(m.type)
    {
        case MRAY_MAT_LAMBERT:
        {
            // LambertMaterial::SampleBxDF (:28-60)
            DirSample ds = SampleCosDirection(rn[0], rn[1]);
            float nDotL = MaxF(ds.value.z, 0.0f);
            out.reflectance = m.albedo * nDotL * kInvPi;
            out.wIDir = QInvRotate(sf.tbn, ds.value);
            out.wIPos = sf.position;
            out.pdf = ds.pdf;
            return out;
        }
        case MRAY_MAT_REFLECT:
        {
            // ReflectMaterial::SampleBxDF (:155-176)
            V3 n = QBasisZ(sf.tbn);
            V3 wI = Reflect(n, wO);
            out.wIDir = Normalize(wI);
            out.wIPos = sf.position;
            out.pdf = 1.0f;
            out.reflectance = Splat4(1.0f);
            return out;
        }
        case MRAY_MAT_REFRACT:
        {
            // RefractMaterial::SampleBxDF (:268-303)
            float fromEta = m.frontIoR, toEta = m.backIoR;
            bool entering = !sf.backSide;
            if(!entering) std::swap(fromEta, toEta);
            V3 nLocal = QBasisZ(sf.tbn);
            float cosTheta = std::fabs(Dot(wO, nLocal));
            float f = FresnelDielectric(cosTheta, fromEta, toEta);
            float xi = rn[0];
            bool doReflection = (xi < f);
            V3 wI;
            if(doReflection) wI = Reflect(nLocal, wO);
            else Refract(wI, nLocal, wO, fromEta, toEta);
            float pdf = doReflection ? f : (1.0f - f);
            out.wIDir = wI;
            out.wIPos = Nudge(sf.position, nLocal);
            out.pdf = pdf;
            out.reflectance = Splat4(pdf);
            out.passedThrough = !doReflection;
            out.dispersed = !doReflection;
            return out;
        }
        default:
        {
            // UnrealMaterial::SampleBxDF (:534-611)
            float alpha = m.roughness * m.roughness;
            float avgAlbedo = Sum4(m.albedo) * 0.3333f;
            float sXi = rn[0];
            float xi0 = rn[1], xi1 = rn[2];
            float misRatio = UnrealMISRatio(m, avgAlbedo);
            float w0 = misRatio, w1 = 1.0f - misRatio;
            bool doDiffuse = (sXi < misRatio);
            V3 V = QRotate(sf.tbn, wO);
            V3 L, H;
            float pdfD, pdfS;
            if(doDiffuse)
            {
                DirSample s = SampleCosDirection(xi0, xi1);
                L = s.value;
                pdfD = s.pdf;
                H = Normalize(L + V);
                float VdH = MaxF(0.0f, Dot(V, H));
                pdfS = VNDFGGXSmithPDF(V, H, alpha);
                pdfS = (VdH == 0.0f) ? 0.0f : (pdfS / (4.0f * VdH));
            }
            else
            {
                DirSample s = VNDFGGXSmithSample(V, alpha, xi0, xi1);
                H = s.value;
                L = Reflect(H, V);
                float VdH = MaxF(0.0f, Dot(V, H));
                pdfS = (VdH == 0.0f) ? 0.0f : (s.pdf / (4.0f * VdH));
                pdfD = PDFCosDirection(L);
            }
            float VdH = MaxF(0.0f, Dot(V, H));
            float LdH = MaxF(0.0f, Dot(L, H));
            float NdH = MaxF(0.0f, H.z);
            float NdV = MaxF(0.0f, V.z);
            float D = DGGX(NdH, alpha);
            float G = GSmithCorrelated(V, L, alpha);
            G = (LdH == 0.0f) ? 0.0f : G;
            G = (VdH == 0.0f) ? 0.0f : G;
            V4 f0 = UnrealF0(m);
            V4 F = FSchlick(VdH, f0);
            V4 specTerm = D * F * G * 0.25f / NdV;
            specTerm = (NdV == 0.0f) ? Splat4(0.0f) : specTerm;
            if(std::isinf(D) || std::isnan(D))
            {
                specTerm = G * F / GSmithSingle(V, alpha);
                pdfS = 1.0f;
            }
            float NdL = MaxF(0.0f, L.z);
            V4 diffAlbedo = (1.0f - m.metallic) * m.albedo;
            V4 diffTerm = NdL * diffAlbedo * kInvPi;
            diffTerm = diffTerm * BurleyDiffuseCorrection(NdL, NdV, LdH, m.roughness);
            float pdfOut = Balance2(pdfD, w0, pdfS, w1);
            L = QInvRotate(sf.tbn, L);
            out.wIDir = L;
            out.wIPos = sf.position;
            out.pdf = pdfOut;
            out.reflectance = diffTerm + specTerm;
            return out;
        }
    }
}

static BxDFEval EvaluateBxDF(const MatInst& m, V3 wIDir, V3 wO)
{
    BxDFEval out{};
    const Surface& sf = m.surf;
    switch(m.type)
    {
        case MRAY_MAT_LAMBERT:
        {
            // LambertMaterial::Evaluate (:72-84)
            V3 wl = QRotate(sf.tbn, wIDir);
            float nDotL = MaxF(wl.z, 0.0f);
            out.reflectance = nDotL * m.albedo * kInvPi;
            return out;
        }
        case MRAY_MAT_REFLECT:
            out.reflectance = Splat4(1.0f);
            return out;
        case MRAY_MAT_REFRACT:
        {
            // RefractMaterial::Evaluate (:313-323)
            bool isFront = Dot(sf.geoNormal, wIDir) > 0.0f;
            out.reflectance = Splat4(1.0f);
            out.passedThrough = !isFront;
            out.dispersed = !isFront;
            return out;
        }
        default:
        {
            // UnrealMaterial::Evaluate (:648-694)
            float alpha = m.roughness * m.roughness;
            V3 V = QRotate(sf.tbn, wO);
            V3 L = QRotate(sf.tbn, wIDir);
            V3 H = Normalize(L + V);
            float LdH = MaxF(0.0f, Dot(L, H));
            float VdH = MaxF(0.0f, Dot(V, H));
            float NdH = MaxF(0.0f, H.z);
            float NdV = MaxF(0.0f, V.z);
            float D = DGGX(NdH, alpha);
            D = (std::isnan(D) || std::isinf(D)) ? 0.0f : D;
            float G = GSmithCorrelated(V, L, alpha);
            G = (LdH == 0.0f) ? 0.0f : G;
            G = (VdH == 0.0f) ? 0.0f : G;
            V4 f0 = UnrealF0(m);
            V4 F = FSchlick(VdH, f0);
            V4 specTerm = D * F * G * 0.25f / NdV;
            specTerm = (NdV == 0.0f) ? Splat4(0.0f) : specTerm;
            float NdL = MaxF(0.0f, L.z);
            V4 diffAlbedo = (1.0f - m.metallic) * m.albedo;
            V4 diffTerm = NdL * diffAlbedo * kInvPi;
            diffTerm = diffTerm * BurleyDiffuseCorrection(NdL, NdV, LdH, m.roughness);
            out.reflectance = diffTerm + specTerm;
            return out;
        }
    }
}

static float PdfBxDF(const MatInst& m, V3 wIDir, V3 wO)
{
    const Surface& sf = m.surf;
    switch(m.type)
    {
        case MRAY_MAT_LAMBERT:
        {
            // LambertMaterial::Pdf (:62-70)
            V3 wl = QRotate(sf.tbn, wIDir);
            return MaxF(PDFCosDirection(wl), 0.0f);
        }
        case MRAY_MAT_REFLECT:
        case MRAY_MAT_REFRACT:
            return 0.0f;
        default:
        {
            // UnrealMaterial::Pdf (:613-646)
            float alpha = m.roughness * m.roughness;
            float avgAlbedo = Sum4(m.albedo) * 0.3333f;
            V3 V = QRotate(sf.tbn, wO);
            V3 L = QRotate(sf.tbn, wIDir);
            V3 H = Normalize(L + V);
            float misRatio = UnrealMISRatio(m, avgAlbedo);
            float w0 = misRatio, w1 = 1.0f - misRatio;
            float p0 = PDFCosDirection(L);
            float NdH = MaxF(0.0f, H.z);
            float VdH = MaxF(0.0f, Dot(V, H));
            float D = DGGX(NdH, alpha);
            float p1 = VNDFGGXSmithPDF(V, H, alpha);
            p1 = (std::isnan(D) || std::isinf(D)) ? 0.0f : p1;
            p1 = (VdH == 0.0f) ? 0.0f : p1 / (4.0f * VdH);
            return Balance2(p0, w0, p1, w1);
        }
    }
}

// ------------------------------------------------------------------------
// Lights (Tracer/LightsDefault.hpp, LightSampler.hpp, MetaLight.hpp:470-520).
// Meta-light list: one entry per emissive triangle in order, then the
// boundary light last.
// ------------------------------------------------------------------------
static inline V3 YUpFromZUp(V3 v) { return {v.y, v.z, v.x}; }  // ZUpToYUp (Matrix.hpp:1227)
static inline V3 ZUpFromYUp(V3 v) { return {v.z, v.x, v.y}; }  // YUpToZUp (Matrix.hpp:1221)

// Graphics::ConcentricOctahedralToDirection (GraphicsFunctions.h:354-381)
static inline V3 CoOctaToDir(float s, float t)
{
    const float piOvr4 = kPi * 0.25f;
    float u = s * 2.0f - 1.0f, v = t * 2.0f - 1.0f;
    float ua = std::fabs(u), va = std::fabs(v);
    float d = 1.0f - (ua + va);
    float radius = 1.0f - std::fabs(d);
    float phiPrime = 0.0f;
    if(radius != 0.0f) phiPrime = ((va - ua) / radius + 1) * piOvr4;
    float sinP, cosP;
    fm::SinCos(phiPrime, sinP, cosP);
    float cosPhi = SignPM1(u) * cosP;
    float sinPhi = SignPM1(v) * sinP;
    float z = SignPM1(d) * (1.0f - radius * radius);
    float xyFactor = radius * std::sqrt(2.0f - radius * radius);
    return {cosPhi * xyFactor, sinPhi * xyFactor, z};
}

// Graphics::DirectionToConcentricOctahedral (GraphicsFunctions.h:318-353):
// Z-up unit direction -> (s, t) in [0, 1]^2; the +Z pole goes to (0, 0) as
// in the reference.
static inline void DirToCoOcta(V3 d, float& s, float& t)
{
    const float twoOvrPi = kInvPi * 2.0f;
    if(d.x == 0.0f && d.y == 0.0f) { s = 0.0f; t = 0.0f; return; }
    float xAbs = std::fabs(d.x), yAbs = std::fabs(d.y);
    float phiPrime = std::atan(yAbs / xAbs);
    float radius = std::sqrt(MaxF(0.0f, 1.0f - std::fabs(d.z)));
    float v = radius * twoOvrPi * phiPrime;
    float u = radius - v;
    if(d.z < 0.0f)
    {
        float uPrime = 1.0f - v;
        float vPrime = 1.0f - u;
        u = uPrime;
        v = vPrime;
    }
    u *= SignPM1(d.x);
    v *= SignPM1(d.y);
    s = (u + 1.0f) * 0.5f;
    t = (v + 1.0f) * 0.5f;
}

static inline bool IsSkysphere(uint32_t t)
{
    return t == MRAY_BOUNDARY_SKYSPHERE_COOCTA || t == MRAY_BOUNDARY_SKYSPHERE_SPHERICAL;
}
// SphericalCoordConverter (LightsDefault.hpp:177-222) over
// CartesianToUnitSpherical / UVToSphericalAngles (GraphicsFunctions.h:253-285,
// 447-470): Z-up azimuth and inclination; u = (azimuth + pi) / (2 pi),
// v = 1 - inclination / pi.
static inline void DirToSpherical(V3 dz, float& u, float& v)
{
    float azimuth = fm::Atan2(dz.y, dz.x);
    float incl = fm::Acos(std::min(std::max(dz.z, -1.0f), 1.0f));
    u = (azimuth + kPi) * 0.5f / kPi;
    v = 1.0f - incl * kInvPi;
}
static inline V3 SphericalToDir(float u, float v, float& sinIncl)
{
    float theta = (u * kPi * 2.0f) - kPi;
    float phi = (1.0f - v) * kPi;
    float sT, cT, sP, cP;
    fm::SinCos(theta, sT, cT);
    fm::SinCos(phi, sP, cP);
    sinIncl = sP;
    return {cT * sP, sT * sP, cP};
}
static inline void SkyUV(uint32_t type, V3 dirYUp, float& u, float& v)
{
    if(type == MRAY_BOUNDARY_SKYSPHERE_SPHERICAL) DirToSpherical(ZUpFromYUp(dirYUp), u, v);
    else DirToCoOcta(ZUpFromYUp(dirYUp), u, v);
}
// ToSolidAnglePdf: CoOcta uvPdf / (4 pi); spherical uvPdf / (2 pi^2 sin),
// zero where sin <= 0 (LightsDefault.hpp:195-222, 270-281)
static inline float SkySolidAnglePdf(uint32_t type, float uvPdf, float sinIncl)
{
    if(type == MRAY_BOUNDARY_SKYSPHERE_SPHERICAL)
        return (sinIncl <= 0.0f) ? 0.0f : uvPdf / (2.0f * kPi * kPi * sinIncl);
    return uvPdf * 0.25f * kInvPi;
}

// Skysphere texture read at a Y-up direction: the reference's default
// texture parameters (TracerI.h:71-82: linear interpolation, wrapped
// edges), texel centres at (i + 1/2) / size, bilinear weights in fp32.
static V3 EnvRadiance(const Scene& s, V3 dirYUp)
{
    float u, v;
    SkyUV(s.boundaryType, dirYUp, u, v);
    const int W = int(s.envW), H = int(s.envH);
    float xf = u * float(W) - 0.5f, yf = v * float(H) - 0.5f;
    float x0f = std::floor(xf), y0f = std::floor(yf);
    float fx = xf - x0f, fy = yf - y0f;
    auto wrap = [](int i, int n) { return ((i % n) + n) % n; };
    int x0 = wrap(int(x0f), W), x1 = wrap(int(x0f) + 1, W);
    int y0 = wrap(int(y0f), H), y1 = wrap(int(y0f) + 1, H);
    const V3 a = s.envTex[size_t(y0) * W + x0], b = s.envTex[size_t(y0) * W + x1];
    const V3 c = s.envTex[size_t(y1) * W + x0], d = s.envTex[size_t(y1) * W + x1];
    const float gx = 1.0f - fx, gy = 1.0f - fy;
    V3 top = {a.x * gx + b.x * fx, a.y * gx + b.y * fx, a.z * gx + b.z * fx};
    V3 bot = {c.x * gx + d.x * fx, c.y * gx + d.y * fx, c.z * gx + d.z * fx};
    return {top.x * gy + bot.x * fy, top.y * gy + bot.y * fy, top.z * gy + bot.z * fy};
}

// DistributionPwC<1>::SampleIndex (Distributions.h:128-148): lower bound of
// xi in the CDF, linear position inside the bin, and the clamp as written
// there (an index of 1 or more steps down one ulp).  Two guards of this
// restatement: the bin index stays below n, and an empty bin gives t = 0.
static float PwcSampleIndex(const float* cdf, uint32_t n, float xi, float& pdf)
{
    uint32_t index = uint32_t(std::lower_bound(cdf, cdf + n, xi) - cdf);
    index = std::min(index, n - 1u);
    float prevCDF = (index == 0) ? 0.0f : cdf[index - 1];
    float myCDF = cdf[index];
    float t = (myCDF > prevCDF) ? (xi - prevCDF) / (myCDF - prevCDF) : 0.0f;
    float indexF = float(index) + t;
    indexF = (indexF < 1.0f) ? indexF : std::nextafter(indexF, -INFINITY);
    pdf = (myCDF - prevCDF) * float(n);
    return indexF;
}
// DistributionPwC<1>::PdfIndex (Distributions.h:158-170)
static float PwcPdfIndex(const float* cdf, uint32_t n, float index)
{
    uint32_t i = uint32_t(index);
    float prevCDF = (i == 0) ? 0.0f : cdf[i - 1];
    return (cdf[i] - prevCDF) * float(n);
}
// DistributionPwC<2>::SampleUV (Distributions.h:181-213)
static void EnvSampleUV(const Scene& s, float xi0, float xi1, float& u, float& v, float& pdf)
{
    float pdfY, pdfX;
    float yF = PwcSampleIndex(s.envCdfY.data(), s.envH, xi1, pdfY);
    uint32_t row = uint32_t(yF);
    float xF = PwcSampleIndex(&s.envCdfX[size_t(row) * s.envW], s.envW, xi0, pdfX);
    u = xF * (1.0f / float(s.envW));
    v = yF * (1.0f / float(s.envH));
    pdf = pdfY * pdfX;
}
// DistributionPwC<2>::PdfUV (Distributions.h:232-243)
static float EnvPdfUV(const Scene& s, float u, float v)
{
    float xF = std::min(u * float(s.envW), float(s.envW) - 1.0f);
    float yF = std::min(v * float(s.envH), float(s.envH) - 1.0f);
    uint32_t row = uint32_t(yF);
    float pdfMarginal = PwcPdfIndex(s.envCdfY.data(), s.envH, yF);
    float pdfConditional = PwcPdfIndex(&s.envCdfX[size_t(row) * s.envW], s.envW, xF);
    return pdfConditional * pdfMarginal;
}
// LightSkysphere::PdfSolidAngle (LightsDefault.hpp:359-370)
static float SkyspherePdf(const Scene& s, V3 dir)
{
    float uvPdf = 1.0f, sinIncl = 1.0f;
    const V3 d = Normalize(dir);
    if(s.boundaryType == MRAY_BOUNDARY_SKYSPHERE_SPHERICAL)
        sinIncl = fm::Sin(fm::Acos(std::min(std::max(d.y, -1.0f), 1.0f)));
    if(s.envW)
    {
        float u, v;
        SkyUV(s.boundaryType, d, u, v);
        uvPdf = EnvPdfUV(s, u, v);
    }
    return SkySolidAnglePdf(s.boundaryType, uvPdf, sinIncl);
}
static V3 SkysphereRadiance(const Scene& s, V3 dirYUp)
{
    return s.envW ? EnvRadiance(s, dirYUp) : s.boundaryRad;
}

// The luminance distribution of a Skysphere texture (LightGroupSkysphere::
// Finalize, LightsDefault.hpp:860-907; DistributionGroupPwC2D::Construct,
// Distributions.cu:465-538): Rec.709 Y per texel, rows scanned in double
// (rounded to float per element), the marginal as a float running sum of
// row totals, each normalized by the float reciprocal of its last entry.
// A zero row keeps a uniform CDF (the reference divides by zero there).
static int BuildEnv(Scene& s, uint32_t W, uint32_t H, const float* rgb)
{
    if(!rgb || W == 0 || H == 0 || W > 8192 || H > 8192) return -1;
    // Y row of the Rec.709 (D65) RGB -> XYZ matrix from the primaries
    const double Pm[3][3] = {{0.64 / 0.33, 0.30 / 0.60, 0.15 / 0.06},
                             {1.0, 1.0, 1.0},
                             {(1 - 0.64 - 0.33) / 0.33, (1 - 0.30 - 0.60) / 0.60,
                              (1 - 0.15 - 0.06) / 0.06}};
    const double Wt[3] = {0.3127 / 0.3290, 1.0, (1 - 0.3127 - 0.3290) / 0.3290};
    auto det3 = [](const double m[3][3])
    {
        return m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
               m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
               m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    };
    float yRow[3];
    for(int c = 0; c < 3; c++)
    {
        double m[3][3];
        std::memcpy(m, Pm, sizeof(m));
        for(int r = 0; r < 3; r++) m[r][c] = Wt[r];
        yRow[c] = float(det3(m) / det3(Pm));
    }
    const size_t P = size_t(W) * H;
    s.envTex.resize(P);
    s.envCdfX.assign(P, 0.0f);
    s.envCdfY.assign(H, 0.0f);
    for(size_t i = 0; i < P; i++) s.envTex[i] = {rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]};
    float ySum = 0.0f;
    for(uint32_t y = 0; y < H; y++)
    {
        double sum = 0.0;
        for(uint32_t x = 0; x < W; x++)
        {
            const V3 t = s.envTex[size_t(y) * W + x];
            const float lum = yRow[0] * t.x + yRow[1] * t.y + yRow[2] * t.z;
            sum += double(std::fabs(lum));
            s.envCdfX[size_t(y) * W + x] = float(sum);
        }
        ySum += s.envCdfX[size_t(y) * W + W - 1];
        s.envCdfY[y] = ySum;
    }
    if(!(ySum > 0.0f)) return -1;
    for(uint32_t y = 0; y < H; y++)
    {
        float* row = &s.envCdfX[size_t(y) * W];
        if(row[W - 1] > 0.0f)
        {
            const float recip = 1.0f / row[W - 1];
            for(uint32_t x = 0; x < W; x++) row[x] *= recip;
        }
        else
            for(uint32_t x = 0; x < W; x++) row[x] = float(x + 1) / float(W);
    }
    const float recip = 1.0f / s.envCdfY[H - 1];
    for(uint32_t y = 0; y < H; y++) s.envCdfY[y] *= recip;
    s.envW = W;
    s.envH = H;
    return 0;
}

struct LightSampleOut
{
    V3 position;
    V4 emission;
    float pdf;
};

static inline V4 TriLightEmitViaSurfacePoint(const Scene& s, uint32_t li, V3 wO, V3 point,
                                             const SpecCtx& sc)
{
    V3 p0, p1, p2;
    s.Tri(s.lightTri[li], p0, p1, p2);
    float ha, hb;
    ProjectedHit(p0, p1, p2, point, ha, hb);
    if(!BaryInRange(ha, hb)) return Splat4(0.0f);
    V3 n = TriNormal(p0, p1, p2);
    float NdL = Dot(n, wO);
    if(!s.lightTwoSided[li] && NdL <= 0.0f) return Splat4(0.0f);
    return ConvertRadiance(sc, s.lightRad[li]);
}

// Power sampling (MRAY_RENDER_LIGHT_POWER, the build's own sampler):
// weights Y(radiance) * area (x2 two-sided), an alias table built with
// Vose's method in double over the emissive triangles; the boundary light
// keeps probability 1 / (n + 1); a triangle light gets n / (n + 1) * w / W.
static void BuildLightPowerTable(Scene& s)
{
    const uint32_t n = uint32_t(s.lightTri.size());
    s.lightProb.assign(n, 1.0f);
    s.lightAlias.assign(n, 0u);
    s.lightPmf.assign(n, 0.0f);
    if(n == 0) return;
    std::vector<double> w(n);
    double total = 0.0;
    for(uint32_t i = 0; i < n; i++)
    {
        V3 p0, p1, p2;
        s.Tri(s.lightTri[i], p0, p1, p2);
        const double ax = double(p1.x) - p0.x, ay = double(p1.y) - p0.y, az = double(p1.z) - p0.z;
        const double bx = double(p2.x) - p0.x, by = double(p2.y) - p0.y, bz = double(p2.z) - p0.z;
        const double cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
        const double area = 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
        const V3 L = s.lightRad[i];
        const double lum = std::fabs(0.2126390058715103 * L.x + 0.7151686787677559 * L.y +
                                     0.0721923153607337 * L.z);
        double wi = lum * area * (s.lightTwoSided[i] ? 2.0 : 1.0);
        if(!std::isfinite(wi)) wi = 0.0;
        w[i] = wi;
        total += wi;
    }
    if(!(total > 0.0) || !std::isfinite(total))
    {
        std::fill(w.begin(), w.end(), 1.0);
        total = double(n);
    }
    std::vector<double> q(n);
    std::vector<uint32_t> under, over;
    for(uint32_t i = 0; i < n; i++)
    {
        q[i] = w[i] * double(n) / total;
        if(q[i] < 1.0) under.push_back(i); else over.push_back(i);
    }
    while(!under.empty() && !over.empty())
    {
        uint32_t a = under.back();
        under.pop_back();
        uint32_t b = over.back();
        s.lightProb[a] = float(q[a]);
        s.lightAlias[a] = b;
        q[b] = (q[b] + q[a]) - 1.0;
        if(q[b] < 1.0)
        {
            over.pop_back();
            under.push_back(b);
        }
    }
    for(uint32_t i : over) { s.lightProb[i] = 1.0f; s.lightAlias[i] = i; }
    for(uint32_t i : under) { s.lightProb[i] = 1.0f; s.lightAlias[i] = i; }
    const double share = double(n) / double(n + 1);
    for(uint32_t i = 0; i < n; i++) s.lightPmf[i] = float(share * w[i] / total);
}

// DirectLightSamplerUniform::SampleLight (LightSampler.hpp:30-74); rn[0..1]
// sample the point, rn[2] picks the light (power: the stratum of a triangle
// light is its alias bucket).
static LightSampleOut SampleLight(const Scene& s, const float* rn, V3 distantPoint, const SpecCtx& sc,
                                  bool power = false)
{
    uint32_t lightCount = uint32_t(s.lightTri.size()) + 1u;
    float lightCountF = float(lightCount);
    float dXi = rn[2];
    float lightIndexF = dXi * lightCountF;
    uint32_t li = uint32_t(lightIndexF);
    li = std::min(std::max(li, 0u), lightCount - 1u);
    float selectionPdf = 1.0f / lightCountF;
    if(power && li < lightCount - 1u)
    {
        const float f = rn[2] * lightCountF - float(li);
        li = (f < s.lightProb[li]) ? li : s.lightAlias[li];
        selectionPdf = s.lightPmf[li];
    }
    V3 point;
    float pointPdf;
    V4 emission;
    if(li < lightCount - 1u)
    {
        // LightPrim::SampleSolidAngle (LightsDefault.hpp:22-44) with
        // Triangle::SampleSurface (PrimitiveDefaultTriangle.hpp:48-77)
        V3 p0, p1, p2;
        s.Tri(s.lightTri[li], p0, p1, p2);
        float r1 = std::sqrt(rn[0]);
        float r2 = rn[1];
        float a = 1 - r1;
        float b = (1 - r2) * r1;
        float c = r1 * r2;
        float area = TriArea(p0, p1, p2);
        float sPdf = 1.0f / area;
        V3 position = (p0 * a + p1 * b) + p2 * c;
        V3 normal = TriNormal(p0, p1, p2);
        V3 sampledDir = distantPoint - position;
        float distSqr = LengthSqr(sampledDir);
        sampledDir = Normalize(sampledDir);
        float NdL = Dot(normal, sampledDir);
        NdL = s.lightTwoSided[li] ? std::fabs(NdL) : MaxF(0.0f, NdL);
        float pdf = (NdL == 0.0f) ? 0.0f : sPdf / NdL;
        pdf *= distSqr;
        point = position;
        pointPdf = pdf;
    }
    else if(IsSkysphere(s.boundaryType))
    {
        // LightSkysphere::SampleSolidAngle (LightsDefault.hpp:340-357):
        // constant radiance: uv = xi, pdf 1; a texture: PWC-2D by
        // luminance; then the mapping's solid-angle pdf
        float u = rn[0], v = rn[1], uvPdf = 1.0f, sinIncl = 1.0f;
        if(s.envW) EnvSampleUV(s, rn[0], rn[1], u, v, uvPdf);
        V3 dirYUp = YUpFromZUp((s.boundaryType == MRAY_BOUNDARY_SKYSPHERE_SPHERICAL)
                                   ? SphericalToDir(u, v, sinIncl) : CoOctaToDir(u, v));
        float pdf = SkySolidAnglePdf(s.boundaryType, uvPdf, sinIncl);
        point = distantPoint + dirYUp * s.sceneDiameter;
        pointPdf = pdf;
    }
    else
    {
        // LightNull::SampleSolidAngle (LightsDefault.hpp:461-470)
        point = {1e10f, 1e10f, 1e10f};
        pointPdf = 0.0f;
    }
    V3 wO = distantPoint - point;
    wO = Normalize(wO);
    if(li < lightCount - 1u)
        emission = TriLightEmitViaSurfacePoint(s, li, wO, point, sc);
    else if(IsSkysphere(s.boundaryType))
        emission = ConvertRadiance(sc, SkysphereRadiance(s, Normalize(-wO)));
    else
        emission = Splat4(0.0f);
    return {point, emission, selectionPdf * pointPdf};
}

// DirectLightSamplerUniform::PdfLight (LightSampler.hpp:76-119) for a hit
// triangle light or the boundary light; rayPos/rayDir of the path ray.
static float PdfLightTri(const Scene& s, uint32_t li, float ha, float hb, V3 rayPos, V3 rayDir,
                         bool power = false)
{
    uint32_t lightCount = uint32_t(s.lightTri.size()) + 1u;
    float selectionPDF = power ? s.lightPmf[li] : 1.0f / float(lightCount);
    V3 p0, p1, p2;
    s.Tri(s.lightTri[li], p0, p1, p2);
    // LightPrim::PdfSolidAngle (LightsDefault.hpp:46-66)
    float pdf;
    if(!BaryInRange(ha, hb)) pdf = 0.0f;
    else
    {
        float hc = 1.0f - ha - hb;
        V3 position = (p0 * ha + p1 * hb) + p2 * hc;
        V3 normal = TriNormal(p0, p1, p2);
        pdf = 1.0f / TriArea(p0, p1, p2);
        float NdL = Dot(normal, -rayDir);
        NdL = s.lightTwoSided[li] ? std::fabs(NdL) : MaxF(0.0f, NdL);
        pdf = (NdL == 0.0f) ? 0.0f : pdf / NdL;
        pdf *= LengthSqr(rayPos - position);
    }
    return pdf * selectionPDF;
}
static float PdfLightBoundary(const Scene& s, V3 rayDir)
{
    uint32_t lightCount = uint32_t(s.lightTri.size()) + 1u;
    float selectionPDF = 1.0f / float(lightCount);
    float pdf = 0.0f;
    if(IsSkysphere(s.boundaryType)) pdf = SkyspherePdf(s, rayDir);
    return pdf * selectionPDF;
}

// ------------------------------------------------------------------------
// Camera (CamerasDefault.hpp:8-36, 97-151) and Gaussian film filter
// (Filters.h:195-220)
// ------------------------------------------------------------------------
struct Camera
{
    V3 position, right, up, gazeDir, bottomLeft;
    V2 planeSize;
    V2 nearFar;
};

static Camera MakeCamera(const MRayCamera& c)
{
    Camera r;
    V3 pos = {c.position[0], c.position[1], c.position[2]};
    V3 up = {c.up[0], c.up[1], c.up[2]};
    V3 gaze = {c.gaze[0], c.gaze[1], c.gaze[2]};
    r.nearFar = {c.nearFar[0], c.nearFar[1]};
    float widthHalf = std::tan(c.fov[0] * 0.5f) * r.nearFar.x;
    float heightHalf = std::tan(c.fov[1] * 0.5f) * r.nearFar.x;
    V3 gazeDir = gaze - pos;
    V3 right = Normalize(Cross(gazeDir, up));
    up = Normalize(Cross(right, gazeDir));
    gazeDir = Normalize(Cross(up, right));
    r.bottomLeft = ((pos - right * widthHalf) - up * heightHalf) + gazeDir * r.nearFar.x;
    r.planeSize = {widthHalf * 2.0f, heightHalf * 2.0f};
    r.position = pos;
    r.right = right;
    r.up = up;
    r.gazeDir = gazeDir;
    return r;
}

struct CamRay { V3 pos, dir; float tMin, tMax, weight; };

// KCGenerateCamRaysStochastic body (RayGenKernels.kt.h:182-225)
static CamRay GenCamRay(const Camera& cam, uint32_t px, uint32_t py, uint32_t W, uint32_t H,
                        float xi0, float xi1, float filterRadius)
{
    float sigma = filterRadius * 0.285714f;
    Sample1 s0 = SampleGaussian(xi0, sigma);
    Sample1 s1 = SampleGaussian(xi1, sigma);
    float offPdf = s0.pdf * s1.pdf;
    float weight = Gaussian(s0.value, sigma) * Gaussian(s1.value, sigma);
    // CameraPinhole::EvaluateRay (CamerasDefault.hpp:97-151)
    V2 delta = {cam.planeSize.x / float(W), cam.planeSize.y / float(H)};
    V2 jitter = {s0.value + 0.5f, s1.value + 0.5f};
    V2 sd = {(float(px) + jitter.x) * delta.x, (float(py) + jitter.y) * delta.y};
    V3 sp = cam.bottomLeft;
    sp = sp + sd.x * cam.right;
    sp = sp + sd.y * cam.up;
    V3 dir = Normalize(sp - cam.position);
    return {cam.position, dir, cam.nearFar.x, cam.nearFar.y, weight / offPdf};
}

// ------------------------------------------------------------------------
// Spectral: hero-wavelength sampling (SpectrumContext.cu:14-135),
// RGB->spectrum (SpectrumContext.hpp:37-150, Jakob & Hanika 2019) and
// spectrum->RGB (SpectrumContext.cu:137-171).
// ------------------------------------------------------------------------
struct SpectralData
{
    bool loaded = false;
    int N = 0;                        // LUT resolution per axis
    std::vector<float> lut;           // [3][N][N][N][4] (c0, c1, c2, 0), x fastest
    std::vector<float> cmf;           // [471][3] observer, 360..830 nm
    std::vector<float> illum;         // [471] illuminant SPD (Y-normalised)
    float xyzToRgb[9] = {0};
};
static SpectralData gSpec;

// Math::InvSmoothstep (Core/Math.h:1224-1232)
static inline float InvSmoothstep(float y)
{
    const float factor = 1.0f / 3.0f;
    return 0.5f - fm::Sin(fm::Asin(1.0f - 2.0f * y) * factor);
}

// 3-D texture fetch with normalised coordinates and texel centres at
// (i + 0.5) / N, i.e. texel space u * (N - 1) for uv = xyz * (N-1)/N + 0.5/N
// (SpectrumContext.hpp:113-118); trilinear, clamp to edge.
static void LutFetch(int ch, float x, float y, float z, float out[3])
{
    const int N = gSpec.N;
    float fx = Clamp(x, 0.0f, 1.0f) * float(N - 1);
    float fy = Clamp(y, 0.0f, 1.0f) * float(N - 1);
    float fz = Clamp(z, 0.0f, 1.0f) * float(N - 1);
    int x0 = std::min(int(fx), N - 2), y0 = std::min(int(fy), N - 2), z0 = std::min(int(fz), N - 2);
    float tx = fx - float(x0), ty = fy - float(y0), tz = fz - float(z0);
    auto C = [&](int zz, int yy, int xx, int k)
    {
        size_t o = ((((size_t(ch) * N + zz) * N + yy) * N + xx) * 4) + size_t(k);
        return gSpec.lut[o];
    };
    for(int k = 0; k < 3; k++)
    {
        float c00 = C(z0, y0, x0, k) * (1 - tx) + C(z0, y0, x0 + 1, k) * tx;
        float c01 = C(z0, y0 + 1, x0, k) * (1 - tx) + C(z0, y0 + 1, x0 + 1, k) * tx;
        float c10 = C(z0 + 1, y0, x0, k) * (1 - tx) + C(z0 + 1, y0, x0 + 1, k) * tx;
        float c11 = C(z0 + 1, y0 + 1, x0, k) * (1 - tx) + C(z0 + 1, y0 + 1, x0 + 1, k) * tx;
        float c0 = c00 * (1 - ty) + c01 * ty;
        float c1 = c10 * (1 - ty) + c11 * ty;
        out[k] = c0 * (1 - tz) + c1 * tz;
    }
}

// Converter::ConvertAlbedo polynomial (SpectrumContext.hpp:40-49)
static inline float EvalSigmoidPoly(const float c[3], float lambda)
{
    float t = std::fma(c[0], lambda, c[1]);
    float x = std::fma(t, lambda, c[2]);
    float denomRecip = 1.0f / std::sqrt(std::fma(x, x, 1.0f));
    return std::fma(0.5f * x, denomRecip, 0.5f);
}

// Converter::ConvertAlbedo (SpectrumContext.hpp:37-125)
static V4 RGBToSpectrumAlbedo(V3 rgb, V4 waves)
{
    float c[3] = {rgb.x, rgb.y, rgb.z};
    int maxI = 0;
    if(c[1] > c[maxI]) maxI = 1;
    if(c[2] > c[maxI]) maxI = 2;
    float maxC = c[maxI];
    float x = 0.0f, y = 0.0f;
    if(maxC > kSmallEps)
    {
        float f = 1.0f / maxC;
        x = c[(maxI + 1) % 3] * f;
        y = c[(maxI + 2) % 3] * f;
    }
    float slice = Clamp(maxC, 0.0f, 1.0f);
    float z = InvSmoothstep(InvSmoothstep(slice));
    float co[3];
    LutFetch(maxI, x, y, z, co);
    return {EvalSigmoidPoly(co, waves.x), EvalSigmoidPoly(co, waves.y),
            EvalSigmoidPoly(co, waves.z), EvalSigmoidPoly(co, waves.w)};
}

// 1-D linear texture over integer nanometres (coordinate lambda + 0.5 - 360)
static inline float Tex1D(const std::vector<float>& t, int stride, int comp, float lambda)
{
    float f = Clamp(lambda - 360.0f, 0.0f, 470.0f);
    int i = std::min(int(f), 469);
    float u = f - float(i);
    return t[size_t(i) * stride + comp] * (1 - u) + t[size_t(i + 1) * stride + comp] * u;
}

static V4 ConvertAlbedo(const SpecCtx& sc, V3 rgb)
{
    if(!sc.spectral) return {rgb.x, rgb.y, rgb.z, 0.0f};
    return RGBToSpectrumAlbedo(rgb, sc.waves);
}
// Converter::ConvertRadiance (SpectrumContext.hpp:127-150)
static V4 ConvertRadiance(const SpecCtx& sc, V3 rgb)
{
    if(!sc.spectral) return {rgb.x, rgb.y, rgb.z, 0.0f};
    float mx = MaxF(MaxF(rgb.x, rgb.y), rgb.z);
    float scale = mx * 2.0f;
    V3 n = (scale == 0.0f) ? V3{0, 0, 0} : V3{rgb.x / scale, rgb.y / scale, rgb.z / scale};
    V4 s = RGBToSpectrumAlbedo(n, sc.waves);
    s.x *= Tex1D(gSpec.illum, 1, 0, sc.waves.x);
    s.y *= Tex1D(gSpec.illum, 1, 0, sc.waves.y);
    s.z *= Tex1D(gSpec.illum, 1, 0, sc.waves.z);
    s.w *= Tex1D(gSpec.illum, 1, 0, sc.waves.w);
    return s * scale;
}

// SingleSampleSpectrumWavelength, HYPERBOLIC_PBRT mode (SpectrumContext.cu:14-135)
static void SampleWavelengths(float xi0, V4& waves, V4& pdf)
{
    float off[4] = {-0.5f, -0.25f, 0.0f, 0.25f};
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

static inline V3 CMFAt(float lambda)
{
    return {Tex1D(gSpec.cmf, 3, 0, lambda), Tex1D(gSpec.cmf, 3, 1, lambda),
            Tex1D(gSpec.cmf, 3, 2, lambda)};
}

// ConvertSpectraToRGBSingle (SpectrumContext.cu:137-171)
static V4 SpectrumToRGB(V4 value, V4 waves, V4 pdf, bool dispersed)
{
    float val[4] = {value.x, value.y, value.z, value.w};
    float wv[4] = {waves.x, waves.y, waves.z, waves.w};
    float pd[4] = {pdf.x, pdf.y, pdf.z, pdf.w};
    int count = dispersed ? 1 : 4;
    float weight = dispersed ? 1.0f : 0.25f;
    V3 xyz = {0, 0, 0};
    for(int i = 0; i < count; i++)
    {
        V3 f = CMFAt(wv[i]);
        float v = (pd[i] == 0.0f) ? 0.0f : val[i] / pd[i];
        xyz = xyz + f * v;
    }
    xyz = xyz * weight;
    const float* M = gSpec.xyzToRgb;
    float r = M[0] * xyz.x + M[1] * xyz.y + M[2] * xyz.z;
    float g = M[3] * xyz.x + M[4] * xyz.y + M[5] * xyz.z;
    float b = M[6] * xyz.x + M[7] * xyz.y + M[8] * xyz.z;
    return {r, g, b, 0.0f};
}

// ------------------------------------------------------------------------
// Path tracer, reference schedule (PathTracerRenderer.cu:303-481,
// 930-1074; PathTracerRendererBase.cu:69-262; PathTracerRendererShaders.h)
// ------------------------------------------------------------------------
enum : uint8_t { ST_ALIVE = 0, ST_DEAD = 1, ST_INVALID = 2 };
enum : uint8_t { RT_CAMERA = 0, RT_PATH = 1, RT_SPECULAR = 2, RT_SHADOW = 3 };

struct Slot
{
    V3 pos, dir;
    float tMin, tMax;
    V4 throughput, radiance;
    uint32_t depth;
    uint8_t status, type;
    bool dispersed;
    float prevMatPdf;
    uint32_t pixel;
    float filterWeight;
    uint32_t rng;
    V4 waves, wavePdf;
};

struct RenderCounters
{
    uint64_t extRays = 0, shadowRays = 0;
};

static inline uint32_t LightIndexOfTri(const Scene& s, uint32_t tri)
{
    uint32_t k = s.surfKey[tri];
    return (k & MRAY_LIGHT_FLAG) ? (k & ~MRAY_LIGHT_FLAG) : kInvalid;
}

// One bounce of one alive slot: cast, then boundary/light work or
// NEE (work 1) + visibility + accumulate + scatter (work 0).
static void BounceSlot(const Scene& s, const MRayRenderParams& p, Slot& sl, RenderCounters& rc)
{
    Hit h = ClosestHitLBVH(s, sl.pos, sl.dir, sl.tMin, sl.tMax, nullptr);
    rc.extRays++;
    SpecCtx sc;
    sc.spectral = (p.spectral != 0);
    sc.waves = sl.waves;
    const uint32_t rr0 = p.rrRange[0], rr1 = p.rrRange[1];
    uint32_t li = (h.prim == kInvalid) ? kInvalid : LightIndexOfTri(s, h.prim);
    bool isBoundary = (h.prim == kInvalid);
    if(isBoundary || li != kInvalid)
    {
        V3 wO = -sl.dir;
        V4 emission;
        if(isBoundary)
        {
            // LightSkysphere::EmitViaHit (LightsDefault.hpp:405-421)
            emission = IsSkysphere(s.boundaryType)
                ? ConvertRadiance(sc, SkysphereRadiance(s, Normalize(sl.dir))) : Splat4(0.0f);
        }
        else
        {
            // LightPrim::EmitViaHit (LightsDefault.hpp:118-139)
            V3 p0, p1, p2;
            s.Tri(h.prim, p0, p1, p2);
            if(!BaryInRange(h.a, h.b)) emission = Splat4(0.0f);
            else
            {
                V3 n = TriNormal(p0, p1, p2);
                float NdL = Dot(n, wO);
                emission = (!s.lightTwoSided[li] && NdL <= 0.0f)
                    ? Splat4(0.0f) : ConvertRadiance(sc, s.lightRad[li]);
            }
        }
        if(p.sampleMode == MRAY_SAMPLE_PURE)
        {
            // LightWorkFunction (PathTracerRendererShaders.h:306-347)
            sl.depth += 1u;
            if(sl.depth <= rr1) sl.radiance = emission * sl.throughput;
            sl.status = ST_DEAD;
            return;
        }
        // LightWorkFunctionWithNEE (PathTracerRendererShaders.h:448-520)
        if(p.sampleMode == MRAY_SAMPLE_NEE && sl.type != RT_CAMERA && sl.type != RT_SPECULAR)
        {
            sl.status = ST_DEAD;
            return;
        }
        bool switchToMIS = (p.sampleMode == MRAY_SAMPLE_NEE_MIS && sl.type == RT_PATH);
        V4 thr = sl.throughput;
        if(switchToMIS)
        {
            float p0 = sl.prevMatPdf;
            float p1 = isBoundary ? PdfLightBoundary(s, sl.dir)
                                  : PdfLightTri(s, li, h.a, h.b, sl.pos, sl.dir,
                                                (p.flags & MRAY_RENDER_LIGHT_POWER) != 0);
            float misPdf = Balance2(p0, 1.0f, p1, 1.0f);
            thr = thr * p0;
            thr = DivideByPDF(thr, misPdf);
        }
        if((sl.depth + 1u) <= rr1) sl.radiance = sl.radiance + emission * thr;
        sl.status = ST_DEAD;
        return;
    }

    // Surface hit
    uint32_t matIndex = s.surfKey[h.prim];
    Surface surf = GenSurface(s, h.prim, h.a, h.b, sl.dir);
    MatInst mat = MakeMaterial(s, matIndex, surf, sc);
    V3 wO = Normalize(-sl.dir);
    float spec = Specularity(mat);
    bool specular = IsSpecular(spec);

    if(p.sampleMode != MRAY_SAMPLE_PURE)
    {
        // WorkFunctionNEE (PathTracerRendererShaders.h:352-443); 3 numbers
        float rn[3];
        for(int k = 0; k < 3; k++) rn[k] = ToFloat01(PcgNext(sl.rng));
        V3 worldPos = surf.position;
        LightSampleOut ls = SampleLight(s, rn, worldPos, sc, (p.flags & MRAY_RENDER_LIGHT_POWER) != 0);
        // LightSampleOutput::SampledRay (LightSampler.hpp:5-21)
        V3 sdir = Normalize(ls.position - worldPos);
        V3 npos = Nudge(ls.position, -sdir);
        float length = Length(npos - worldPos);
        float stMin = kEps, stMax = length * (1.0f - kLargeEps);
        BxDFEval ev = EvaluateBxDF(mat, sdir, wO);
        V4 thr = sl.throughput * ev.reflectance;
        float pdf;
        if(p.sampleMode == MRAY_SAMPLE_NEE_MIS)
        {
            float bxdfPdf = PdfBxDF(mat, sdir, wO);
            pdf = Balance2(bxdfPdf, 1.0f, ls.pdf, 1.0f);
        }
        else pdf = ls.pdf;
        V4 shadowRad = thr * ls.emission;
        shadowRad = DivideByPDF(shadowRad, pdf);
        if(sc.spectral && ev.dispersed)
        {
            float first = shadowRad.x;
            shadowRad = Splat4(0.0f);
            shadowRad.x = first;
        }
        if(specular) sl.type = RT_SPECULAR;
        else
        {
            V3 nn = surf.geoNormal;
            if(ev.passedThrough) nn = -nn;
            V3 spos = Nudge(worldPos, nn);
            sl.type = RT_SHADOW;
            // CastVisibilityRays + KCAccumulateShadowRaysPT (PathTracerRenderer.cu:11-32)
            rc.shadowRays++;
            bool visible = !AnyHitLBVH(s, spos, sdir, stMin, stMax);
            if((sl.depth + 2u) <= rr1 && visible) sl.radiance = sl.radiance + shadowRad;
        }
    }

    // WorkFunction (PathTracerRendererShaders.h:201-301)
    uint32_t matRN = MatRNCount(mat.type);
    float rn[4];
    for(uint32_t k = 0; k < matRN + 1u; k++) rn[k] = ToFloat01(PcgNext(sl.rng));
    BxDFSample bs = SampleBxDF(mat, wO, rn);
    bs.wIDir = Normalize(bs.wIDir);
    V4 thr = sl.throughput * bs.reflectance;
    if(sc.spectral && bs.dispersed && !sl.dispersed)
    {
        // SpectrumConverter::DisperseWaves: keep the hero wavelength only
        sl.dispersed = true;
    }
    sl.depth += 1u;
    bool dead = (sl.depth >= rr1);
    if(!dead && sl.depth >= rr0 && !specular)
    {
        const float chInv = sc.spectral ? 0.25f : 0.33333333f;
        float rrXi = rn[matRN];
        float rrFactor = Sum4(thr) * chInv;
        // Common::RussianRoulette (DistributionFunctions.h:942-953)
        float prob = Clamp(rrFactor, 0.1f, 1.0f);
        if(rrXi >= prob) dead = true;
        else thr = thr * (1.0f / prob);
    }
    sl.type = specular ? RT_SPECULAR : RT_PATH;
    if(!dead)
    {
        sl.throughput = DivideByPDF(thr, bs.pdf);
        if(p.sampleMode == MRAY_SAMPLE_NEE_MIS) sl.prevMatPdf = bs.pdf;
        V3 nn = surf.geoNormal;
        if(bs.passedThrough) nn = -nn;
        sl.pos = Nudge(bs.wIPos, nn);
        sl.dir = bs.wIDir;
        sl.tMin = kLargeEps;
        sl.tMax = FLT_MAX;
    }
    else sl.status = ST_DEAD;
}

struct RenderOut
{
    uint64_t deadPaths = 0;
    uint64_t extRays = 0;
    uint64_t shadowRays = 0;
    uint32_t iterations = 0;
};

template<class F>
static void ParallelFor(uint32_t n, uint32_t threads, F&& f)
{
    if(threads <= 1 || n < 256)
    {
        for(uint32_t i = 0; i < n; i++) f(i, 0u);
        return;
    }
    std::atomic<uint32_t> next{0};
    std::vector<std::thread> pool;
    const uint32_t chunk = 64;
    for(uint32_t t = 0; t < threads; t++)
    {
        pool.emplace_back([&, t]()
        {
            for(;;)
            {
                uint32_t b = next.fetch_add(chunk);
                if(b >= n) break;
                uint32_t e = std::min(n, b + chunk);
                for(uint32_t i = b; i < e; i++) f(i, t);
            }
        });
    }
    for(auto& th : pool) th.join();
}

// Film add (TextureFilter.cu:672-697): NaN/Inf -> BIG_MAGENTA, weight x128
static inline void FilmAdd(float* film, uint32_t W, uint32_t H, uint32_t pixel, V4 rad, float w)
{
    size_t P = size_t(W) * H;
    float r = rad.x, g = rad.y, b = rad.z;
    if(!(std::isfinite(r) && std::isfinite(g) && std::isfinite(b) && std::isfinite(rad.w)))
    {
        r = 1e7f; g = 0.0f; b = 1e7f;
        w = w * 128.0f;
    }
    film[pixel] += r;
    film[P + pixel] += g;
    film[2 * P + pixel] += b;
    film[3 * P + pixel] += w;
}

static RenderOut Render(const Scene& s, const MRayRenderParams& p, float* film, uint32_t threads,
                        uint32_t maxIterations, uint64_t pathLimitOverride)
{
    RenderOut out;
    const uint32_t W = p.width, H = p.height;
    const uint64_t P = uint64_t(W) * H;
    uint32_t N = p.maxRayCount ? p.maxRayCount : uint32_t(P);
    if(N > P) N = uint32_t(P);
    const float filterRadius = (p.filterRadius > 0.0f) ? p.filterRadius : 1.0f;
    Camera cam = MakeCamera(s.cam);
    std::vector<uint32_t> states(P);
    SeedStates(p.seed, uint32_t(P), states.data());
    std::vector<Slot> slots(N);
    for(uint32_t i = 0; i < N; i++)
    {
        std::memset(&slots[i], 0, sizeof(Slot));
        slots[i].status = ST_INVALID;   // ResetAllPaths (PathTracerRendererBase.cu:205-213)
        slots[i].rng = states[i];
    }
    std::memset(film, 0, sizeof(float) * 4 * P);
    uint64_t counter = 0;
    const uint64_t limit = pathLimitOverride ? pathLimitOverride : uint64_t(p.totalSPP) * P;
    std::vector<RenderCounters> ctr(std::max(threads, 1u));
    std::vector<uint32_t> alive;
    alive.reserve(N);
    for(uint32_t it = 0; it < maxIterations; it++)
    {
        // ReloadPaths: stable partition, dead slots ascending get the next
        // consecutive pixel indices (PathTracerRendererBase.cu:69-203)
        uint64_t k = 0;
        for(uint32_t i = 0; i < N; i++)
        {
            Slot& sl = slots[i];
            if(sl.status == ST_ALIVE) continue;
            if(counter + k < limit)
            {
                uint64_t pix = (counter + k) % P;
                uint32_t px = uint32_t(pix % W), py = uint32_t(pix / W);
                float xi0 = ToFloat01(PcgNext(sl.rng));
                float xi1 = ToFloat01(PcgNext(sl.rng));
                CamRay cr = GenCamRay(cam, px, py, W, H, xi0, xi1, filterRadius);
                sl.pos = cr.pos; sl.dir = cr.dir; sl.tMin = cr.tMin; sl.tMax = cr.tMax;
                sl.filterWeight = cr.weight;
                sl.pixel = uint32_t(pix);
                // KCInitPathStateIndirect
                sl.radiance = Splat4(0.0f);
                sl.throughput = Splat4(1.0f);
                sl.depth = 0;
                sl.status = ST_ALIVE;
                sl.type = RT_CAMERA;
                sl.dispersed = false;
                sl.prevMatPdf = 0.0f;
                if(p.spectral)
                {
                    float xw = ToFloat01(PcgNext(sl.rng));
                    SampleWavelengths(xw, sl.waves, sl.wavePdf);
                }
                k++;
            }
            else
            {
                sl.status = ST_INVALID;
            }
        }
        counter += k;
        alive.clear();
        for(uint32_t i = 0; i < N; i++) if(slots[i].status == ST_ALIVE) alive.push_back(i);
        if(alive.empty()) break;
        out.iterations++;
        for(auto& c : ctr) c = RenderCounters{};
        ParallelFor(uint32_t(alive.size()), threads, [&](uint32_t i, uint32_t t)
        {
            BounceSlot(s, p, slots[alive[i]], ctr[t]);
        });
        for(auto& c : ctr) { out.extRays += c.extRays; out.shadowRays += c.shadowRays; }
        // Dead paths to the film (AddRadianceToRenderBufferThroughput)
        for(uint32_t i : alive)
        {
            Slot& sl = slots[i];
            if(sl.status != ST_DEAD) continue;
            V4 rad = sl.radiance;
            if(p.spectral) rad = SpectrumToRGB(rad, sl.waves, sl.wavePdf, sl.dispersed);
            FilmAdd(film, W, H, sl.pixel, rad, sl.filterWeight);
            sl.status = ST_INVALID;
            out.deadPaths++;
        }
    }
    return out;
}

// ------------------------------------------------------------------------
// Tiled rendering: ImageTiler (Tracer/RenderImage.cpp:20-136) and
// PathTracerRendererT::DoLatencyRender (TracerDLL/PathTracerRenderer.cu:
// 1076-1180) with RNGGroupIndependent::SetupRange (Random.cu:728-774).
// ------------------------------------------------------------------------
// ImageTiler::FindOptimumTileSize (RenderImage.cpp:20-52); a hint below one
// pixel per axis is taken as 1 (the reference would divide by zero).
static void TileSize(uint32_t W, uint32_t H, uint32_t hint, uint32_t out[2])
{
    float aspectRatio = float(W) / float(H);
    float factor = std::sqrt(float(hint) / aspectRatio);
    uint32_t tileHint[2] = {std::max(1u, uint32_t(std::round(aspectRatio * factor))),
                            std::max(1u, uint32_t(std::round(factor)))};
    uint32_t fbSize[2] = {W, H};
    for(int i = 0; i < 2; i++)
    {
        if(fbSize[i] < tileHint[i])
        {
            out[i] = fbSize[i];
            continue;
        }
        float tileCountF = float(fbSize[i]) / float(tileHint[i]);
        uint32_t tileCount = uint32_t(std::round(tileCountF));
        uint32_t result = fbSize[i] / tileCount;
        uint32_t residual = fbSize[i] % tileCount;
        residual = (residual + tileCount - 1) / tileCount;
        out[i] = result + residual;
    }
}

// CameraPinhole::GenerateSubCamera (CamerasDefault.hpp:226-247)
static Camera SubCamera(const Camera& c, uint32_t sx, uint32_t sy, uint32_t cx, uint32_t cy)
{
    Camera t = c;
    V2 delta = {c.planeSize.x / float(cx), c.planeSize.y / float(cy)};
    V2 dist = {float(sx) * delta.x, float(sy) * delta.y};
    t.bottomLeft = c.bottomLeft + ((dist.x * c.right) + (dist.y * c.up));
    t.planeSize = delta;
    return t;
}

struct TileCall { uint32_t tile, x0, y0, x1, y1, spp; };

// One call per passCounts[c]: that many more samples of the current tile,
// run until every one of its paths is in the film, then the next tile
// (row-major, x fastest).  Slot j holds the random state of the tile's
// pixel (x0 + j mod w, y0 + j / w): loaded when the tile becomes current,
// stored back when another tile replaces it; slots past a narrower edge
// tile keep whatever they held.  Path g of a tile is its pixel g mod (w h),
// rendered through the tile's sub-camera at w x h.  film [4][H][W]
// accumulates every call.
static RenderOut RenderTiled(const Scene& s, const MRayRenderParams& p, const uint32_t* passCounts,
                             uint32_t calls, float* film, uint32_t threads, TileCall* out)
{
    RenderOut ro;
    const uint32_t W = p.width, H = p.height;
    const uint64_t P = uint64_t(W) * H;
    uint32_t ts[2];
    TileSize(W, H, p.maxRayCount ? p.maxRayCount : (1u << 21), ts);
    const uint32_t cx = (W + ts[0] - 1) / ts[0], cy = (H + ts[1] - 1) / ts[1];
    const uint32_t T = cx * cy;
    const uint32_t N = ts[0] * ts[1];
    const float filterRadius = (p.filterRadius > 0.0f) ? p.filterRadius : 1.0f;
    const Camera cam = MakeCamera(s.cam);
    std::vector<uint32_t> pixRng(P);
    SeedStates(p.seed, uint32_t(P), pixRng.data());
    std::vector<uint32_t> slotSeeds(N);
    SeedStates(p.seed, N, slotSeeds.data());
    std::vector<Slot> slots(N);
    for(uint32_t i = 0; i < N; i++)
    {
        std::memset(&slots[i], 0, sizeof(Slot));
        slots[i].status = ST_INVALID;
        slots[i].rng = slotSeeds[i];
    }
    std::memset(film, 0, sizeof(float) * 4 * P);
    std::vector<uint64_t> tilePaths(T, 0);
    std::vector<uint32_t> tileSPP(T, 0);
    uint32_t cur = 0;
    int loaded = -1;
    auto rect = [&](uint32_t t, uint32_t& x0, uint32_t& y0, uint32_t& w, uint32_t& h)
    {
        x0 = (t % cx) * ts[0];
        y0 = (t / cx) * ts[1];
        w = std::min(ts[0], W - x0);
        h = std::min(ts[1], H - y0);
    };
    std::vector<RenderCounters> ctr(std::max(threads, 1u));
    std::vector<uint32_t> alive;
    alive.reserve(N);
    for(uint32_t c = 0; c < calls; c++)
    {
        const uint32_t t = cur;
        uint32_t x0, y0, w, h;
        rect(t, x0, y0, w, h);
        if(loaded != int(t))
        {
            if(loaded >= 0)
            {
                uint32_t a0, b0, aw, bh;
                rect(uint32_t(loaded), a0, b0, aw, bh);
                for(uint32_t j = 0; j < aw * bh; j++)
                    pixRng[size_t(b0 + j / aw) * W + a0 + j % aw] = slots[j].rng;
            }
            for(uint32_t j = 0; j < w * h; j++)
                slots[j].rng = pixRng[size_t(y0 + j / w) * W + x0 + j % w];
            loaded = int(t);
        }
        const Camera sub = SubCamera(cam, t % cx, t / cx, cx, cy);
        const uint64_t tp = uint64_t(w) * h;
        const uint32_t spp = tileSPP[t] + passCounts[c];
        const uint64_t limit = uint64_t(spp) * tp;
        uint64_t counter = tilePaths[t];
        for(;;)
        {
            // ReloadPaths over the tile's path indices (PathTracerRendererBase.cu:69-203)
            uint64_t k = 0;
            for(uint32_t i = 0; i < N; i++)
            {
                Slot& sl = slots[i];
                if(sl.status == ST_ALIVE) continue;
                if(counter + k < limit)
                {
                    uint64_t local = (counter + k) % tp;
                    uint32_t px = uint32_t(local % w), py = uint32_t(local / w);
                    float xi0 = ToFloat01(PcgNext(sl.rng));
                    float xi1 = ToFloat01(PcgNext(sl.rng));
                    CamRay cr = GenCamRay(sub, px, py, w, h, xi0, xi1, filterRadius);
                    sl.pos = cr.pos; sl.dir = cr.dir; sl.tMin = cr.tMin; sl.tMax = cr.tMax;
                    sl.filterWeight = cr.weight;
                    sl.pixel = (y0 + py) * W + x0 + px;
                    sl.radiance = Splat4(0.0f);
                    sl.throughput = Splat4(1.0f);
                    sl.depth = 0;
                    sl.status = ST_ALIVE;
                    sl.type = RT_CAMERA;
                    sl.dispersed = false;
                    sl.prevMatPdf = 0.0f;
                    if(p.spectral)
                    {
                        float xw = ToFloat01(PcgNext(sl.rng));
                        SampleWavelengths(xw, sl.waves, sl.wavePdf);
                    }
                    k++;
                }
                else sl.status = ST_INVALID;
            }
            counter += k;
            alive.clear();
            for(uint32_t i = 0; i < N; i++) if(slots[i].status == ST_ALIVE) alive.push_back(i);
            if(alive.empty()) break;
            ro.iterations++;
            for(auto& cc : ctr) cc = RenderCounters{};
            ParallelFor(uint32_t(alive.size()), threads, [&](uint32_t i, uint32_t th)
            {
                BounceSlot(s, p, slots[alive[i]], ctr[th]);
            });
            for(auto& cc : ctr) { ro.extRays += cc.extRays; ro.shadowRays += cc.shadowRays; }
            for(uint32_t i : alive)
            {
                Slot& sl = slots[i];
                if(sl.status != ST_DEAD) continue;
                V4 rad = sl.radiance;
                if(p.spectral) rad = SpectrumToRGB(rad, sl.waves, sl.wavePdf, sl.dispersed);
                FilmAdd(film, W, H, sl.pixel, rad, sl.filterWeight);
                sl.status = ST_INVALID;
                ro.deadPaths++;
            }
        }
        tilePaths[t] = limit;
        tileSPP[t] = spp;
        if(out) out[c] = TileCall{t, x0, y0, x0 + w, y0 + h, spp};
        cur = (t + 1) % T;
    }
    return ro;
}

// Batch of views over several scenes (mray_batch_create semantics; no
// reference counterpart): one slot pool, global path indices taken view by
// view (view k owns [sum of earlier W*H*spp, + its own)), each slot shaded
// in its view's scene and splatted into its view's film.  The schedule is
// Render's: ascending dead slots take the next consecutive indices.
struct BatchViewRT
{
    const Scene* s;
    Camera cam;
    uint32_t W, H;
    uint64_t P, begin, end;
    float* film;
};

static RenderOut RenderBatch(const std::vector<BatchViewRT>& views, const MRayRenderParams& p,
                             uint32_t threads)
{
    RenderOut out;
    uint64_t pixels = 0;
    for(const BatchViewRT& v : views) pixels += v.P;
    const uint64_t limit = views.back().end;
    uint32_t N = p.maxRayCount ? p.maxRayCount : uint32_t(std::min<uint64_t>(pixels, 1u << 21));
    if(N > pixels) N = uint32_t(pixels);
    const float filterRadius = (p.filterRadius > 0.0f) ? p.filterRadius : 1.0f;
    std::vector<uint32_t> states(N);
    SeedStates(p.seed, N, states.data());
    std::vector<Slot> slots(N);
    std::vector<uint32_t> slotView(N, 0);
    for(uint32_t i = 0; i < N; i++)
    {
        std::memset(&slots[i], 0, sizeof(Slot));
        slots[i].status = ST_INVALID;
        slots[i].rng = states[i];
    }
    for(const BatchViewRT& v : views) std::memset(v.film, 0, sizeof(float) * 4 * v.P);
    uint64_t counter = 0;
    std::vector<RenderCounters> ctr(std::max(threads, 1u));
    std::vector<uint32_t> alive;
    alive.reserve(N);
    for(;;)
    {
        uint64_t k = 0;
        for(uint32_t i = 0; i < N; i++)
        {
            Slot& sl = slots[i];
            if(sl.status == ST_ALIVE) continue;
            if(counter + k < limit)
            {
                const uint64_t g = counter + k;
                uint32_t vi = 0;
                while(vi + 1 < views.size() && views[vi + 1].begin <= g) vi++;
                const BatchViewRT& v = views[vi];
                uint64_t pix = (g - v.begin) % v.P;
                uint32_t px = uint32_t(pix % v.W), py = uint32_t(pix / v.W);
                float xi0 = ToFloat01(PcgNext(sl.rng));
                float xi1 = ToFloat01(PcgNext(sl.rng));
                CamRay cr = GenCamRay(v.cam, px, py, v.W, v.H, xi0, xi1, filterRadius);
                sl.pos = cr.pos; sl.dir = cr.dir; sl.tMin = cr.tMin; sl.tMax = cr.tMax;
                sl.filterWeight = cr.weight;
                sl.pixel = uint32_t(pix);
                sl.radiance = Splat4(0.0f);
                sl.throughput = Splat4(1.0f);
                sl.depth = 0;
                sl.status = ST_ALIVE;
                sl.type = RT_CAMERA;
                sl.dispersed = false;
                sl.prevMatPdf = 0.0f;
                if(p.spectral)
                {
                    float xw = ToFloat01(PcgNext(sl.rng));
                    SampleWavelengths(xw, sl.waves, sl.wavePdf);
                }
                slotView[i] = vi;
                k++;
            }
            else sl.status = ST_INVALID;
        }
        counter += k;
        alive.clear();
        for(uint32_t i = 0; i < N; i++) if(slots[i].status == ST_ALIVE) alive.push_back(i);
        if(alive.empty()) break;
        out.iterations++;
        for(auto& c : ctr) c = RenderCounters{};
        ParallelFor(uint32_t(alive.size()), threads, [&](uint32_t i, uint32_t t)
        {
            uint32_t si = alive[i];
            BounceSlot(*views[slotView[si]].s, p, slots[si], ctr[t]);
        });
        for(auto& c : ctr) { out.extRays += c.extRays; out.shadowRays += c.shadowRays; }
        for(uint32_t i : alive)
        {
            Slot& sl = slots[i];
            if(sl.status != ST_DEAD) continue;
            const BatchViewRT& v = views[slotView[i]];
            V4 rad = sl.radiance;
            if(p.spectral) rad = SpectrumToRGB(rad, sl.waves, sl.wavePdf, sl.dispersed);
            FilmAdd(v.film, v.W, v.H, sl.pixel, rad, sl.filterWeight);
            sl.status = ST_INVALID;
            out.deadPaths++;
        }
    }
    return out;
}

} // namespace rc

// ==========================================================================
// C ABI for the tests (ctypes)
// ==========================================================================
using namespace rc;

extern "C" {

const char* rc_version(void) { return "refcpu-1"; }

uint64_t rc_morton63_compose(uint32_t x, uint32_t y, uint32_t z) { return Compose3D_64(x, y, z); }
uint32_t rc_morton30_compose(uint32_t x, uint32_t y, uint32_t z) { return Compose3D_32(x, y, z); }
void rc_morton63_decompose(uint64_t c, uint32_t* out)
{
    out[0] = Shrink3D_64(c);
    out[1] = Shrink3D_64(c >> 1);
    out[2] = Shrink3D_64(c >> 2);
}
void rc_morton30_decompose(uint32_t c, uint32_t* out)
{
    out[0] = Shrink3D_32(c);
    out[1] = Shrink3D_32(c >> 1);
    out[2] = Shrink3D_32(c >> 2);
}

void rc_prim_centers_aabbs(const float* pos, const uint32_t* idx, uint32_t n, float* centers,
                           float* aabbs)
{
    for(uint32_t t = 0; t < n; t++)
    {
        const float* a = pos + 3 * size_t(idx[3 * t + 0]);
        const float* b = pos + 3 * size_t(idx[3 * t + 1]);
        const float* c = pos + 3 * size_t(idx[3 * t + 2]);
        V3 p0{a[0], a[1], a[2]}, p1{b[0], b[1], b[2]}, p2{c[0], c[1], c[2]};
        V3 ctr = TriCenter(p0, p1, p2);
        if(centers) { centers[3 * t] = ctr.x; centers[3 * t + 1] = ctr.y; centers[3 * t + 2] = ctr.z; }
        if(aabbs)
        {
            float* bb = aabbs + 6 * size_t(t);
            bb[0] = MinF(MinF(p0.x, p1.x), p2.x);
            bb[1] = MinF(MinF(p0.y, p1.y), p2.y);
            bb[2] = MinF(MinF(p0.z, p1.z), p2.z);
            bb[3] = MaxF(MaxF(p0.x, p1.x), p2.x);
            bb[4] = MaxF(MaxF(p0.y, p1.y), p2.y);
            bb[5] = MaxF(MaxF(p0.z, p1.z), p2.z);
        }
    }
}

void rc_morton63_codes(const float* centers, uint32_t n, const float* aabb, uint64_t* codes)
{
    V3 mn{aabb[0], aabb[1], aabb[2]}, mx{aabb[3], aabb[4], aabb[5]};
    for(uint32_t i = 0; i < n; i++)
        codes[i] = Morton63Of(V3{centers[3 * i], centers[3 * i + 1], centers[3 * i + 2]}, mn, mx);
}
void rc_morton30_codes(const float* centers, uint32_t n, const float* aabb, uint32_t* codes)
{
    V3 mn{aabb[0], aabb[1], aabb[2]}, mx{aabb[3], aabb[4], aabb[5]};
    for(uint32_t i = 0; i < n; i++)
        codes[i] = Morton30Of(V3{centers[3 * i], centers[3 * i + 1], centers[3 * i + 2]}, mn, mx);
}

// Karras construction + refit from sorted codes (AcceleratorLBVH.cu:170-435)
void rc_lbvh_from_sorted(const uint64_t* sortedCodes, const uint32_t* leafOrder,
                         const float* leafAabbs, uint32_t n, uint32_t* nodes, float* boxes)
{
    uint32_t nodeCount = (n > 1) ? n - 1 : 1;
    for(size_t i = 0; i < size_t(nodeCount) * 3; i++) nodes[i] = kInvalid;
    std::vector<uint32_t> leafParent(n, kInvalid);
    std::vector<uint32_t> range(size_t(nodeCount) * 2, 0u);
    KarrasBuild(sortedCodes, leafOrder, int32_t(n), nodes, leafParent.data(), range.data(), 0);
    // boxes only for a well-formed tree (a malformed one has no defined refit)
    if(ValidateTree(nodes, range.data(), n))
        RefitBoxes(nodes, int32_t(nodeCount), leafAabbs, boxes);
}

int rc_lbvh_tree_valid(const uint64_t* sortedCodes, const uint32_t* leafOrder, uint32_t n,
                       int strict)
{
    uint32_t nodeCount = (n > 1) ? n - 1 : 1;
    std::vector<uint32_t> nodes(size_t(nodeCount) * 3, kInvalid), range(size_t(nodeCount) * 2);
    std::vector<uint32_t> leafParent(n, kInvalid);
    KarrasBuild(sortedCodes, leafOrder, int32_t(n), nodes.data(), leafParent.data(), range.data(),
                strict);
    return ValidateTree(nodes.data(), range.data(), n) ? 1 : 0;
}

void* rc_scene_create(const MRaySceneDesc* d)
{
    Scene* s = new Scene();
    s->vCount = d->vertexCount;
    s->tCount = d->triangleCount;
    s->pos.resize(d->vertexCount);
    s->tbn.resize(d->vertexCount);
    for(uint32_t v = 0; v < d->vertexCount; v++)
    {
        s->pos[v] = {d->hPositions[3 * v], d->hPositions[3 * v + 1], d->hPositions[3 * v + 2]};
        if(d->hTbnQuats)
            s->tbn[v] = {d->hTbnQuats[4 * v], d->hTbnQuats[4 * v + 1], d->hTbnQuats[4 * v + 2],
                         d->hTbnQuats[4 * v + 3]};
        else s->tbn[v] = {1, 0, 0, 0};
    }
    s->idx.assign(d->hIndices, d->hIndices + 3 * size_t(d->triangleCount));
    if(d->hSurfaceKeys) s->surfKey.assign(d->hSurfaceKeys, d->hSurfaceKeys + d->triangleCount);
    else s->surfKey.assign(d->triangleCount, 0u);
    if(d->hCullBackFace) s->cull.assign(d->hCullBackFace, d->hCullBackFace + d->triangleCount);
    else s->cull.assign(d->triangleCount, 0u);
    for(uint32_t m = 0; m < d->materialCount; m++)
    {
        const MRayMaterial& M = d->hMaterials[m];
        Material mm;
        mm.type = M.type;
        mm.albedo = {M.albedo[0], M.albedo[1], M.albedo[2]};
        mm.roughness = M.roughness;
        mm.metallic = M.metallic;
        mm.specular = M.specular;
        mm.cauchyFront = {M.cauchyFront[0], M.cauchyFront[1], M.cauchyFront[2]};
        mm.cauchyBack = {M.cauchyBack[0], M.cauchyBack[1], M.cauchyBack[2]};
        s->mats.push_back(mm);
    }
    for(uint32_t l = 0; l < d->lightCount; l++)
    {
        s->lightTri.push_back(d->hLightTriangles[l]);
        s->lightRad.push_back({d->hLightRadiance[3 * l], d->hLightRadiance[3 * l + 1],
                               d->hLightRadiance[3 * l + 2]});
        s->lightTwoSided.push_back(d->hLightTwoSided ? d->hLightTwoSided[l] : 0);
    }
    s->boundaryType = d->boundaryType;
    s->boundaryRad = {d->boundaryRadiance[0], d->boundaryRadiance[1], d->boundaryRadiance[2]};
    s->sceneDiameter = d->sceneDiameter;
    s->cam = d->camera;
    if(s->tCount) BuildLBVH(*s);
    BuildLightPowerTable(*s);
    return s;
}

void rc_scene_destroy(void* h) { delete static_cast<Scene*>(h); }

void rc_scene_lbvh(void* h, uint64_t* codes, uint32_t* leafOrder, uint32_t* nodes, float* boxes,
                   float* sceneAabb)
{
    const Scene& s = *static_cast<Scene*>(h);
    size_t n = s.tCount;
    if(codes) std::memcpy(codes, s.codes.data(), n * 8);
    if(leafOrder) std::memcpy(leafOrder, s.leafOrder.data(), n * 4);
    if(nodes) std::memcpy(nodes, s.nodes.data(), s.nodes.size() * 4);
    if(boxes) std::memcpy(boxes, s.boxes.data(), s.boxes.size() * 4);
    if(sceneAabb)
    {
        sceneAabb[0] = s.sceneMin.x; sceneAabb[1] = s.sceneMin.y; sceneAabb[2] = s.sceneMin.z;
        sceneAabb[3] = s.sceneMax.x; sceneAabb[4] = s.sceneMax.y; sceneAabb[5] = s.sceneMax.z;
    }
}

// rays [n][8] = {pos, tMin, dir, tMax}; mode 0 = LBVH, 1 = brute force
void rc_cast_closest(void* h, const float* rays, uint32_t n, int brute, uint32_t* prim, float* t,
                     float* bary, uint64_t* counters, uint32_t threads)
{
    const Scene& s = *static_cast<Scene*>(h);
    std::vector<TravCounters> tc(std::max(threads, 1u));
    ParallelFor(n, threads, [&](uint32_t i, uint32_t th)
    {
        const float* r = rays + 8 * size_t(i);
        V3 pos{r[0], r[1], r[2]}, dir{r[4], r[5], r[6]};
        Hit hh = brute ? ClosestHitBrute(s, pos, dir, r[3], r[7])
                       : ClosestHitLBVH(s, pos, dir, r[3], r[7], counters ? &tc[th] : nullptr);
        prim[i] = hh.prim;
        if(t) t[i] = (hh.prim == kInvalid) ? r[7] : hh.t;
        if(bary) { bary[2 * i] = hh.a; bary[2 * i + 1] = hh.b; }
    });
    if(counters)
    {
        counters[0] = counters[1] = 0;
        for(auto& c : tc) { counters[0] += c.nodes; counters[1] += c.tris; }
    }
}

void rc_cast_visibility(void* h, const float* rays, uint32_t n, uint32_t* visible, uint32_t threads)
{
    const Scene& s = *static_cast<Scene*>(h);
    ParallelFor(n, threads, [&](uint32_t i, uint32_t)
    {
        const float* r = rays + 8 * size_t(i);
        visible[i] = AnyHitLBVH(s, V3{r[0], r[1], r[2]}, V3{r[4], r[5], r[6]}, r[3], r[7]) ? 0u : 1u;
    });
}

void rc_seed_states(uint64_t seed, uint32_t count, uint32_t* out) { SeedStates(seed, count, out); }
uint32_t rc_pcg_next(uint32_t* state) { return PcgNext(*state); }
float rc_to_float01(uint32_t v) { return ToFloat01(v); }

// First-iteration camera rays of slots 0..n-1 (pixel i, PCG state i).
void rc_camera_rays(void* h, const MRayRenderParams* p, uint32_t n, float* rays, float* weights)
{
    const Scene& s = *static_cast<Scene*>(h);
    Camera cam = MakeCamera(s.cam);
    uint64_t P = uint64_t(p->width) * p->height;
    std::vector<uint32_t> st(P);
    SeedStates(p->seed, uint32_t(P), st.data());
    float fr = (p->filterRadius > 0.0f) ? p->filterRadius : 1.0f;
    for(uint32_t i = 0; i < n; i++)
    {
        uint32_t pix = uint32_t(i % P);
        uint32_t r = st[i % P];
        float xi0 = ToFloat01(PcgNext(r));
        float xi1 = ToFloat01(PcgNext(r));
        CamRay cr = GenCamRay(cam, pix % p->width, pix / p->width, p->width, p->height, xi0, xi1, fr);
        float* o = rays + 8 * size_t(i);
        o[0] = cr.pos.x; o[1] = cr.pos.y; o[2] = cr.pos.z; o[3] = cr.tMin;
        o[4] = cr.dir.x; o[5] = cr.dir.y; o[6] = cr.dir.z; o[7] = cr.tMax;
        if(weights) weights[i] = cr.weight;
    }
}

// Full render; film [4][H][W].  maxIterations 0 = until done.  stats[4] =
// dead paths, extension rays, shadow rays, iterations.
int rc_render(void* h, const MRayRenderParams* p, float* film, uint32_t threads,
              uint32_t maxIterations, uint64_t pathLimit, uint64_t* stats)
{
    const Scene& s = *static_cast<Scene*>(h);
    if(p->spectral && !gSpec.loaded) return 1;
    if(threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    RenderOut o = Render(s, *p, film, threads, maxIterations ? maxIterations : 0xFFFFFFFFu, pathLimit);
    if(stats)
    {
        stats[0] = o.deadPaths;
        stats[1] = o.extRays;
        stats[2] = o.shadowRays;
        stats[3] = o.iterations;
    }
    return 0;
}

// Tiled render: calls[c] = passCount of call c; film [4][H][W] accumulates
// every call; tiles[6 * c] = tile, x0, y0, x1, y1, tile spp of call c;
// tileInfo[4] = tile size (2), tile count (2).  stats[4] as rc_render.
int rc_render_tiled(void* h, const MRayRenderParams* p, const uint32_t* calls, uint32_t nCalls,
                    float* film, uint32_t threads, uint32_t* tiles, uint32_t* tileInfo,
                    uint64_t* stats)
{
    const Scene& s = *static_cast<Scene*>(h);
    if(p->spectral && !gSpec.loaded) return 1;
    if(threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    std::vector<TileCall> tc(nCalls);
    RenderOut o = RenderTiled(s, *p, calls, nCalls, film, threads, tc.data());
    uint32_t ts[2];
    TileSize(p->width, p->height, p->maxRayCount ? p->maxRayCount : (1u << 21), ts);
    if(tileInfo)
    {
        tileInfo[0] = ts[0];
        tileInfo[1] = ts[1];
        tileInfo[2] = (p->width + ts[0] - 1) / ts[0];
        tileInfo[3] = (p->height + ts[1] - 1) / ts[1];
    }
    if(tiles)
        for(uint32_t c = 0; c < nCalls; c++)
        {
            const TileCall& t = tc[c];
            uint32_t* o6 = tiles + 6 * size_t(c);
            o6[0] = t.tile; o6[1] = t.x0; o6[2] = t.y0; o6[3] = t.x1; o6[4] = t.y1; o6[5] = t.spp;
        }
    if(stats)
    {
        stats[0] = o.deadPaths;
        stats[1] = o.extRays;
        stats[2] = o.shadowRays;
        stats[3] = o.iterations;
    }
    return 0;
}

// Batch render; views as MRayBatchView (scene index into `scenes`), films[v]
// [4][H][W] per view.  stats[4] as rc_render.
int rc_render_batch(void* const* scenes, uint32_t nScenes, const MRayBatchView* views,
                    uint32_t nViews, const MRayRenderParams* p, float* const* films,
                    uint32_t threads, uint64_t* stats)
{
    if(p->spectral && !gSpec.loaded) return 1;
    if(nViews == 0) return 2;
    if(threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
    std::vector<BatchViewRT> vs(nViews);
    uint64_t paths = 0;
    for(uint32_t k = 0; k < nViews; k++)
    {
        const MRayBatchView& v = views[k];
        if(v.scene >= nScenes) return 2;
        const Scene* s = static_cast<const Scene*>(scenes[v.scene]);
        BatchViewRT& o = vs[k];
        o.s = s;
        o.cam = MakeCamera(v.useCamera ? v.camera : s->cam);
        o.W = v.width;
        o.H = v.height;
        o.P = uint64_t(v.width) * v.height;
        o.begin = paths;
        o.end = paths + o.P * v.spp;
        o.film = films[k];
        paths = o.end;
    }
    RenderOut o = RenderBatch(vs, *p, threads);
    if(stats)
    {
        stats[0] = o.deadPaths;
        stats[1] = o.extRays;
        stats[2] = o.shadowRays;
        stats[3] = o.iterations;
    }
    return 0;
}

// Spectral tables.  The observer, illuminant and XYZ->RGB matrix are built
// here from the CIE 1931 / D65 data (public standard data, the same values the
// reference renders with), restating the reference's load-time normalisation
// (Tracer/SpectrumContext.cu:401-426): each observer channel times the float
// reciprocal of its integral, the illuminant times integral(ybar) / 10566.864;
// and Color::GenRGBToXYZ (Core/ColorFunctions.h:347-361) in float for Rec.709
// with the D65 white (0.31272, 0.32903).  Only the coefficient LUT comes from
// the caller: [3][N][N][N][4] (c0, c1, c2, 0), x fastest.
static void Inverse3(const float m[9], float out[9])
{
    const float a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6], h = m[7],
                i = m[8];
    const float A = e * i - f * h, B = f * g - d * i, Cc = d * h - e * g;
    const float det = a * A + b * B + c * Cc;
    const float r[9] = {A, c * h - b * i, b * f - c * e, B, a * i - c * g, c * d - a * f,
                        Cc, b * g - a * h, a * e - b * d};
    const float inv = 1.0f / det;
    for(int k = 0; k < 9; k++) out[k] = r[k] * inv;
}

static void BuildColourTables()
{
    using namespace mray;
    gSpec.cmf.resize(size_t(kCieRows) * 3);
    gSpec.illum.resize(kCieRows);
    const float w[3] = {1.0f / kCieIntegralX, 1.0f / kCieIntegralY, 1.0f / kCieIntegralZ};
    float illumFactor = kCieIntegralY;
    illumFactor /= kD65NormFactor;
    for(int i = 0; i < kCieRows; i++)
    {
        for(int k = 0; k < 3; k++) gSpec.cmf[3 * i + k] = kCieObserverD65[i][k] * w[k];
        gSpec.illum[i] = kCieObserverD65[i][3] * illumFactor;
    }
    // YxyToXYZ of each primary and the white at Y = 1 (ColorFunctions.h:249-257)
    auto xyz = [](double xd, double yd, float out[3])
    {
        const float x = float(xd), y = float(yd);
        const float yy = 1.0f / y;
        out[0] = x * yy;
        out[1] = 1.0f;
        out[2] = (1.0f - x - y) * yy;
    };
    const double* p = kRec709Primaries;
    float r[3], g[3], b[3], wt[3];
    xyz(p[0], p[1], r);
    xyz(p[2], p[3], g);
    xyz(p[4], p[5], b);
    xyz(p[6], p[7], wt);
    // columns = primaries; scaled so RGB (1, 1, 1) is the white
    const float P[9] = {r[0], g[0], b[0], r[1], g[1], b[1], r[2], g[2], b[2]};
    float Pi[9];
    Inverse3(P, Pi);
    float S[3];
    for(int k = 0; k < 3; k++) S[k] = Pi[3 * k] * wt[0] + Pi[3 * k + 1] * wt[1] + Pi[3 * k + 2] * wt[2];
    float toXyz[9];
    for(int row = 0; row < 3; row++)
        for(int k = 0; k < 3; k++) toXyz[3 * row + k] = P[3 * row + k] * S[k];
    Inverse3(toXyz, gSpec.xyzToRgb);
}

void rc_spectral_load(int N, const float* lut)
{
    gSpec.N = N;
    gSpec.lut.assign(lut, lut + size_t(3) * N * N * N * 4);
    BuildColourTables();
    gSpec.loaded = true;
}

// The oracle's observer [471][3], illuminant [471] and XYZ->RGB [9].
void rc_spectral_colour_tables(float* cmf, float* illum, float* xyzToRgb)
{
    BuildColourTables();
    std::memcpy(cmf, gSpec.cmf.data(), gSpec.cmf.size() * sizeof(float));
    std::memcpy(illum, gSpec.illum.data(), gSpec.illum.size() * sizeof(float));
    std::memcpy(xyzToRgb, gSpec.xyzToRgb, 9 * sizeof(float));
}

void rc_sample_wavelengths(const float* xi, uint32_t n, float* waves, float* pdfs)
{
    for(uint32_t i = 0; i < n; i++)
    {
        V4 w, p;
        SampleWavelengths(xi[i], w, p);
        waves[4 * i] = w.x; waves[4 * i + 1] = w.y; waves[4 * i + 2] = w.z; waves[4 * i + 3] = w.w;
        pdfs[4 * i] = p.x; pdfs[4 * i + 1] = p.y; pdfs[4 * i + 2] = p.z; pdfs[4 * i + 3] = p.w;
    }
}

void rc_rgb_to_spectrum(const float* rgb, const float* waves, uint32_t n, int radiance, float* out)
{
    for(uint32_t i = 0; i < n; i++)
    {
        SpecCtx sc;
        sc.spectral = true;
        sc.waves = {waves[4 * i], waves[4 * i + 1], waves[4 * i + 2], waves[4 * i + 3]};
        V3 c{rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]};
        V4 r = radiance ? ConvertRadiance(sc, c) : ConvertAlbedo(sc, c);
        out[4 * i] = r.x; out[4 * i + 1] = r.y; out[4 * i + 2] = r.z; out[4 * i + 3] = r.w;
    }
}

void rc_spectrum_to_rgb(const float* value, const float* waves, const float* pdfs, uint32_t n,
                        float* out)
{
    for(uint32_t i = 0; i < n; i++)
    {
        V4 v{value[4 * i], value[4 * i + 1], value[4 * i + 2], value[4 * i + 3]};
        V4 w{waves[4 * i], waves[4 * i + 1], waves[4 * i + 2], waves[4 * i + 3]};
        V4 p{pdfs[4 * i], pdfs[4 * i + 1], pdfs[4 * i + 2], pdfs[4 * i + 3]};
        V4 r = SpectrumToRGB(v, w, p, false);
        out[3 * i] = r.x; out[3 * i + 1] = r.y; out[3 * i + 2] = r.z;
    }
}

// BSDF probes for the closed-form tests: material m on a surface whose
// shading frame is `tbn` (w,x,y,z), geometric normal n.  op 0 = sample
// (rn[4] used) -> out {dir xyz, pdf, refl rgbw}; op 1 = evaluate(wI) ->
// refl rgbw; op 2 = pdf(wI).
void rc_bsdf_probe(const MRayMaterial* M, const float* tbn, const float* geoN, int backSide,
                   const float* wO, const float* wI, const float* rn, int op, float* out)
{
    Scene dummy;
    Material mm;
    mm.type = M->type;
    mm.albedo = {M->albedo[0], M->albedo[1], M->albedo[2]};
    mm.roughness = M->roughness;
    mm.metallic = M->metallic;
    mm.specular = M->specular;
    mm.cauchyFront = {M->cauchyFront[0], M->cauchyFront[1], M->cauchyFront[2]};
    mm.cauchyBack = {M->cauchyBack[0], M->cauchyBack[1], M->cauchyBack[2]};
    dummy.mats.push_back(mm);
    Surface sf;
    sf.position = {0, 0, 0};
    sf.geoNormal = {geoN[0], geoN[1], geoN[2]};
    sf.tbn = {tbn[0], tbn[1], tbn[2], tbn[3]};
    sf.backSide = backSide != 0;
    SpecCtx sc;
    MatInst m = MakeMaterial(dummy, 0, sf, sc);
    V3 o{wO[0], wO[1], wO[2]};
    if(op == 0)
    {
        BxDFSample b = SampleBxDF(m, o, rn);
        out[0] = b.wIDir.x; out[1] = b.wIDir.y; out[2] = b.wIDir.z; out[3] = b.pdf;
        out[4] = b.reflectance.x; out[5] = b.reflectance.y; out[6] = b.reflectance.z;
        out[7] = b.reflectance.w;
    }
    else if(op == 1)
    {
        BxDFEval e = EvaluateBxDF(m, V3{wI[0], wI[1], wI[2]}, o);
        out[0] = e.reflectance.x; out[1] = e.reflectance.y; out[2] = e.reflectance.z;
        out[3] = e.reflectance.w;
    }
    else
    {
        out[0] = PdfBxDF(m, V3{wI[0], wI[1], wI[2]}, o);
    }
}

// Batched BxDF evaluation of one material at one outgoing direction over
// n incoming directions: out [n][4] = f * cos (the reflectance EvaluateBxDF
// returns), pdf [n] (may be NULL).
void rc_bsdf_eval_many(const MRayMaterial* M, const float* tbn, const float* geoN, int backSide,
                       const float* wO, const float* wI, uint32_t n, float* out, float* pdf)
{
    for(uint32_t i = 0; i < n; i++)
    {
        rc_bsdf_probe(M, tbn, geoN, backSide, wO, wI + 3 * size_t(i), nullptr, 1, out + 4 * size_t(i));
        if(pdf) rc_bsdf_probe(M, tbn, geoN, backSide, wO, wI + 3 * size_t(i), nullptr, 2, pdf + i);
    }
}

// The shared elementary functions (fmath.h) at n points, for their accuracy
// tests: fn 0 sin, 1 cos, 2 exp, 3 log, 4 atan2(x, y), 5 acos, 6 asin,
// 7 atanh, 8 cosh.
void rc_fmath(int fn, const float* x, const float* y, uint32_t n, float* out)
{
    for(uint32_t i = 0; i < n; i++)
    {
        float a = x[i];
        switch(fn)
        {
            case 0: out[i] = fm::Sin(a); break;
            case 1: out[i] = fm::Cos(a); break;
            case 2: out[i] = fm::Exp(a); break;
            case 3: out[i] = fm::Log(a); break;
            case 4: out[i] = fm::Atan2(a, y[i]); break;
            case 5: out[i] = fm::Acos(a); break;
            case 6: out[i] = fm::Asin(a); break;
            case 7: out[i] = fm::Atanh(a); break;
            default: out[i] = fm::Cosh(a); break;
        }
    }
}

// InvErf (Filters.h Gaussian sampling) at n points.
void rc_inv_erf(const float* x, uint32_t n, float* out)
{
    for(uint32_t i = 0; i < n; i++) out[i] = InvErf(x[i]);
}

// SampleGaussian(xi, sigma) at n points: value [n], pdf [n].
void rc_sample_gaussian(const float* xi, uint32_t n, float sigma, float* value, float* pdf)
{
    for(uint32_t i = 0; i < n; i++)
    {
        Sample1 s = SampleGaussian(xi[i], sigma);
        value[i] = s.value;
        pdf[i] = s.pdf;
    }
}

// Latitude-longitude map: uv [n][2] <-> Z-up unit directions [n][3];
// sinIncl [n] (may be NULL) is the sine of the inclination the pdf divides by.
void rc_spherical_to_dir(const float* uv, uint32_t n, float* dirs, float* sinIncl)
{
    for(uint32_t i = 0; i < n; i++)
    {
        float s = 0.0f;
        V3 d = SphericalToDir(uv[2 * i], uv[2 * i + 1], s);
        dirs[3 * i] = d.x; dirs[3 * i + 1] = d.y; dirs[3 * i + 2] = d.z;
        if(sinIncl) sinIncl[i] = s;
    }
}
void rc_dir_to_spherical(const float* dirs, uint32_t n, float* uv)
{
    for(uint32_t i = 0; i < n; i++)
        DirToSpherical(V3{dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]}, uv[2 * i], uv[2 * i + 1]);
}
// Solid-angle pdf of the Skysphere mapping for a uv pdf (SkySolidAnglePdf).
float rc_sky_solid_angle_pdf(uint32_t type, float uvPdf, float sinIncl)
{
    return SkySolidAnglePdf(type, uvPdf, sinIncl);
}

// Light-sample probe: NEE sample from a shading point (rn[3]) ->
// {point xyz, pdf, emission rgbw}; and the MIS pdf of hitting light li
// at bary (a, b) from rayPos along rayDir.
void rc_light_probe(void* h, const float* point, const float* rn, float* out)
{
    const Scene& s = *static_cast<Scene*>(h);
    SpecCtx sc;
    LightSampleOut ls = SampleLight(s, rn, V3{point[0], point[1], point[2]}, sc);
    out[0] = ls.position.x; out[1] = ls.position.y; out[2] = ls.position.z; out[3] = ls.pdf;
    out[4] = ls.emission.x; out[5] = ls.emission.y; out[6] = ls.emission.z; out[7] = ls.emission.w;
}
float rc_light_pdf(void* h, uint32_t li, float a, float b, const float* rayPos, const float* rayDir)
{
    const Scene& s = *static_cast<Scene*>(h);
    if(li >= s.lightTri.size()) return PdfLightBoundary(s, V3{rayDir[0], rayDir[1], rayDir[2]});
    return PdfLightTri(s, li, a, b, V3{rayPos[0], rayPos[1], rayPos[2]},
                       V3{rayDir[0], rayDir[1], rayDir[2]});
}

// The power-sampling tables of a scene's lights.
void rc_light_power_table(void* h, float* prob, uint32_t* alias, float* pmf)
{
    const Scene& s = *static_cast<Scene*>(h);
    std::memcpy(prob, s.lightProb.data(), s.lightProb.size() * 4);
    std::memcpy(alias, s.lightAlias.data(), s.lightAlias.size() * 4);
    std::memcpy(pmf, s.lightPmf.data(), s.lightPmf.size() * 4);
}

// Textured Skysphere of a scene and probes of its parts.
int rc_scene_set_environment(void* h, uint32_t W, uint32_t H, const float* rgb)
{
    Scene& s = *static_cast<Scene*>(h);
    if(!IsSkysphere(s.boundaryType)) return -1;
    return BuildEnv(s, W, H, rgb);
}
void rc_env_cdfs(void* h, float* cdfX, float* cdfY)
{
    const Scene& s = *static_cast<Scene*>(h);
    std::memcpy(cdfX, s.envCdfX.data(), s.envCdfX.size() * 4);
    std::memcpy(cdfY, s.envCdfY.data(), s.envCdfY.size() * 4);
}
// xi [n][2] -> out [n][3] = (u, v, pdf over uv)
void rc_env_sample(void* h, const float* xi, uint32_t n, float* out)
{
    const Scene& s = *static_cast<Scene*>(h);
    for(uint32_t i = 0; i < n; i++)
        EnvSampleUV(s, xi[2 * i], xi[2 * i + 1], out[3 * i], out[3 * i + 1], out[3 * i + 2]);
}
// uv [n][2] -> pdf over uv [n]
void rc_env_pdf_uv(void* h, const float* uv, uint32_t n, float* out)
{
    const Scene& s = *static_cast<Scene*>(h);
    for(uint32_t i = 0; i < n; i++) out[i] = EnvPdfUV(s, uv[2 * i], uv[2 * i + 1]);
}
// Y-up directions [n][3] -> solid-angle pdf [n] and radiance [n][3]
void rc_env_eval(void* h, const float* dirs, uint32_t n, float* pdf, float* rgb)
{
    const Scene& s = *static_cast<Scene*>(h);
    for(uint32_t i = 0; i < n; i++)
    {
        V3 d{dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]};
        pdf[i] = SkyspherePdf(s, d);
        V3 r = SkysphereRadiance(s, Normalize(d));
        rgb[3 * i] = r.x; rgb[3 * i + 1] = r.y; rgb[3 * i + 2] = r.z;
    }
}
// concentric octahedral map: uv [n][2] <-> Z-up unit directions [n][3]
void rc_coocta_to_dir(const float* uv, uint32_t n, float* dirs)
{
    for(uint32_t i = 0; i < n; i++)
    {
        V3 d = CoOctaToDir(uv[2 * i], uv[2 * i + 1]);
        dirs[3 * i] = d.x; dirs[3 * i + 1] = d.y; dirs[3 * i + 2] = d.z;
    }
}
void rc_dir_to_coocta(const float* dirs, uint32_t n, float* uv)
{
    for(uint32_t i = 0; i < n; i++)
        DirToCoOcta(V3{dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]}, uv[2 * i], uv[2 * i + 1]);
}

} // extern "C"
