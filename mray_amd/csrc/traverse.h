// Closest-hit / any-hit traversal of the 4-wide BVH on gfx950.
//
// One lane per ray.  The traversal stack keeps its first kLdsStack entries in
// LDS (thread-strided, conflict-free) and spills deeper entries to a global
// per-lane area, so any tree depth is handled without scratch.
// Triangle test: Ray::IntersectsTriangle (Core/Ray.hpp:121-167), same
// operation order and explicit FMAs as the reference.
#pragma once

#include "common.h"
#include "lbvh.h"

namespace mray
{

static constexpr uint32_t kTraceTPB = 256;
static constexpr uint32_t kLdsStack = 16;

struct TraceCtx
{
    const BNode4* __restrict__ nodes;
    const SoupTri* __restrict__ soup;
    uint32_t* __restrict__ gstack;   // [gstackDepth][gridThreads]
    uint32_t gstackDepth;            // entries beyond kLdsStack per lane
    uint32_t gridThreads;
    uint32_t maxVisits;              // node count: no correct traversal visits more
};

struct HitRec
{
    uint32_t prim;
    float t, a, b;
};

// Moller-Trumbore with the reference's thresholds; writes (t, a=1-u-v, b=u).
__device__ __forceinline__ bool IntersectTri(V3 pos, V3 dir, V3 t0, V3 t1, V3 t2, bool cullFace,
                                             float& tOut, float& aOut, float& bOut)
{
    const float kSmallEps = 1e-7f;
    V3 e0 = t1 - t0;
    V3 e1 = t2 - t0;
    V3 p = Cross(dir, e1);
    float det = Dot(e0, p);
    bool backfaceHit = (det < kSmallEps);
    bool nearlyParallel = (fabsf(det) < kSmallEps);
    if((cullFace && backfaceHit) || nearlyParallel) return false;
    float invDet = 1.0f / det;
    V3 tVec = pos - t0;
    float u = Dot(tVec, p) * invDet;
    if(u < 0.0f || u > 1.0f) return false;
    V3 qVec = Cross(tVec, e0);
    float v = Dot(dir, qVec) * invDet;
    if((v < 0.0f) || (v + u) > 1.0f) return false;
    float t = Dot(e1, qVec) * invDet;
    if(t <= kSmallEps) return false;
    tOut = t;
    aOut = 1.0f - u - v;
    bOut = u;
    return true;
}

struct RayPre
{
    V3 o, d, inv;
    // float4 offsets (0 or 1) of the near plane per axis inside BNode4:
    // lo when the direction is positive, hi when negative (or -0).
    uint32_t sx, sy, sz;
};

__device__ __forceinline__ RayPre MakeRayPre(V3 o, V3 d)
{
    RayPre r;
    r.o = o;
    r.d = d;
    // exact IEEE reciprocals; a zero component gives +-inf
    r.inv = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    r.sx = signbit(r.inv.x) ? 1u : 0u;
    r.sy = signbit(r.inv.y) ? 1u : 0u;
    r.sz = signbit(r.inv.z) ? 1u : 0u;
    return r;
}

// Conservative slab test of the four child boxes (Ize 2013, "Robust BVH Ray
// Traversal"): near/far planes picked by direction sign, t = (plane - o) * inv
// with one rounding per operation, and the far bound scaled by 1 + 2*gamma(4).
// A ray parallel to an axis whose origin lies exactly on a slab plane makes
// (plane - o) * inv = 0 * inf = NaN; fmaxf/fminf drop the NaN operand, so
// that plane leaves the interval unbounded -- the ray lies in the box face and
// must reach its triangles.  tNear[k] = +inf on a miss.
__device__ __forceinline__ void SlabTest4(const BNode4* __restrict__ np, const RayPre& r,
                                          float tMin, float tMax, float tNear[4], uint4& child)
{
    const float4* q = reinterpret_cast<const float4*>(np);
    float4 nx = q[0 + r.sx], fx = q[1 - r.sx];
    float4 ny = q[2 + r.sy], fy = q[3 - r.sy];
    float4 nz = q[4 + r.sz], fz = q[5 - r.sz];
    child = reinterpret_cast<const uint4*>(np)[6];
    const float grow = 1.0f + 2.0f * 5.96e-8f * 4.0f;
    float nX[4] = {nx.x, nx.y, nx.z, nx.w}, fX[4] = {fx.x, fx.y, fx.z, fx.w};
    float nY[4] = {ny.x, ny.y, ny.z, ny.w}, fY[4] = {fy.x, fy.y, fy.z, fy.w};
    float nZ[4] = {nz.x, nz.y, nz.z, nz.w}, fZ[4] = {fz.x, fz.y, fz.z, fz.w};
    #pragma unroll
    for(int k = 0; k < 4; k++)
    {
        float x0 = (nX[k] - r.o.x) * r.inv.x, x1 = (fX[k] - r.o.x) * r.inv.x;
        float y0 = (nY[k] - r.o.y) * r.inv.y, y1 = (fY[k] - r.o.y) * r.inv.y;
        float z0 = (nZ[k] - r.o.z) * r.inv.z, z1 = (fZ[k] - r.o.z) * r.inv.z;
        float tn = MaxF(MaxF(tMin, x0), MaxF(y0, z0));
        float tf = MinF(MinF(tMax, x1), MinF(y1, z1)) * grow;
        // empty slots store lo = +inf, hi = -inf: their near bound is +inf
        // and far bound -inf for every direction sign, so they always miss
        tNear[k] = (tn <= tf) ? tn : INFINITY;
    }
}

template<bool ANY_HIT>
__device__ __forceinline__ bool TestLeafRange(const SoupTri* __restrict__ soup, uint32_t leaf,
                                              const RayPre& r, float tMin, float tMax,
                                              HitRec& best, uint32_t* triTests)
{
    uint32_t first = leaf & kLeafStartMask;
    uint32_t cnt = ((leaf & ~kLeafFlag) >> kLeafCountShift) + 1u;
    for(uint32_t k = 0; k < cnt; k++)
    {
        const float4* tp = reinterpret_cast<const float4*>(soup + first + k);
        float4 a = tp[0], b = tp[1], c = tp[2];
        uint32_t prim = __float_as_uint(a.w);
        bool cull = __float_as_uint(b.w) != 0u;
        if(triTests) (*triTests)++;
        float t, ba, bb;
        if(!IntersectTri(r.o, r.d, V3{a.x, a.y, a.z}, V3{b.x, b.y, b.z}, V3{c.x, c.y, c.z}, cull,
                         t, ba, bb))
            continue;
        // IntersectionCheck IsInRange (AcceleratorLBVH.hpp:233-236)
        if(!(t >= tMin && t < tMax)) continue;
        if(ANY_HIT) return true;
        if(t < best.t || (t == best.t && prim < best.prim))
        {
            best.prim = prim;
            best.t = t;
            best.a = ba;
            best.b = bb;
        }
    }
    return false;
}

// Closest hit; ties in t go to the smaller triangle index.
// `lstack` = this block's LDS stack base; `gtid` = grid-wide lane id.
template<bool ANY_HIT, bool COUNT>
__device__ __forceinline__ HitRec Trace(const TraceCtx& ctx, uint32_t* lstack, uint32_t gtid,
                                        V3 o, V3 d, float tMin, float tMax,
                                        uint32_t* nodeVisits = nullptr,
                                        uint32_t* triTests = nullptr)
{
    HitRec best{kInvalid, INFINITY, 0.0f, 0.0f};
    RayPre r = MakeRayPre(o, d);
    const uint32_t tid = threadIdx.x;
    const uint32_t B = blockDim.x;
    uint32_t sp = 0;
    uint32_t node = 0;
    const uint32_t spMax = kLdsStack + ctx.gstackDepth;
    for(uint32_t visits = 0; visits < ctx.maxVisits; visits++)
    {
        if(COUNT) (*nodeVisits)++;
        float curMax = (best.prim == kInvalid) ? tMax : best.t;
        float tn[4];
        uint4 ch;
        SlabTest4(ctx.nodes + node, r, tMin, curMax, tn, ch);
        uint32_t c[4] = {ch.x, ch.y, ch.z, ch.w};
        // sort (tNear, child) ascending: 5-comparator network
        auto cs = [&](int i, int j)
        {
            if(tn[j] < tn[i])
            {
                float tt = tn[i]; tn[i] = tn[j]; tn[j] = tt;
                uint32_t cc = c[i]; c[i] = c[j]; c[j] = cc;
            }
        };
        if(!ANY_HIT)
        {
            cs(0, 1); cs(2, 3); cs(0, 2); cs(1, 3); cs(1, 2);
        }
        uint32_t next = kInvalid;
        uint32_t inner[4];
        int innerCount = 0;
        #pragma unroll
        for(int k = 0; k < 4; k++)
        {
            if(tn[k] == INFINITY) continue;
            if(c[k] & kLeafFlag)
            {
                float lim = (best.prim == kInvalid) ? tMax : best.t;
                if(!ANY_HIT && tn[k] > lim * 1.0000005f) continue;
                bool h = TestLeafRange<ANY_HIT>(ctx.soup, c[k], r, tMin, tMax, best,
                                                COUNT ? triTests : nullptr);
                if(ANY_HIT && h)
                {
                    best.prim = 0;
                    return best;
                }
            }
            else inner[innerCount++] = c[k];
        }
        if(innerCount > 0)
        {
            next = inner[0];
            for(int k = innerCount - 1; k >= 1; k--)
            {
                uint32_t v = inner[k];
                if(sp >= spMax) break;   // unreachable: sized 3 per level
                if(sp < kLdsStack) lstack[sp * B + tid] = v;
                else ctx.gstack[size_t(sp - kLdsStack) * ctx.gridThreads + gtid] = v;
                sp++;
            }
        }
        if(next == kInvalid)
        {
            if(sp == 0) break;
            sp--;
            next = (sp < kLdsStack) ? lstack[sp * B + tid]
                                    : ctx.gstack[size_t(sp - kLdsStack) * ctx.gridThreads + gtid];
        }
        node = next;
    }
    if(ANY_HIT) best.prim = kInvalid;
    return best;
}

} // namespace mray
