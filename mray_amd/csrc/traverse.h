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
    V3 o, d, inv, oi;
};

__device__ __forceinline__ RayPre MakeRayPre(V3 o, V3 d)
{
    RayPre r;
    r.o = o;
    r.d = d;
    const float tiny = 1e-20f;
    float dx = (fabsf(d.x) > tiny) ? d.x : copysignf(tiny, d.x);
    float dy = (fabsf(d.y) > tiny) ? d.y : copysignf(tiny, d.y);
    float dz = (fabsf(d.z) > tiny) ? d.z : copysignf(tiny, d.z);
    r.inv = {1.0f / dx, 1.0f / dy, 1.0f / dz};
    r.oi = {o.x * r.inv.x, o.y * r.inv.y, o.z * r.inv.z};
    return r;
}

// Conservative slab test of the four child boxes.  tNear[k] = +inf on a miss.
__device__ __forceinline__ void SlabTest4(const BNode4* __restrict__ np, const RayPre& r,
                                          float tMin, float tMax, float tNear[4], uint4& child)
{
    const float4* q = reinterpret_cast<const float4*>(np);
    float4 lox = q[0], hix = q[1], loy = q[2], hiy = q[3], loz = q[4], hiz = q[5];
    child = reinterpret_cast<const uint4*>(np)[6];
    const float grow = 1.0f + 2.0f * 5.96e-8f * 4.0f;
    float lx[4] = {lox.x, lox.y, lox.z, lox.w}, hx[4] = {hix.x, hix.y, hix.z, hix.w};
    float ly[4] = {loy.x, loy.y, loy.z, loy.w}, hy[4] = {hiy.x, hiy.y, hiy.z, hiy.w};
    float lz[4] = {loz.x, loz.y, loz.z, loz.w}, hz[4] = {hiz.x, hiz.y, hiz.z, hiz.w};
    #pragma unroll
    for(int k = 0; k < 4; k++)
    {
        float x0 = Fma(lx[k], r.inv.x, -r.oi.x), x1 = Fma(hx[k], r.inv.x, -r.oi.x);
        float y0 = Fma(ly[k], r.inv.y, -r.oi.y), y1 = Fma(hy[k], r.inv.y, -r.oi.y);
        float z0 = Fma(lz[k], r.inv.z, -r.oi.z), z1 = Fma(hz[k], r.inv.z, -r.oi.z);
        float tn = MaxF(MaxF(MinF(x0, x1), MinF(y0, y1)), MaxF(MinF(z0, z1), tMin));
        float tf = MinF(MinF(MaxF(x0, x1), MaxF(y0, y1)), MinF(MaxF(z0, z1), tMax * grow)) * grow;
        // empty slots have lo > hi, so tn > tf for any ray
        tNear[k] = (tn <= tf && lx[k] <= hx[k]) ? tn : INFINITY;
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
