// Closest-hit / any-hit traversal of the wide BVH on gfx950: 4-wide nodes
// with fp32 boxes (BNode4) or compressed 8-wide nodes (CNode8).
//
// One lane per ray.  The traversal stack holds (node, tNear) pairs: its first
// kLdsStack entries live in LDS (rows of 64 per wave, conflict-free), deeper ones
// spill to a global per-lane area, so any tree depth is handled without
// scratch.  A popped entry whose tNear is beyond the closest hit found so far
// is dropped without fetching the node.
// Triangle test: Ray::IntersectsTriangle (Core/Ray.hpp:121-167), same
// operation order and explicit FMAs as the reference.
#pragma once

#include "common.h"
#include "lbvh.h"

namespace mray
{

static constexpr uint32_t kTraceTPB = 256;
static constexpr uint32_t kLdsStack = 12;

// One instance of a mesh in a two-level scene: a TRS transform (rotation
// quaternion, per-axis scale, translation) object -> world, its inverse as
// three rows, the mesh tree's root and the instance's first light in the
// world light table.
struct InstGPU
{
    float4 rot;          // (w, x, y, z)
    float4 transScale;   // translation xyz, w = 1 if the scale is uniform else 0
    float4 scale;        // per-axis scale xyz
    float4 inv[3];       // world -> object, rows (3x3 | translation)
    uint32_t root;
    uint32_t lightOff;
    uint32_t mesh;
    uint32_t flags;       // MRAY_INST_ROTATE_FRAMES
};

struct TraceCtx
{
    const BNode4* __restrict__ nodes;   // width 4
    const CNode8* __restrict__ cnodes;  // width 8
    const SoupTri* __restrict__ soup;
    uint2* __restrict__ gstack;      // [gstackDepth][gridThreads] (node, tNear bits)
    uint32_t gstackDepth;            // entries beyond kLdsStack per lane
    uint32_t gridThreads;
    uint32_t maxVisits;              // node count: no correct traversal visits more
    uint32_t width;                  // 4 or 8
    // two-level scenes only: instance tree over world boxes, its leaf
    // positions' instances, the instance table
    const BNode4* __restrict__ tlas;
    const uint32_t* __restrict__ tlasOrder;
    const InstGPU* __restrict__ inst;
    uint32_t tlasVisits;
    uint32_t instWalkVisits;         // both levels together: no correct walk visits more
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

// Compressed 8-wide node: decode each child box exactly as the builder
// verified it (fmaf(q, 2^e, origin) contains the true box), then the same
// robust slab test as SlabTest4.
__device__ __forceinline__ float Pow2(int e) { return __uint_as_float(uint32_t(e + 127) << 23); }
__device__ __forceinline__ float QByte(uint32_t lo4, uint32_t hi4, int k)
{
    uint32_t w = (k < 4) ? lo4 : hi4;
    return float((w >> (8 * (k & 3))) & 0xFFu);
}

__device__ __forceinline__ void SlabTest8(const CNode8* __restrict__ np, const RayPre& r,
                                          float tMin, float tMax, float tNear[8], uint32_t child[8])
{
    const uint4* q = reinterpret_cast<const uint4*>(np);
    uint4 h = q[0], bx = q[1], by = q[2], bz = q[3], c0 = q[4], c1 = q[5];
    const float ox = __uint_as_float(h.x), oy = __uint_as_float(h.y), oz = __uint_as_float(h.z);
    const float sx = Pow2(int(int8_t(h.w & 0xFFu)));
    const float sy = Pow2(int(int8_t((h.w >> 8) & 0xFFu)));
    const float sz = Pow2(int(int8_t((h.w >> 16) & 0xFFu)));
    child[0] = c0.x; child[1] = c0.y; child[2] = c0.z; child[3] = c0.w;
    child[4] = c1.x; child[5] = c1.y; child[6] = c1.z; child[7] = c1.w;
    const float grow = 1.0f + 2.0f * 5.96e-8f * 4.0f;
    #pragma unroll
    for(int k = 0; k < 8; k++)
    {
        float lx = __builtin_fmaf(QByte(bx.x, bx.y, k), sx, ox);
        float hx = __builtin_fmaf(QByte(bx.z, bx.w, k), sx, ox);
        float ly = __builtin_fmaf(QByte(by.x, by.y, k), sy, oy);
        float hy = __builtin_fmaf(QByte(by.z, by.w, k), sy, oy);
        float lz = __builtin_fmaf(QByte(bz.x, bz.y, k), sz, oz);
        float hz = __builtin_fmaf(QByte(bz.z, bz.w, k), sz, oz);
        float nx = r.sx ? hx : lx, fx = r.sx ? lx : hx;
        float ny = r.sy ? hy : ly, fy = r.sy ? ly : hy;
        float nz = r.sz ? hz : lz, fz = r.sz ? lz : hz;
        float x0 = (nx - r.o.x) * r.inv.x, x1 = (fx - r.o.x) * r.inv.x;
        float y0 = (ny - r.o.y) * r.inv.y, y1 = (fy - r.o.y) * r.inv.y;
        float z0 = (nz - r.o.z) * r.inv.z, z1 = (fz - r.o.z) * r.inv.z;
        float tn = MaxF(MaxF(tMin, x0), MaxF(y0, z0));
        float tf = MinF(MinF(tMax, x1), MinF(y1, z1)) * grow;
        // an empty slot holds kInvalid; its 255/0 bytes do not decode to an
        // empty box when the node is flat on an axis (exponent -126 makes
        // 255 steps vanish against the origin), so the child word decides
        bool valid = child[k] != kInvalid;
        tNear[k] = (valid && tn <= tf) ? tn : INFINITY;
    }
}

// LDS stack layout: each wave owns kLdsStack rows of 64 entries, entry sp
// of a lane at sp * 64 + StackLane() (a shift, no multiply by the block
// size; a row is 64 consecutive 8-byte entries, conflict-free).
__device__ __forceinline__ uint32_t StackLane()
{
    return (threadIdx.x >> 6) * (kLdsStack * 64u) + (threadIdx.x & 63u);
}

// One traversal-stack entry: the first kLdsStack levels in LDS, deeper ones
// in the lane's global spill area.
// While no lane of the wave is past the LDS levels (nearly always) the wave
// takes a branch-free LDS path, chosen by one wave-uniform test; otherwise
// each lane picks its memory.  The empty asm statement (a compiler barrier;
// it emits nothing) keeps that slow path's LDS and global accesses apart:
// merged, they would become one access through a generic pointer.
__device__ __forceinline__ void StackPut(const TraceCtx& ctx, uint2* lstack, uint32_t sp,
                                         uint32_t lb, uint32_t gtid, uint2 v)
{
    if(__builtin_expect(__ballot(sp >= kLdsStack) == 0ull, 1)) lstack[sp * 64u + lb] = v;
    else if(sp < kLdsStack) lstack[sp * 64u + lb] = v;
    else
    {
        ctx.gstack[size_t(sp - kLdsStack) * ctx.gridThreads + gtid] = v;
        asm volatile("" ::: "memory");
    }
}

__device__ __forceinline__ uint2 StackGet(const TraceCtx& ctx, const uint2* lstack, uint32_t sp,
                                          uint32_t lb, uint32_t gtid)
{
    uint2 v;
    if(__builtin_expect(__ballot(sp >= kLdsStack) == 0ull, 1)) v = lstack[sp * 64u + lb];
    else if(sp < kLdsStack) v = lstack[sp * 64u + lb];
    else
    {
        v = ctx.gstack[size_t(sp - kLdsStack) * ctx.gridThreads + gtid];
        asm volatile("" ::: "memory");
    }
    return v;
}

// A node's stack candidates (bit k of cand: child k, in ascending tNear
// order), farthest first: each pushes the one found before it, so the
// nearest is left in next and the stack pops near-first.  While every lane
// of the wave has room for W-1 more entries in LDS (one wave-uniform test)
// the pushes are branch-free: each candidate writes the free slot and the
// stack grows only on a real push (a write above the top is dead).  The
// 8-wide walk and the two-level walk use the guarded pushes only (their
// branch-free form spilled SGPRs, and measured slower, respectively).
template<int W, bool BRANCH_FREE = true>
__device__ __forceinline__ void PushFarFirst(const TraceCtx& ctx, uint2* lstack, uint32_t& sp,
                                             uint32_t spMax, uint32_t lb, uint32_t gtid, const float (&tn)[W],
                                             const uint32_t (&c)[W], uint32_t cand,
                                             uint32_t& next, float& nextT)
{
    next = kInvalid;
    nextT = 0.0f;
    if(BRANCH_FREE && __builtin_expect(__ballot(sp + uint32_t(W - 1) > kLdsStack) == 0ull, 1))
    {
        #pragma unroll
        for(int k = W - 1; k >= 0; k--)
        {
            const bool ck = (cand >> k) & 1u;
            if(k < W - 1)
            {
                lstack[sp * 64u + lb] = make_uint2(next, __float_as_uint(nextT));
                sp += (ck && next != kInvalid) ? 1u : 0u;
            }
            next = ck ? c[k] : next;
            nextT = ck ? tn[k] : nextT;
        }
    }
    else
    {
        #pragma unroll
        for(int k = W - 1; k >= 0; k--)
        {
            if(!((cand >> k) & 1u)) continue;
            if(next != kInvalid && sp < spMax)   // sp < spMax: sized for every level
            {
                StackPut(ctx, lstack, sp, lb, gtid, make_uint2(next, __float_as_uint(nextT)));
                sp++;
            }
            next = c[k];
            nextT = tn[k];
        }
    }
}

// Children in ascending tNear (optimal networks: 5 / 19 comparators), each
// compare-exchange as selects (no exec-mask branch per comparator).  Any-hit
// walks near-first as well: occluders of a shadow ray cluster near its
// origin, and the order cut the shadow pass by about 30 %.
template<int W>
__device__ __forceinline__ void SortNear(float (&tn)[W], uint32_t (&c)[W])
{
    auto cs = [&](int i, int j)
    {
        const bool s = tn[j] < tn[i];
        const float a = s ? tn[j] : tn[i], b = s ? tn[i] : tn[j];
        const uint32_t ca = s ? c[j] : c[i], cb = s ? c[i] : c[j];
        tn[i] = a; tn[j] = b; c[i] = ca; c[j] = cb;
    };
    if constexpr(W == 8)
    {
        cs(0, 2); cs(1, 3); cs(4, 6); cs(5, 7);
        cs(0, 4); cs(1, 5); cs(2, 6); cs(3, 7);
        cs(0, 1); cs(2, 3); cs(4, 5); cs(6, 7);
        cs(2, 4); cs(3, 5);
        cs(1, 4); cs(3, 6);
        cs(1, 2); cs(3, 4); cs(5, 6);
    }
    else
    {
        cs(0, 1); cs(2, 3); cs(0, 2); cs(1, 3); cs(1, 2);
    }
}

template<bool ANY_HIT>
__device__ __forceinline__ bool TestLeafRange(const SoupTri* __restrict__ soup, uint32_t leaf,
                                              const RayPre& r, float tMin, float tMax,
                                              HitRec& best, uint32_t* triTests)
{
    uint32_t first = leaf & kLeafStartMask;
    // every tree is collapsed with kLeafMax triangles per leaf at most
    uint32_t cnt = (kLeafMax == 1u) ? 1u : ((leaf & ~kLeafFlag) >> kLeafCountShift) + 1u;
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

// One ray's walk of a wide tree, kept between steps: the lanes of a kernel
// that hands a lane a new ray as soon as its walk ends (KTraceRefill) keep
// this state per lane.  WalkStep visits one node exactly as an iteration of
// TraceW's loop does, so a walk finds TraceW's hit.
struct Walk
{
    RayPre r;
    float tMin, tMax;
    HitRec best;
    uint32_t node;       // the node the next step visits
    uint32_t sp;         // stack height
};

__device__ __forceinline__ void WalkInit(Walk& w, V3 o, V3 d, float tMin, float tMax, uint32_t root,
                                         uint32_t sp0)
{
    w.r = MakeRayPre(o, d);
    w.tMin = tMin;
    w.tMax = tMax;
    w.best = HitRec{kInvalid, INFINITY, 0.0f, 0.0f};
    w.node = root;
    w.sp = sp0;
}

// Visit w.node: test its children, test leaf triangles, push the farther
// inner children, pop the next node (dropping subtrees beyond the closest
// hit).  Returns true when the walk is over: the stack is empty, or (any
// hit) a triangle was hit, which sets w.best.prim to 0.
// `lstack` = this block's LDS stack base; `gtid` = grid-wide lane id; `sp0`
// = the stack's floor (entries below belong to an enclosing walk).
template<int W, bool ANY_HIT, bool COUNT>
__device__ __forceinline__ bool WalkStep(const TraceCtx& ctx, uint2* lstack, uint32_t gtid, Walk& w,
                                         uint32_t sp0, uint32_t* nodeVisits, uint32_t* triTests)
{
    const uint32_t lb = StackLane();
    const uint32_t spMax = kLdsStack + ctx.gstackDepth;
    const float kGrow = 1.0000005f;
    if(COUNT) (*nodeVisits)++;
    HitRec& best = w.best;
    float curMax = (best.prim == kInvalid) ? w.tMax : best.t;
    float tn[W];
    uint32_t c[W];
    if constexpr(W == 8) SlabTest8(ctx.cnodes + w.node, w.r, w.tMin, curMax, tn, c);
    else
    {
        uint4 ch;
        SlabTest4(ctx.nodes + w.node, w.r, w.tMin, curMax, tn, ch);
        c[0] = ch.x; c[1] = ch.y; c[2] = ch.z; c[3] = ch.w;
    }
    SortNear<W>(tn, c);
    // as in TraceW: inner children farthest first, each pushing the one
    // found before it; then the leaves nearest first
    uint32_t next = kInvalid;
    float nextT = 0.0f;
    uint32_t sp = w.sp;
    #pragma unroll
    for(int k = W - 1; k >= 0; k--)
    {
        if(tn[k] == INFINITY || (c[k] & kLeafFlag)) continue;
        if(next != kInvalid && sp < spMax)   // sp < spMax: sized (W-1) per level
        {
            const uint2 v = make_uint2(next, __float_as_uint(nextT));
            StackPut(ctx, lstack, sp, lb, gtid, v);
            sp++;
        }
        next = c[k];
        nextT = tn[k];
    }
    #pragma unroll
    for(int k = 0; k < W; k++)
    {
        if(tn[k] == INFINITY || !(c[k] & kLeafFlag)) continue;
        float lim = (best.prim == kInvalid) ? w.tMax : best.t;
        if(!ANY_HIT && tn[k] > lim * kGrow) continue;
        bool h = TestLeafRange<ANY_HIT>(ctx.soup, c[k], w.r, w.tMin, w.tMax, best,
                                        COUNT ? triTests : nullptr);
        if(ANY_HIT && h)
        {
            best.prim = 0;
            w.sp = sp;
            return true;
        }
    }
    while(next == kInvalid && sp > sp0)
    {
        sp--;
        uint2 v;
        v = StackGet(ctx, lstack, sp, lb, gtid);
        // drop subtrees that start beyond the closest hit so far
        if(!ANY_HIT && best.prim != kInvalid && __uint_as_float(v.y) > best.t * kGrow) continue;
        next = v.x;
    }
    w.sp = sp;
    w.node = next;
    return next == kInvalid;
}

// Closest hit; ties in t go to the smaller triangle index.
// `lstack` = this block's LDS stack base; `gtid` = grid-wide lane id;
// `root` = the tree's root node (0, or a scene's first node in a batch);
// `sp0` = the stack's floor (entries below belong to an enclosing walk).
template<int W, bool ANY_HIT, bool COUNT>
__device__ __forceinline__ HitRec TraceW(const TraceCtx& ctx, uint2* lstack, uint32_t gtid,
                                         V3 o, V3 d, float tMin, float tMax, uint32_t root,
                                         uint32_t* nodeVisits = nullptr,
                                         uint32_t* triTests = nullptr, uint32_t sp0 = 0)
{
    HitRec best{kInvalid, INFINITY, 0.0f, 0.0f};
    RayPre r = MakeRayPre(o, d);
    uint32_t sp = sp0;
    uint32_t node = root;
    const uint32_t lb = StackLane();
    const uint32_t spMax = kLdsStack + ctx.gstackDepth;
    const float kGrow = 1.0000005f;
    for(uint32_t visits = 0; visits < ctx.maxVisits; visits++)
    {
        if(COUNT) (*nodeVisits)++;
        float curMax = (best.prim == kInvalid) ? tMax : best.t;
        float tn[W];
        uint32_t c[W];
        if constexpr(W == 8) SlabTest8(ctx.cnodes + node, r, tMin, curMax, tn, c);
        else
        {
            uint4 ch;
            SlabTest4(ctx.nodes + node, r, tMin, curMax, tn, ch);
            c[0] = ch.x; c[1] = ch.y; c[2] = ch.z; c[3] = ch.w;
        }
        SortNear<W>(tn, c);
        // hit children: leaves are tested below, inner ones go on the stack
        uint32_t leafMask = 0, cand = 0;
        #pragma unroll
        for(int k = 0; k < W; k++)
        {
            const bool hit = tn[k] != INFINITY, leaf = (c[k] & kLeafFlag) != 0u;
            leafMask |= (hit && leaf) ? (1u << k) : 0u;
            cand |= (hit && !leaf) ? (1u << k) : 0u;
        }
        uint32_t next;
        float nextT;
        PushFarFirst<W, W == 4>(ctx, lstack, sp, spMax, lb, gtid, tn, c, cand, next, nextT);
        // the node's leaves, nearest first: a wave runs as many triangle
        // tests as its lane with the most leaves here, not one per child
        // position that any lane has a leaf in
        while(leafMask)
        {
            const int k = __builtin_ctz(leafMask);
            leafMask &= leafMask - 1u;
            uint32_t ck = c[0];
            float tk = tn[0];
            #pragma unroll
            for(int j = 1; j < W; j++)
            {
                ck = (k == j) ? c[j] : ck;
                tk = (k == j) ? tn[j] : tk;
            }
            const float lim = (best.prim == kInvalid) ? tMax : best.t;
            if(!ANY_HIT && tk > lim * kGrow) continue;
            bool h = TestLeafRange<ANY_HIT>(ctx.soup, ck, r, tMin, tMax, best,
                                            COUNT ? triTests : nullptr);
            if(ANY_HIT && h)
            {
                best.prim = 0;
                return best;
            }
        }
        while(next == kInvalid && sp > sp0)
        {
            sp--;
            uint2 v = StackGet(ctx, lstack, sp, lb, gtid);
            // drop subtrees that start beyond the closest hit so far
            if(!ANY_HIT && best.prim != kInvalid && __uint_as_float(v.y) > best.t * kGrow) continue;
            next = v.x;
        }
        if(next == kInvalid) break;
        node = next;
    }
    if(ANY_HIT) best.prim = kInvalid;
    return best;
}

// Two-level closest / any hit (the reference's instance tree then mesh tree,
// AcceleratorLBVH.cu:437-495 and AcceleratorWork.kt.h:177-244), in one loop
// over both levels: each step visits one node, of the 4-wide instance tree
// or of the mesh tree of the instance being walked, so the lanes of a wave
// keep stepping together whichever level they are in.  (A walk that calls
// the mesh walk from the instance walk leaves the lanes outside an instance
// idle while the others walk its tree; this loop renders the two-level
// workload 1.8x faster.)  An instance is a stack candidate like an inner
// node, taken near-first; entering it pushes a return mark, moves the ray
// to object space (the transform is affine, so t is unchanged) and
// continues at the mesh tree's root; popping the mark restores the
// world-space ray.  instOut names the hit's instance.  Exact ties in t
// between two instances go to the one walked first.
constexpr uint32_t kInstEnter = kLeafFlag;       // | tlasOrder position
constexpr uint32_t kInstReturn = 0xFFFFFFFEu;
template<bool ANY_HIT, bool COUNT>
__device__ __forceinline__ HitRec TraceInstFlat(const TraceCtx& ctx, uint2* lstack, uint32_t gtid,
                                                V3 o, V3 d, float tMin, float tMax, uint32_t& instOut,
                                                uint32_t* nodeVisits = nullptr,
                                                uint32_t* triTests = nullptr)
{
    HitRec best{kInvalid, INFINITY, 0.0f, 0.0f};
    instOut = kInvalid;
    const RayPre world = MakeRayPre(o, d);
    RayPre r = world;
    uint32_t inst = kInvalid;                 // instance whose mesh tree is walked
    const BNode4* tree = ctx.tlas;
    uint32_t sp = 0;
    uint32_t node = 0;
    const uint32_t lb = StackLane();
    const uint32_t spMax = kLdsStack + ctx.gstackDepth;
    const float kGrow = 1.0000005f;
    for(uint32_t visits = 0; visits < ctx.instWalkVisits; visits++)
    {
        if(COUNT) (*nodeVisits)++;
        const float curMax = (best.prim == kInvalid) ? tMax : best.t;
        float tn[4];
        uint4 ch;
        SlabTest4(tree + node, r, tMin, curMax, tn, ch);
        uint32_t c[4] = {ch.x, ch.y, ch.z, ch.w};
        SortNear<4>(tn, c);
        // stack candidates: an inner node of either tree, or an instance of
        // the instance tree (kInstEnter | its leaf position); mesh leaves
        // are tested below
        uint32_t cand = 0;
        #pragma unroll
        for(int k = 0; k < 4; k++)
        {
            const bool meshLeaf = (c[k] & kLeafFlag) && inst != kInvalid;
            cand |= (tn[k] != INFINITY && !meshLeaf) ? (1u << k) : 0u;
        }
        uint32_t next;
        float nextT;
        PushFarFirst<4, false>(ctx, lstack, sp, spMax, lb, gtid, tn, c, cand, next, nextT);
        #pragma unroll
        for(int k = 0; k < 4; k++)
        {
            if(tn[k] == INFINITY) continue;
            if((c[k] & kLeafFlag) && inst != kInvalid)
            {
                // a mesh leaf: its triangles, in object space
                const float lim = (best.prim == kInvalid) ? tMax : best.t;
                if(!ANY_HIT && tn[k] > lim * kGrow) continue;
                const uint32_t before = best.prim;
                const float beforeT = best.t;
                bool h = TestLeafRange<ANY_HIT>(ctx.soup, c[k], r, tMin, tMax, best,
                                                COUNT ? triTests : nullptr);
                if(ANY_HIT && h)
                {
                    instOut = inst;
                    best.prim = 0;
                    return best;
                }
                if(best.prim != before || best.t != beforeT) instOut = inst;
            }
        }
        while(next == kInvalid && sp > 0)
        {
            sp--;
            uint2 v;
            v = StackGet(ctx, lstack, sp, lb, gtid);
            if(v.x == kInstReturn)
            {
                // back from an instance: world ray, instance tree
                r = world;
                inst = kInvalid;
                tree = ctx.tlas;
                continue;
            }
            // drop candidates that start beyond the closest hit so far
            if(!ANY_HIT && best.prim != kInvalid && __uint_as_float(v.y) > best.t * kGrow) continue;
            next = v.x;
        }
        if(next == kInvalid) break;
        if((next & kLeafFlag) && inst == kInvalid)
        {
            // enter the instance: return mark, object-space ray, mesh root
            const uint32_t id = ctx.tlasOrder[next & kLeafStartMask];
            const InstGPU& I = ctx.inst[id];
            if(sp >= spMax) break;       // unreachable: the mark has its slot
            const uint2 mark = make_uint2(kInstReturn, 0u);
            StackPut(ctx, lstack, sp, lb, gtid, mark);
            sp++;
            const float4 a = I.inv[0], b = I.inv[1], e = I.inv[2];
            const V3 lo = {a.x * o.x + a.y * o.y + a.z * o.z + a.w,
                           b.x * o.x + b.y * o.y + b.z * o.z + b.w,
                           e.x * o.x + e.y * o.y + e.z * o.z + e.w};
            const V3 ld = {a.x * d.x + a.y * d.y + a.z * d.z,
                           b.x * d.x + b.y * d.y + b.z * d.z,
                           e.x * d.x + e.y * d.y + e.z * d.z};
            r = MakeRayPre(lo, ld);
            inst = id;
            tree = ctx.nodes;
            node = I.root;
        }
        else node = next;
    }
    if(ANY_HIT) best.prim = kInvalid;
    return best;
}

template<bool ANY_HIT, bool COUNT>
__device__ __forceinline__ HitRec Trace(const TraceCtx& ctx, uint2* lstack, uint32_t gtid,
                                        V3 o, V3 d, float tMin, float tMax,
                                        uint32_t* nodeVisits = nullptr,
                                        uint32_t* triTests = nullptr)
{
    return TraceW<4, ANY_HIT, COUNT>(ctx, lstack, gtid, o, d, tMin, tMax, 0u, nodeVisits, triTests);
}

} // namespace mray
