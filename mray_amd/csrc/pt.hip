// Wavefront path tracer on gfx950 ("PathTracerRGB" / "PathTracerSpectral").
//
// One iteration = one DoRenderWork() bounce of every path slot
// (TracerDLL/PathTracerRenderer.cu:930-1074), as six kernels that read every
// count from device memory, so iterations queue back-to-back with no host
// synchronisation:
//   KRefillCount / KRefillScan / KRefillApply  ReloadPaths: finished paths
//        go to the film with float atomics (AddRadianceToRenderBufferThroughput,
//        PathTracerRendererBase.cu:215-262); then free slots, in ascending slot
//        order, take consecutive pixel indices from the global path counter
//        (PathTracerRendererBase.cu:69-203 with the stable partition of
//        Device/CPU/AlgBinaryPartitionCPU.h:31-48).  Live slots form the
//        active list in ascending slot order; every position comes from the
//        per-block scan, so no atomics and a deterministic order.
//   KTrace    closest hit of every active slot (CastRays)
//   KShade    boundary/light work or NEE (work 1) + BxDF scatter (work 0)
//             (PathTracerRendererShaders.h:201-520); slots with a shadow ray
//             are appended to the shadow queue (one atomic per wave)
//   KShadow   any-hit visibility over the shadow queue only, then
//             KCAccumulateShadowRaysPT.
// Each slot consumes PCG32 numbers in the reference's per-slot order, so with
// the same seed every path sample equals the reference schedule's.
#include "common.h"
#include "scene.h"
#include "shading.h"
#include "spectral.h"
#include "traverse.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <chrono>
#include <random>
#include <vector>

namespace mray
{

int DevAlloc(DeviceArena& arena, void** p, size_t bytes);

enum : uint32_t { ST_ALIVE = 0, ST_DEAD = 1, ST_INVALID = 2 };
enum : uint32_t { RT_CAMERA = 0, RT_PATH = 1, RT_SPECULAR = 2, RT_SHADOW = 3 };

__device__ __forceinline__ uint32_t MDepth(uint32_t m) { return m & 0xFFFFu; }
__device__ __forceinline__ uint32_t MStatus(uint32_t m) { return (m >> 16) & 3u; }
__device__ __forceinline__ uint32_t MType(uint32_t m) { return (m >> 18) & 3u; }
__device__ __forceinline__ bool MDispersed(uint32_t m) { return (m >> 20) & 1u; }
__device__ __forceinline__ uint32_t MPack(uint32_t depth, uint32_t status, uint32_t type,
                                          bool disp, bool pending)
{
    return (depth & 0xFFFFu) | (status << 16) | (type << 18) | (uint32_t(disp) << 20) |
           (uint32_t(pending) << 21);
}

struct SlotSoA
{
    float4* rayO;      // pos, tMin
    float4* rayD;      // dir, tMax
    uint4* hit;        // prim, t, a, b
    float4* thr;
    float4* rad;
    uint32_t* meta;
    float* prevPdf;
    uint32_t* pixel;
    float* filterW;
    uint32_t* rng;
    float4* sO;        // shadow ray pos, tMin
    float4* sD;        // shadow ray dir, tMax
    float4* sRad;
    float4* waves;     // spectral only
    float4* wpdf;
    float2* spread;    // radiance cache: (a0, sum of sqrt spreads) of the path
    uint32_t* view;    // batch: viewport of the path
    uint32_t* inst;    // instanced scenes: instance of the hit
};

// Radiance-cache state of the renderer: one cache, or one per view of a
// batch renderer (queries, training rings and the scene box are per view;
// a path's view picks them).  One path slot in 2^trainShift is a training
// path: it never terminates into the cache and records every surface vertex
// (query record, throughput, radiance so far, four channels: RGB, or the
// four hero wavelengths); when it ends, each vertex yields a training pair
// (record, reflected radiance = (final - so far) / throughput, converted to
// RGB in spectral mode).  Other paths terminate into the cache once their
// spread is large (Mueller et al. 2021, sec. 4).
constexpr uint32_t kNrcRec = 11;
constexpr uint32_t kNrcVert = kNrcRec + 8;   // record, throughput, radiance so far
struct NrcRenderState
{
    uint32_t on;
    uint32_t trainShift;
    uint32_t maxVerts;
    uint32_t ringSize;        // training pairs kept per view (power of two)
    uint32_t minDepth;        // terminal vertices have depth >= minDepth
    uint32_t neeAtEnd;        // terminal vertices take their NEE; the cache the rest
    uint32_t views;           // caches (1, or the batch's view count)
    uint32_t queryCap;        // entries of one view's query queue (N)
    float spreadC;
    const float4* box;        // [views] scene box min xyz, 1 / extent
    float* trainStore;        // [N >> trainShift][maxVerts][kNrcVert]
    uint32_t* trainCount;     // [N >> trainShift]
    float* queryRec;          // [views][queryCap][11]
    uint32_t* querySlot;      // [views][queryCap]
    float* queryOut;          // [views][queryCap][3]
    unsigned long long* ringCount;  // [views] training pairs written so far
    float* ringRec;           // [views][ringSize][11]
    float* ringTgt;           // [views][ringSize][3]
};

struct Sched
{
    unsigned long long counter;   // next global path index
    unsigned long long limit;     // totalSPP * pixels
    unsigned long long base;      // counter at this iteration's refill
    unsigned long long deadTotal;
    unsigned long long extRays;
    unsigned long long shadowRays;
    unsigned long long nodeVisits;  // counting mode only
    unsigned long long triTests;    // counting mode only
    uint32_t fill;
    uint32_t activeCount;
    uint32_t iterations;
    uint32_t spare;
    unsigned long long spare2;
    unsigned long long queries;     // radiance-cache queries so far
    unsigned long long startClock;  // wall clock at the first refill
    uint32_t viewsDone;             // batch: views [0, viewsDone) are complete
    uint32_t viewBase;              // batch: view of the first path this refill issues
    unsigned long long shadowNodeVisits;  // counting mode only: any-hit walks
    unsigned long long shadowTriTests;
};

// Work and queue counters.  Atomics on one address serialise at about
// 90 M/s on this GPU, so every queue and work list is split into kShards
// shards with their own counters, 128 bytes apart.  Shard s of the active
// list is its s-th contiguous sixteenth in 64-item chunks (ShardOfChunk);
// the shard follows an item into the material queues and the shadow queue,
// so no shard of any queue holds more than ShardCap(N) items, and a shard
// stays one region of the image.
constexpr uint32_t kShards = 16;
constexpr uint32_t kCtrStride = 32;      // uint32 words between counters
constexpr uint32_t kMaxNrcViews = 32;    // radiance caches per renderer (one per batch view)
constexpr uint32_t kFetch = 256;         // items a wave takes per fetch atomic
enum : uint32_t
{
    CTR_TRACE = 0,                         // fetch, per shard of the active list
    CTR_SHADOW_FETCH = kShards,            // fetch, per shadow-queue shard
    CTR_SHADOW = 2 * kShards,              // shadow-queue shard sizes
    CTR_QFETCH = 3 * kShards,              // [5 material queues][kShards] fetch
    CTR_QCOUNT = 8 * kShards,              // [5 material queues][kShards] sizes
    CTR_NRCQ = 13 * kShards,               // [views] radiance-cache queries this bounce
    CTR_COUNT = 13 * kShards + kMaxNrcViews
};
__host__ __device__ constexpr uint32_t ShardCap(uint32_t N)
{
    return 64u * ((((N + 63u) / 64u) + kShards - 1u) / kShards);
}
// shadow-queue shards: the chunk bound plus 63 items per material queue a
// wave straddling two queue shards may add
__host__ __device__ constexpr uint32_t ShadowShardCap(uint32_t N) { return ShardCap(N) + 64u * 5u; }
__device__ __forceinline__ uint32_t* Ctr(uint32_t* ctr, uint32_t idx) { return ctr + size_t(idx) * kCtrStride; }
// shard of 64-item chunk c of an n-item list (C chunks): contiguous ranges
__device__ __forceinline__ uint32_t ShardOfChunk(uint32_t c, uint32_t n)
{
    const uint32_t C = (n + 63u) / 64u;
    return uint32_t((uint64_t(c) * kShards) / C);
}
// Dynamic work fetch, 64 items per atomic, over kXcds counters: the list
// [0, n) is cut into kXcds contiguous parts, a wave starts on the part of its
// XCD (workgroups are dispatched round-robin over the 8 XCDs, each with its
// own L2, so an XCD keeps to one region of the image) and then takes from
// the others.  A plain load first skips a part already used up; every wave
// exits after one pass over the parts.  Wave-uniform values stay scalar.
// f(k) runs for each item k < n.
constexpr uint32_t kXcds = 8;
template<class F>
__device__ __forceinline__ void FetchLoop(uint32_t* ctr, uint32_t n, F&& f)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t C = (n + 63u) / 64u;                       // 64-item chunks
    uint32_t part = blockIdx.x % kXcds;
    for(uint32_t t = 0; t < kXcds; t++)
    {
        const uint32_t c0 = __builtin_amdgcn_readfirstlane(uint32_t((uint64_t(C) * part) / kXcds));
        const uint32_t c1 = __builtin_amdgcn_readfirstlane(uint32_t((uint64_t(C) * (part + 1u)) / kXcds));
        uint32_t* fc = ctr + size_t(part) * kCtrStride;
        for(;;)
        {
            uint32_t c = c1 - c0;
            if(lane == 0)
            {
                c = __hip_atomic_load(fc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if(c < c1 - c0) c = atomicAdd(fc, 1u);
            }
            c = __builtin_amdgcn_readfirstlane(c);
            if(c >= c1 - c0) break;
            const uint32_t k = (c0 + c) * 64u + lane;
            if(k < n) f(k);
        }
        part = (part + 1u) % kXcds;
    }
}

// kShards queue shards read as one list: shard s holds items
// [s * cap, s * cap + count_s); virtual index k maps to its shard and slot.
struct Shards
{
    uint32_t pre[kShards + 1];   // exclusive prefix of the shard sizes
};
__device__ __forceinline__ void LoadShards(uint32_t* ctr, uint32_t countBase, Shards& g)
{
    uint32_t acc = 0;
    #pragma unroll
    for(uint32_t sh = 0; sh < kShards; sh++)
    {
        g.pre[sh] = acc;
        acc += __builtin_amdgcn_readfirstlane(*Ctr(ctr, countBase + sh));
    }
    g.pre[kShards] = acc;
}
__device__ __forceinline__ uint32_t ShardItem(const Shards& g, uint32_t k, uint32_t cap, uint32_t& sh)
{
    uint32_t s0 = 0, first = 0;
    #pragma unroll
    for(uint32_t t = 1; t < kShards; t++)
        if(k >= g.pre[t]) { s0 = t; first = g.pre[t]; }
    sh = s0;
    return s0 * cap + (k - first);
}

// One viewport of a batch: its camera and film, its range of the global path
// index space, and the scene it renders inside the merged scene arrays (tree
// root, light range, boundary light).
struct ViewGPU
{
    CameraGPU cam;
    float* film;                      // [4][P]
    unsigned long long pathBegin;     // global path indices [pathBegin, pathEnd)
    unsigned long long pathEnd;
    uint32_t P, W;
    uint32_t root;
    uint32_t lightOff, lightCount;
    uint32_t boundaryType;
    float boundaryRad[3];
    float sceneDiameter;
    const float4* envTex;             // textured Skysphere (envW != 0)
    const float* envCdfX;
    const float* envCdfY;
    uint32_t envW, envH;
};

// The current tile of a tiled renderer (MRAY_RENDER_TILED): the tile's
// sub-camera (CameraPinhole::GenerateSubCamera, CamerasDefault.hpp:226-247)
// at the tile's own resolution, and where the tile sits in the film.
struct TileGPU
{
    CameraGPU cam;
    uint32_t x0, y0;      // first film pixel
    uint32_t w;           // width of this tile (edge tiles may be narrower)
    uint32_t P;           // pixels of this tile
};

struct RenderConst
{
    SceneView scene;
    CameraGPU cam;
    SpectralView spec;
    uint32_t N;
    uint32_t P;
    uint32_t W;
    uint32_t rr0, rr1;
    uint32_t mode;
    uint32_t spectral;
    float* film;
    NrcRenderState nrc;
    // batch kernels only
    const InstGPU* inst;              // instanced scenes: instance table
    const ViewGPU* views;             // ordered by pathBegin
    const unsigned long long* viewBegin;  // [viewCount + 1] path index ranges
    uint32_t viewCount;
    unsigned long long* viewClock;    // wall clock when a view completed
    // frame order (mray_renderer_render_frame on a batch): path g is pixel
    // (g mod pixTotal) of the views laid end to end, so every frame gives
    // every view its samples at once
    uint32_t frameMajor;
    unsigned long long pixTotal;
    const unsigned long long* viewPix;    // [viewCount + 1] pixel prefix sums
    // tiled renderers: the current tile (device memory, rewritten between
    // tiles), NULL otherwise
    const TileGPU* tile;
};

// View whose path range holds global path index g: the last v in [lo, hi]
// with begin[v] <= g (views ordered by their first path index).
__device__ __forceinline__ uint32_t ViewOfPath(const unsigned long long* __restrict__ begin,
                                               uint32_t lo, uint32_t hi, unsigned long long g)
{
    while(lo < hi)
    {
        uint32_t mid = (lo + hi + 1u) >> 1;
        if(begin[mid] <= g) lo = mid;
        else hi = mid - 1u;
    }
    return lo;
}

// Scene of slot i: the whole scene, or in a batch the merged arrays with the
// slot's view's lights and boundary light.
template<bool BATCH>
__device__ __forceinline__ SceneView SlotScene(const RenderConst& rc, const SlotSoA& S, uint32_t i)
{
    SceneView s = rc.scene;
    if constexpr(BATCH)
    {
        const ViewGPU& v = rc.views[S.view[i]];
        s.lightPos += size_t(v.lightOff) * 9u;
        s.lightRad += v.lightOff;
        s.lightCount = v.lightCount;
        s.boundaryType = v.boundaryType;
        s.boundaryRad[0] = v.boundaryRad[0];
        s.boundaryRad[1] = v.boundaryRad[1];
        s.boundaryRad[2] = v.boundaryRad[2];
        s.sceneDiameter = v.sceneDiameter;
        s.envTex = v.envTex;
        s.envCdfX = v.envCdfX;
        s.envCdfY = v.envCdfY;
        s.envW = v.envW;
        s.envH = v.envH;
        if(s.lightAlias)
        {
            s.lightAlias += v.lightOff;
            s.lightPmf += v.lightOff;
        }
        if(s.lightCoeff) s.lightCoeff += v.lightOff;
    }
    return s;
}

static constexpr uint32_t kSlotTPB = 256;
static constexpr uint32_t kMatQueues = 5;   // boundary/light + 4 material types
static constexpr uint32_t kGraphIters = 8;  // bounces per captured hipGraph
// Bursts queued ahead of the one whose scheduler snapshot the host reads: the
// GPU always holds kLookahead bursts of queued work while the host decides
// whether the frame is done.
static constexpr uint32_t kLookahead = 2;
static constexpr uint32_t kRingBursts = kLookahead + 1;

// ---------------------------------------------------------------------------
// Refill
// ---------------------------------------------------------------------------
// Per-block counts: x = free slots (dead or invalid), y = live slots,
// z = finished paths waiting for the film.  KRefillScan turns x and y into
// exclusive prefixes.
template<bool BATCH>
__global__ void __launch_bounds__(kSlotTPB)
KRefillCount(const uint32_t* __restrict__ meta, const uint32_t* __restrict__ view, uint32_t N,
             uint4* __restrict__ blockCnt)
{
    __shared__ uint32_t wsum[4][kSlotTPB / 64];
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t st = (i < N) ? MStatus(meta[i]) : ST_ALIVE;
    // batch: lowest view among the paths still running in this block
    uint32_t vmin = 0xFFFFFFFFu;
    if constexpr(BATCH)
    {
        if(i < N && st == ST_ALIVE) vmin = view[i];
        for(int o = 32; o > 0; o >>= 1) vmin = min(vmin, uint32_t(__shfl_xor(int(vmin), o)));
    }
    uint64_t bf = __ballot((i < N) && st != ST_ALIVE);
    uint64_t bl = __ballot((i < N) && st == ST_ALIVE);
    uint64_t bd = __ballot((i < N) && st == ST_DEAD);
    const uint32_t w = threadIdx.x >> 6;
    if((threadIdx.x & 63u) == 0)
    {
        wsum[0][w] = uint32_t(__popcll(bf));
        wsum[1][w] = uint32_t(__popcll(bl));
        wsum[2][w] = uint32_t(__popcll(bd));
        wsum[3][w] = vmin;
    }
    __syncthreads();
    if(threadIdx.x == 0)
    {
        uint4 c = make_uint4(0, 0, 0, 0xFFFFFFFFu);
        for(uint32_t k = 0; k < kSlotTPB / 64; k++)
        {
            c.x += wsum[0][k];
            c.y += wsum[1][k];
            c.z += wsum[2][k];
            c.w = min(c.w, wsum[3][k]);
        }
        blockCnt[blockIdx.x] = c;
    }
}

// One block: exclusive scans of the free and live counts, then the refill
// amount (ReloadPaths with TotalSampleLimit, PathTracerRendererBase.h:59-67).
// Batch views complete when every path below both the lowest view still
// running and the view of the next unissued path has reached the film:
// those are stamped with the wall clock here (their last paths are splatted
// by the refill that follows).
__global__ void __launch_bounds__(1024)
KRefillScan(uint4* __restrict__ blockCnt, uint32_t count, Sched* __restrict__ sched,
            const unsigned long long* __restrict__ viewBegin, uint32_t viewCount,
            unsigned long long* __restrict__ viewClock, uint32_t* __restrict__ ctr)
{
    // this bounce's work and queue counters start empty
    if(threadIdx.x < CTR_COUNT) *Ctr(ctr, threadIdx.x) = 0u;
    // exclusive scans of the free (x) and live (y) counts: 8 blocks per
    // thread, a wave scan of the thread totals, then a scan of the 16 wave
    // totals (two barriers per 8192 blocks)
    constexpr uint32_t kPer = 8;
    __shared__ uint32_t wsF[16], wsL[16];
    __shared__ uint32_t vmin;
    __shared__ unsigned long long accum;
    if(threadIdx.x == 0) { accum = 0; vmin = 0xFFFFFFFFu; }
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t carryF = 0, carryL = 0;
    unsigned long long myAccum = 0;
    uint32_t myMin = 0xFFFFFFFFu;
    for(uint32_t base = 0; base < count; base += 1024u * kPer)
    {
        const uint32_t i0 = base + threadIdx.x * kPer;
        uint32_t f[kPer], l[kPer], z[kPer];
        uint32_t tf = 0, tl = 0;
        #pragma unroll
        for(uint32_t k = 0; k < kPer; k++)
        {
            const uint32_t i = i0 + k;
            const uint4 v = (i < count) ? blockCnt[i] : make_uint4(0, 0, 0, 0xFFFFFFFFu);
            myAccum += v.z;
            myMin = min(myMin, v.w);
            f[k] = v.x; l[k] = v.y; z[k] = v.z;
            tf += v.x; tl += v.y;
        }
        uint32_t sF = tf, sL = tl;      // inclusive scan over the wave
        for(uint32_t o = 1; o < 64; o <<= 1)
        {
            const uint32_t a = __shfl_up(sF, o), b = __shfl_up(sL, o);
            if(lane >= o) { sF += a; sL += b; }
        }
        if(lane == 63) { wsF[wave] = sF; wsL[wave] = sL; }
        __syncthreads();
        if(wave == 0)
        {
            uint32_t x = (lane < 16) ? wsF[lane] : 0u, y = (lane < 16) ? wsL[lane] : 0u;
            for(uint32_t o = 1; o < 16; o <<= 1)
            {
                const uint32_t a = __shfl_up(x, o), b = __shfl_up(y, o);
                if(lane >= o) { x += a; y += b; }
            }
            if(lane < 16) { wsF[lane] = x; wsL[lane] = y; }
        }
        __syncthreads();
        uint32_t exF = carryF + (wave ? wsF[wave - 1] : 0u) + sF - tf;
        uint32_t exL = carryL + (wave ? wsL[wave - 1] : 0u) + sL - tl;
        #pragma unroll
        for(uint32_t k = 0; k < kPer; k++)
        {
            const uint32_t i = i0 + k;
            if(i < count) blockCnt[i] = make_uint4(exF, exL, z[k], 0u);
            exF += f[k];
            exL += l[k];
        }
        carryF += wsF[15];
        carryL += wsL[15];
        __syncthreads();                // wsF / wsL are rewritten next pass
    }
    // number of finished paths: wave reduction, then one LDS atomic per wave
    for(uint32_t o = 32; o > 0; o >>= 1) myAccum += __shfl_down(myAccum, o);
    if((threadIdx.x & 63u) == 0) atomicAdd(&accum, myAccum);
    if(viewClock)
    {
        for(int o = 32; o > 0; o >>= 1) myMin = min(myMin, uint32_t(__shfl_xor(int(myMin), o)));
        if((threadIdx.x & 63u) == 0) atomicMin(&vmin, myMin);
    }
    __syncthreads();
    if(threadIdx.x == 0 && viewClock)
    {
        const unsigned long long next = sched->counter;
        uint32_t vb = (next >= viewBegin[viewCount]) ? viewCount
                                                     : ViewOfPath(viewBegin, 0u, viewCount - 1u, next);
        sched->viewBase = min(vb, viewCount - 1u);
        uint32_t bound = min(vb, vmin);
        if(bound > viewCount) bound = viewCount;
        const unsigned long long now = wall_clock64();
        for(uint32_t v = sched->viewsDone; v < bound; v++) viewClock[v] = now;
        if(bound > sched->viewsDone) sched->viewsDone = bound;
    }
    if(threadIdx.x == 0)
    {
        unsigned long long freeSlots = carryF;
        unsigned long long remaining = sched->limit - sched->counter;
        unsigned long long fill = (freeSlots < remaining) ? freeSlots : remaining;
        sched->base = sched->counter;
        sched->fill = uint32_t(fill);
        sched->counter += fill;
        sched->activeCount = carryL + uint32_t(fill);
        sched->deadTotal += accum;
        // a bounce after the frame has drained (queued ahead by the host)
        // does nothing and is not counted
        if(sched->activeCount != 0 || accum != 0)
        {
            if(sched->iterations == 0) sched->startClock = wall_clock64();
            sched->iterations += 1;
        }
    }
}

// Finished path -> film (AddRadianceToRenderBufferThroughput,
// PathTracerRendererBase.cu:215-262; accumulation as TextureFilter.cu:672-697):
// spectral paths are converted to RGB first; a non-finite sample is written
// as magenta with a large weight so it stays visible.
template<bool SPECTRAL, bool BATCH>
__device__ __forceinline__ void SplatToFilm(const SlotSoA& S, const RenderConst& rc, uint32_t i,
                                            uint32_t m)
{
    float4 r4 = S.rad[i];
    V4 rad = {r4.x, r4.y, r4.z, r4.w};
    if constexpr(SPECTRAL)
    {
        float4 w4 = S.waves[i], p4 = S.wpdf[i];
        rad = SpectrumToRGB(rc.spec, rad, V4{w4.x, w4.y, w4.z, w4.w},
                            V4{p4.x, p4.y, p4.z, p4.w}, MDispersed(m));
    }
    float w = S.filterW[i];
    float r = rad.x, g = rad.y, b = rad.z;
    if(!(isfinite(r) && isfinite(g) && isfinite(b) && isfinite(rad.w)))
    {
        r = 1e7f; g = 0.0f; b = 1e7f;
        w = w * 128.0f;
    }
    uint32_t px = S.pixel[i];
    float* film = rc.film;
    size_t P = rc.P;
    if constexpr(BATCH)
    {
        const ViewGPU& v = rc.views[S.view[i]];
        film = v.film;
        P = v.P;
    }
    atomicAdd(film + px, r);
    atomicAdd(film + P + px, g);
    atomicAdd(film + 2 * P + px, b);
    atomicAdd(film + 3 * P + px, w);
}

// Training pairs of the training paths that finished in the last bounce:
// target = (final radiance - radiance before the vertex) / throughput, per
// channel; in spectral mode the four wavelengths' targets become one RGB
// value as the film converts a path (SpectrumToRGB).  One lane per (path,
// recorded vertex), G lanes per path (G a power of two >= maxVerts, dividing
// 64); each path's first lane reserves its entries in the ring of its view
// (block-wide per view in LDS, then one global atomic per view and block).
template<bool SPECTRAL, bool BATCH>
__global__ void __launch_bounds__(256)
KNrcEmit(SlotSoA S, RenderConst rc, Sched* __restrict__ sched, uint32_t log2G)
{
    __shared__ uint32_t sCnt[kMaxNrcViews];
    __shared__ unsigned long long sBase[kMaxNrcViews];
    const NrcRenderState& nr = rc.nrc;
    const uint32_t G = 1u << log2G;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    if(threadIdx.x < kMaxNrcViews) sCnt[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t ts = t >> log2G, v = t & (G - 1u);
    const uint32_t i = ts << nr.trainShift;
    const bool fin = (ts << nr.trainShift) < rc.N && (i >> nr.trainShift) == ts &&
                     MStatus(S.meta[i]) == ST_DEAD;
    const uint32_t nv = fin ? nr.trainCount[ts] : 0u;
    const uint32_t view = (BATCH && fin) ? S.view[i] : 0u;
    uint32_t mine = 0;
    if(v == 0u && nv > 0u) mine = atomicAdd(&sCnt[view], nv);
    __syncthreads();
    if(threadIdx.x < nr.views && sCnt[threadIdx.x] > 0u)
        sBase[threadIdx.x] = atomicAdd(&nr.ringCount[threadIdx.x], (unsigned long long)sCnt[threadIdx.x]);
    __syncthreads();
    const int src = int(lane & ~(G - 1u));
    const uint32_t first = uint32_t(__shfl(int(mine), src));
    if(!fin || v >= nv) return;
    const unsigned long long base = sBase[view] + first;
    const float* vx = nr.trainStore + (size_t(ts) * nr.maxVerts + v) * kNrcVert;
    const size_t slot = size_t(view) * nr.ringSize + size_t((base + v) & (nr.ringSize - 1u));
    for(uint32_t k = 0; k < kNrcRec; k++) nr.ringRec[slot * kNrcRec + k] = vx[k];
    const float4 fr = S.rad[i];
    const float fin4[4] = {fr.x, fr.y, fr.z, fr.w};
    float tgt[4];
    for(uint32_t c = 0; c < 4; c++)
    {
        float thr = vx[kNrcRec + c], before = vx[kNrcRec + 4 + c];
        float x = (thr > 1e-12f) ? (fin4[c] - before) / thr : 0.0f;
        tgt[c] = isfinite(x) ? x : 0.0f;
    }
    if constexpr(SPECTRAL)
    {
        const float4 w4 = S.waves[i], p4 = S.wpdf[i];
        V4 rgb = SpectrumToRGB(rc.spec, V4{tgt[0], tgt[1], tgt[2], tgt[3]},
                               V4{w4.x, w4.y, w4.z, w4.w}, V4{p4.x, p4.y, p4.z, p4.w},
                               MDispersed(S.meta[i]));
        tgt[0] = rgb.x; tgt[1] = rgb.y; tgt[2] = rgb.z;
    }
    for(uint32_t c = 0; c < 3; c++) nr.ringTgt[slot * 3 + c] = isfinite(tgt[c]) ? MaxF(tgt[c], 0.0f) : 0.0f;
}

// Camera path of global pixel `pix` of a W-wide image seen through camera c
// into slot i (KCGenerateCamRaysStochastic, RayGenKernels.kt.h:155-226, and
// KCInitPathStateIndirect, PathTracerRendererBase.cu:35-57).
template<bool SPECTRAL>
__device__ __forceinline__ void StartCameraPath(const SlotSoA& S, const CameraGPU& c, uint32_t W,
                                                uint32_t i, uint32_t pix)
{
    uint32_t px = pix % W, py = pix / W;
    uint32_t st = S.rng[i];
    float xi0 = NextRN(st), xi1 = NextRN(st);
    float o0 = SampleGaussian(xi0, c.sigma);
    float o1 = SampleGaussian(xi1, c.sigma);
    float offPdf = Gaussian(o0, c.sigma) * Gaussian(o1, c.sigma);
    float weight = Gaussian(o0, c.sigma) * Gaussian(o1, c.sigma);
    float dx = c.planeSize[0] / float(c.width), dy = c.planeSize[1] / float(c.height);
    float sdx = (float(px) + (o0 + 0.5f)) * dx;
    float sdy = (float(py) + (o1 + 0.5f)) * dy;
    V3 sp = {c.bottomLeft[0], c.bottomLeft[1], c.bottomLeft[2]};
    sp = sp + sdx * V3{c.right[0], c.right[1], c.right[2]};
    sp = sp + sdy * V3{c.up[0], c.up[1], c.up[2]};
    V3 cp = {c.position[0], c.position[1], c.position[2]};
    V3 dir = Normalize(sp - cp);
    S.rayO[i] = make_float4(cp.x, cp.y, cp.z, c.nearFar[0]);
    S.rayD[i] = make_float4(dir.x, dir.y, dir.z, c.nearFar[1]);
    S.filterW[i] = weight / offPdf;
    S.pixel[i] = pix;
    S.rad[i] = make_float4(0, 0, 0, 0);
    S.thr[i] = make_float4(1, 1, 1, 1);
    S.prevPdf[i] = 0.0f;
    if constexpr(SPECTRAL)
    {
        float xw = NextRN(st);
        V4 w, p;
        SampleWavelengths(xw, w, p);
        S.waves[i] = make_float4(w.x, w.y, w.z, w.w);
        S.wpdf[i] = make_float4(p.x, p.y, p.z, p.w);
    }
    S.rng[i] = st;
    S.meta[i] = MPack(0, ST_ALIVE, RT_CAMERA, false, false);
}

template<bool SPECTRAL, bool BATCH>
__global__ void __launch_bounds__(kSlotTPB)
KRefillApply(SlotSoA S, RenderConst rc, const uint4* __restrict__ blockOff,
             Sched* __restrict__ sched, uint32_t* __restrict__ active)
{
    __shared__ uint32_t wsum[kSlotTPB / 64], asum[kSlotTPB / 64];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool inRange = i < rc.N;
    uint32_t m = inRange ? S.meta[i] : MPack(0, ST_INVALID, 0, false, false);
    const uint32_t status = MStatus(m);
    const bool freeSlot = inRange && status != ST_ALIVE;
    if(inRange && status == ST_DEAD)
    {
        SplatToFilm<SPECTRAL, BATCH>(S, rc, i, m);
    }
    uint64_t b = __ballot(freeSlot);
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    if(lane == 0) wsum[wave] = uint32_t(__popcll(b));
    __syncthreads();
    uint32_t waveOff = 0;
    for(uint32_t w = 0; w < wave; w++) waveOff += wsum[w];
    const uint4 off = blockOff[blockIdx.x];
    const uint32_t fill = sched->fill;
    bool alive = inRange && !freeSlot;
    if(freeSlot)
    {
        uint32_t rank = off.x + waveOff + uint32_t(__popcll(b & lt));
        if(rank < fill)
        {
            const unsigned long long g = sched->base + rank;
            if constexpr(BATCH)
            {
                uint32_t lo, pix;
                if(rc.frameMajor)
                {
                    // frame order: the view whose pixel range holds g mod pixTotal
                    const unsigned long long pr = g % rc.pixTotal;
                    lo = ViewOfPath(rc.viewPix, 0u, rc.viewCount - 1u, pr);
                    pix = uint32_t(pr - rc.viewPix[lo]);
                }
                else
                {
                    // the view whose path range holds g (at or after this refill's first)
                    lo = ViewOfPath(rc.viewBegin, sched->viewBase, rc.viewCount - 1u, g);
                    pix = uint32_t((g - rc.views[lo].pathBegin) % rc.views[lo].P);
                }
                const ViewGPU& v = rc.views[lo];
                S.view[i] = lo;
                StartCameraPath<SPECTRAL>(S, v.cam, v.W, i, pix);
            }
            else if(rc.tile)
            {
                // path g of the tile is its pixel g mod P, rendered through
                // the tile's sub-camera and splatted at its film position
                const TileGPU& t = *rc.tile;
                const uint32_t local = uint32_t(g % t.P);
                StartCameraPath<SPECTRAL>(S, t.cam, t.w, i, local);
                S.pixel[i] = (t.y0 + local / t.w) * rc.W + t.x0 + local % t.w;
            }
            else StartCameraPath<SPECTRAL>(S, rc.cam, rc.W, i, uint32_t(g % rc.P));
            if(rc.nrc.on)
            {
                S.spread[i] = make_float2(0.0f, 0.0f);
                if((i & ((1u << rc.nrc.trainShift) - 1u)) == 0u)
                    rc.nrc.trainCount[i >> rc.nrc.trainShift] = 0;
            }
            alive = true;
        }
        else if(status != ST_INVALID)
        {
            S.meta[i] = MPack(MDepth(m), ST_INVALID, MType(m), MDispersed(m), false);
        }
    }
    // active list position: live slots of earlier blocks, slots refilled in
    // earlier blocks, then this block's earlier live slots (ascending order)
    uint64_t am = __ballot(alive);
    if(lane == 0) asum[wave] = uint32_t(__popcll(am));
    __syncthreads();
    if(alive)
    {
        uint32_t pos = off.y + min(off.x, fill);
        for(uint32_t w = 0; w < wave; w++) pos += asum[w];
        active[pos + uint32_t(__popcll(am & lt))] = i;
    }
}

// ---------------------------------------------------------------------------
// Closest hit
// ---------------------------------------------------------------------------
template<int W, bool COUNT, bool BATCH, int TPB = kTraceTPB, bool INST = false>
__global__ void __launch_bounds__(TPB)
KTrace(TraceCtx ctx, SlotSoA S, const uint32_t* __restrict__ active, Sched* __restrict__ sched,
       const ViewGPU* __restrict__ views, uint32_t* __restrict__ ctr)
{
    __shared__ uint2 lstack[kLdsStack * TPB];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = sched->activeCount;
    if(gtid == 0) sched->extRays += n;
    uint32_t nodes = 0, tris = 0;
    FetchLoop(Ctr(ctr, CTR_TRACE), n, [&](uint32_t k)
    {
        const uint32_t i = active[k];
        const float4 o = S.rayO[i], d = S.rayD[i];
        HitRec h;
        if constexpr(INST)
        {
            uint32_t instId;
            h = TraceInstFlat<false, COUNT>(ctx, lstack, gtid, V3{o.x, o.y, o.z}, V3{d.x, d.y, d.z},
                                        o.w, d.w, instId, &nodes, &tris);
            S.inst[i] = instId;
        }
        else
        {
            const uint32_t root = BATCH ? views[S.view[i]].root : 0u;
            h = TraceW<W, false, COUNT>(ctx, lstack, gtid, V3{o.x, o.y, o.z}, V3{d.x, d.y, d.z},
                                        o.w, d.w, root, &nodes, &tris);
        }
        S.hit[i] = make_uint4(h.prim, __float_as_uint(h.t), __float_as_uint(h.a),
                              __float_as_uint(h.b));
    });
    if(COUNT)
    {
        unsigned long long nv = nodes, nt = tris;
        for(int o = 32; o > 0; o >>= 1)
        {
            nv += __shfl_xor(nv, o);
            nt += __shfl_xor(nt, o);
        }
        if((threadIdx.x & 63u) == 0 && (nv | nt))
        {
            atomicAdd(&sched->nodeVisits, nv);
            atomicAdd(&sched->triTests, nt);
        }
    }
}

// Wave-uniform source of 64-item chunks of an n-item list: the chunks of
// this workgroup's XCD part first, then the other parts (the order FetchLoop
// hands them out in).
struct ChunkSource
{
    uint32_t* ctr;
    uint32_t n, C, part, tries;
    __device__ __forceinline__ ChunkSource(uint32_t* c, uint32_t n_)
        : ctr(c), n(n_), C((n_ + 63u) / 64u), part(blockIdx.x % kXcds), tries(0) {}
    // next chunk [begin, end) or false when every part is used up
    __device__ __forceinline__ bool Next(uint32_t& begin, uint32_t& end)
    {
        const uint32_t lane = threadIdx.x & 63u;
        while(tries < kXcds)
        {
            const uint32_t c0 = __builtin_amdgcn_readfirstlane(uint32_t((uint64_t(C) * part) / kXcds));
            const uint32_t c1 = __builtin_amdgcn_readfirstlane(uint32_t((uint64_t(C) * (part + 1u)) / kXcds));
            uint32_t* fc = ctr + size_t(part) * kCtrStride;
            uint32_t c = c1 - c0;
            if(lane == 0)
            {
                c = __hip_atomic_load(fc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if(c < c1 - c0) c = atomicAdd(fc, 1u);
            }
            c = __builtin_amdgcn_readfirstlane(c);
            if(c < c1 - c0)
            {
                begin = (c0 + c) * 64u;
                end = min(begin + 64u, n);
                return true;
            }
            part = (part + 1u) % kXcds;
            tries++;
        }
        return false;
    }
};

// Closest hit with lane refill (Aila and Laine 2009, "Understanding the
// Efficiency of Ray Traversal on GPUs": persistent threads that fetch new
// rays): a lane whose walk has ended takes the next ray of the wave's chunk
// once a quarter of the wave is idle, so a wave no longer waits for the
// longest of its 64 rays.  Each ray's walk is the same as KTrace's, so hits
// are identical; only the order rays are walked in changes.
constexpr uint32_t kRefillAt = 16;
template<bool COUNT, bool BATCH>
__global__ void __launch_bounds__(kTraceTPB)
KTraceRefill(TraceCtx ctx, SlotSoA S, const uint32_t* __restrict__ active, Sched* __restrict__ sched,
             const ViewGPU* __restrict__ views, uint32_t* __restrict__ ctr)
{
    __shared__ uint2 lstack[kLdsStack * kTraceTPB];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t below = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    const uint32_t n = sched->activeCount;
    if(gtid == 0) sched->extRays += n;
    ChunkSource src(Ctr(ctr, CTR_TRACE), n);
    uint32_t pool = 0, poolEnd = 0;        // wave-uniform: items not yet handed out
    bool more = true;                      // wave-uniform: chunks may remain
    bool has = false;
    uint32_t slot = 0, visits = 0;
    Walk w;
    uint32_t nodes = 0, tris = 0;
    for(;;)
    {
        uint64_t idle = __ballot(!has);
        if(more && uint32_t(__popcll(idle)) >= kRefillAt)
        {
            while(idle != 0ull)
            {
                if(pool >= poolEnd)
                {
                    uint32_t b = 0, e = 0;
                    more = src.Next(b, e);
                    if(!more) break;
                    pool = b;
                    poolEnd = e;
                    continue;
                }
                const uint32_t take = min(poolEnd - pool, uint32_t(__popcll(idle)));
                if(!has)
                {
                    const uint32_t rank = uint32_t(__popcll(idle & below));
                    if(rank < take)
                    {
                        slot = active[pool + rank];
                        const float4 o = S.rayO[slot], d = S.rayD[slot];
                        const uint32_t root = BATCH ? views[S.view[slot]].root : 0u;
                        WalkInit(w, V3{o.x, o.y, o.z}, V3{d.x, d.y, d.z}, o.w, d.w, root, 0u);
                        visits = 0;
                        has = true;
                    }
                }
                pool += take;
                idle = __ballot(!has);
            }
        }
        if(__ballot(has) == 0ull)
        {
            if(!more) break;
            continue;
        }
        if(has)
        {
            // a walk visits at most maxVisits nodes (as TraceW)
            bool over = WalkStep<4, false, COUNT>(ctx, lstack, gtid, w, 0u, &nodes, &tris);
            if(++visits >= ctx.maxVisits) over = true;
            if(over)
            {
                S.hit[slot] = make_uint4(w.best.prim, __float_as_uint(w.best.t),
                                         __float_as_uint(w.best.a), __float_as_uint(w.best.b));
                has = false;
            }
        }
    }
    if(COUNT)
    {
        unsigned long long nv = nodes, nt = tris;
        for(int o = 32; o > 0; o >>= 1)
        {
            nv += __shfl_xor(nv, o);
            nt += __shfl_xor(nt, o);
        }
        if((threadIdx.x & 63u) == 0 && (nv | nt))
        {
            atomicAdd(&sched->nodeVisits, nv);
            atomicAdd(&sched->triTests, nt);
        }
    }
}

// ---------------------------------------------------------------------------
// Lights (LightsDefault.hpp, LightSampler.hpp, MetaLight.hpp:470-520)
// ---------------------------------------------------------------------------
__device__ __forceinline__ V3 CoOctaToDir(float s, float t)
{
    const float piOvr4 = kPi * 0.25f;
    float u = s * 2.0f - 1.0f, v = t * 2.0f - 1.0f;
    float ua = fabsf(u), va = fabsf(v);
    float d = 1.0f - (ua + va);
    float radius = 1.0f - fabsf(d);
    float phiPrime = 0.0f;
    if(radius != 0.0f) phiPrime = ((va - ua) / radius + 1) * piOvr4;
    float sinP, cosP;
    fm::SinCos(phiPrime, sinP, cosP);
    float cosPhi = (signbit(u) ? -1.0f : 1.0f) * cosP;
    float sinPhi = (signbit(v) ? -1.0f : 1.0f) * sinP;
    float z = (signbit(d) ? -1.0f : 1.0f) * (1.0f - radius * radius);
    float xyFactor = radius * sqrtf(2.0f - radius * radius);
    return {cosPhi * xyFactor, sinPhi * xyFactor, z};
}

// Graphics::DirectionToConcentricOctahedral (GraphicsFunctions.h:318-353),
// Z-up unit direction -> uv in [0, 1]^2 (the reference's quirk kept: the
// +Z pole maps to uv (0, 0)).
__device__ __forceinline__ void DirToCoOcta(V3 d, float& s, float& t)
{
    const float twoOvrPi = kInvPi * 2.0f;
    if(d.x == 0.0f && d.y == 0.0f) { s = 0.0f; t = 0.0f; return; }
    float xAbs = fabsf(d.x), yAbs = fabsf(d.y);
    float phiPrime = atanf(yAbs / xAbs);
    float radius = sqrtf(MaxF(0.0f, 1.0f - fabsf(d.z)));
    float v = radius * twoOvrPi * phiPrime;
    float u = radius - v;
    if(d.z < 0.0f)
    {
        float uP = 1.0f - v, vP = 1.0f - u;
        u = uP;
        v = vP;
    }
    u *= signbit(d.x) ? -1.0f : 1.0f;
    v *= signbit(d.y) ? -1.0f : 1.0f;
    s = (u + 1.0f) * 0.5f;
    t = (v + 1.0f) * 0.5f;
}

__device__ __forceinline__ bool IsSkysphere(uint32_t t)
{
    return t == MRAY_BOUNDARY_SKYSPHERE_COOCTA || t == MRAY_BOUNDARY_SKYSPHERE_SPHERICAL;
}
// SphericalCoordConverter (LightsDefault.hpp:177-222): Z-up spherical
// angles, u = (azimuth + pi) / (2 pi), v = 1 - inclination / pi.
__device__ __forceinline__ void DirToSpherical(V3 dz, float& u, float& v)
{
    float azimuth = fm::Atan2(dz.y, dz.x);
    float incl = fm::Acos(ClampF(dz.z, -1.0f, 1.0f));
    u = (azimuth + kPi) * 0.5f / kPi;
    v = 1.0f - incl * kInvPi;
}
__device__ __forceinline__ V3 SphericalToDir(float u, float v, float& sinIncl)
{
    float theta = (u * kPi * 2.0f) - kPi;
    float phi = (1.0f - v) * kPi;
    float sT, cT, sP, cP;
    fm::SinCos(theta, sT, cT);
    fm::SinCos(phi, sP, cP);
    sinIncl = sP;
    return {cT * sP, sT * sP, cP};
}
// Skysphere texture coordinates of a Y-up direction for the boundary type
__device__ __forceinline__ void SkyUV(uint32_t type, V3 dirYUp, float& u, float& v)
{
    const V3 dz = {dirYUp.z, dirYUp.x, dirYUp.y};           // YUpToZUp
    if(type == MRAY_BOUNDARY_SKYSPHERE_SPHERICAL) DirToSpherical(dz, u, v);
    else DirToCoOcta(dz, u, v);
}
// Solid-angle pdf of a uv pdf: equal-area CoOcta 1 / (4 pi); spherical
// 1 / (2 pi^2 sin(inclination)), zero at the poles
__device__ __forceinline__ float SkySolidAnglePdf(uint32_t type, float uvPdf, float sinIncl)
{
    if(type == MRAY_BOUNDARY_SKYSPHERE_SPHERICAL)
        return (sinIncl <= 0.0f) ? 0.0f : uvPdf / (2.0f * kPi * kPi * sinIncl);
    return uvPdf * 0.25f * kInvPi;
}

// Textured Skysphere radiance at Y-up direction dirYUp: bilinear, wrapped
// edges, texel centres at (i + 1/2) / size (the reference's default
// texture parameters, TracerI.h:71-82), evaluated in fp32.
__device__ __forceinline__ V3 EnvRadiance(const SceneView& s, V3 dirYUp)
{
    float u, v;
    SkyUV(s.boundaryType, dirYUp, u, v);
    const int W = int(s.envW), H = int(s.envH);
    float xf = u * float(W) - 0.5f, yf = v * float(H) - 0.5f;
    float x0f = floorf(xf), y0f = floorf(yf);
    float fx = xf - x0f, fy = yf - y0f;
    int x0 = int(x0f), y0 = int(y0f);
    int x1 = x0 + 1, y1 = y0 + 1;
    x0 = ((x0 % W) + W) % W; x1 = ((x1 % W) + W) % W;
    y0 = ((y0 % H) + H) % H; y1 = ((y1 % H) + H) % H;
    const float4 a = s.envTex[y0 * W + x0], b = s.envTex[y0 * W + x1];
    const float4 c = s.envTex[y1 * W + x0], d = s.envTex[y1 * W + x1];
    const float gx = 1.0f - fx, gy = 1.0f - fy;
    V3 top = {a.x * gx + b.x * fx, a.y * gx + b.y * fx, a.z * gx + b.z * fx};
    V3 bot = {c.x * gx + d.x * fx, c.y * gx + d.y * fx, c.z * gx + d.z * fx};
    return {top.x * gy + bot.x * fy, top.y * gy + bot.y * fy, top.z * gy + bot.z * fy};
}

// rocprim::lower_bound: the first i with cdf[i] >= xi (n if none).
__device__ __forceinline__ uint32_t LowerBoundF(const float* cdf, uint32_t n, float xi)
{
    uint32_t lo = 0, len = n;
    while(len > 0)
    {
        uint32_t half = len >> 1;
        if(cdf[lo + half] < xi) { lo += half + 1; len -= half + 1; }
        else len = half;
    }
    return lo;
}

// DistributionPwC<1>::SampleIndex (Distributions.h:128-148), with its
// clamp kept as written: an index of 1 or more steps down one ulp.
__device__ __forceinline__ float PwcSampleIndex(const float* cdf, uint32_t n, float xi, float& pdf)
{
    uint32_t index = LowerBoundF(cdf, n, xi);
    index = min(index, n - 1u);
    float prevCDF = (index == 0) ? 0.0f : cdf[index - 1];
    float myCDF = cdf[index];
    float t = (xi - prevCDF) / (myCDF - prevCDF);
    float indexF = float(index) + t;
    indexF = (indexF < 1.0f) ? indexF : __uint_as_float(__float_as_uint(indexF) - 1u);
    pdf = (myCDF - prevCDF) * float(n);
    return indexF;
}
__device__ __forceinline__ float PwcPdfIndex(const float* cdf, uint32_t n, float index)
{
    uint32_t i = uint32_t(index);
    float prevCDF = (i == 0) ? 0.0f : cdf[i - 1];
    return (cdf[i] - prevCDF) * float(n);
}

// DistributionPwC<2>::SampleUV (Distributions.h:181-213): the row from
// xi.y on the marginal, the column from xi.x on that row.
__device__ __forceinline__ void EnvSampleUV(const SceneView& s, float xi0, float xi1,
                                            float& u, float& v, float& pdf)
{
    float pdfY, pdfX;
    float yF = PwcSampleIndex(s.envCdfY, s.envH, xi1, pdfY);
    uint32_t row = uint32_t(yF);
    float xF = PwcSampleIndex(s.envCdfX + size_t(row) * s.envW, s.envW, xi0, pdfX);
    u = xF * (1.0f / float(s.envW));
    v = yF * (1.0f / float(s.envH));
    pdf = pdfY * pdfX;
}

// DistributionPwC<2>::PdfUV (Distributions.h:232-243)
__device__ __forceinline__ float EnvPdfUV(const SceneView& s, float u, float v)
{
    float xF = MinF(u * float(s.envW), float(s.envW) - 1.0f);
    float yF = MinF(v * float(s.envH), float(s.envH) - 1.0f);
    uint32_t row = uint32_t(yF);
    float pdfMarginal = PwcPdfIndex(s.envCdfY, s.envH, yF);
    float pdfConditional = PwcPdfIndex(s.envCdfX + size_t(row) * s.envW, s.envW, xF);
    return pdfConditional * pdfMarginal;
}

// LightSkysphere::PdfSolidAngle (LightsDefault.hpp:359-370) of direction dir
__device__ __forceinline__ float SkyspherePdf(const SceneView& s, V3 dir)
{
    float uvPdf = 1.0f, sinIncl = 1.0f;
    const V3 d = Normalize(dir);
    if(s.boundaryType == MRAY_BOUNDARY_SKYSPHERE_SPHERICAL)
        sinIncl = fm::Sin(fm::Acos(ClampF(d.y, -1.0f, 1.0f)));    // Z-up z = Y-up y
    if(s.envW != 0)
    {
        float u, v;
        SkyUV(s.boundaryType, d, u, v);
        uvPdf = EnvPdfUV(s, u, v);
    }
    return SkySolidAnglePdf(s.boundaryType, uvPdf, sinIncl);
}

template<bool SPECTRAL>
__device__ __forceinline__ V4 ConvRadiance(const RenderConst& rc, V3 rgb, V4 waves)
{
    if constexpr(SPECTRAL) return SpectralRadiance(rc.spec, rgb, waves);
    else return V4{rgb.x, rgb.y, rgb.z, 0.0f};
}
template<bool SPECTRAL>
__device__ __forceinline__ V4 ConvAlbedo(const RenderConst& rc, V3 rgb, V4 waves)
{
    if constexpr(SPECTRAL) return SpectralAlbedo(rc.spec, rgb, waves);
    else return V4{rgb.x, rgb.y, rgb.z, 0.0f};
}
// A light's radiance at the path's wavelengths: spectral renderers evaluate
// the coefficients converted once per renderer (KSpectralCoeffs), the same
// numbers SpectralRadiance computes for that colour on every call.
template<bool SPECTRAL>
__device__ __forceinline__ V4 LightRadiance(const RenderConst& rc, const SceneView& s, uint32_t li,
                                            float4 lr, V4 waves)
{
    if constexpr(SPECTRAL) return EvalRadianceCoeffs(rc.spec, s.lightCoeff[li], waves);
    else return V4{lr.x, lr.y, lr.z, 0.0f};
}

// Per material the albedo's coefficients, per light its radiance's
// coefficients and scale (spectral renderers, once at creation).
__global__ void __launch_bounds__(256)
KSpectralCoeffs(SpectralView sv, const MRayMaterial* __restrict__ mats, uint32_t nm,
                const float4* __restrict__ lightRad, uint32_t nl, float4* __restrict__ matCoeff,
                float4* __restrict__ lightCoeff)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < nm)
    {
        const MRayMaterial& m = mats[i];
        V3 co = AlbedoCoeffs(sv, V3{m.albedo[0], m.albedo[1], m.albedo[2]});
        matCoeff[i] = make_float4(co.x, co.y, co.z, 0.0f);
    }
    if(i < nl)
    {
        const float4 lr = lightRad[i];
        lightCoeff[i] = RadianceCoeffs(sv, V3{lr.x, lr.y, lr.z});
    }
}

template<bool SPECTRAL>
__device__ __forceinline__ V4 TriLightEmitViaPoint(const RenderConst& rc, const SceneView& s,
                                                   uint32_t li, V3 wO, V3 point, V4 waves)
{
    V3 p0, p1, p2;
    LoadLightTri(s, li, p0, p1, p2);
    float ha, hb;
    ProjectedHit(p0, p1, p2, point, ha, hb);
    if(!BaryInRange(ha, hb)) return Splat4(0.0f);
    V3 n = TriNormal(p0, p1, p2);
    float NdL = Dot(n, wO);
    float4 lr = s.lightRad[li];
    bool twoSided = lr.w != 0.0f;
    if(!twoSided && NdL <= 0.0f) return Splat4(0.0f);
    return LightRadiance<SPECTRAL>(rc, s, li, lr, waves);
}

struct LightSample { V3 position; V4 emission; float pdf; };

// DirectLightSamplerUniform::SampleLight (LightSampler.hpp:30-74)
template<bool SPECTRAL>
__device__ __forceinline__ LightSample SampleLight(const RenderConst& rc, const SceneView& s,
                                                   const float* rn, V3 distantPoint, V4 waves)
{
    uint32_t lightCount = s.lightCount + 1u;
    float lightCountF = float(lightCount);
    uint32_t li = uint32_t(rn[2] * lightCountF);
    li = min(li, lightCount - 1u);
    float selectionPdf = 1.0f / lightCountF;
    if(s.lightAlias && li < lightCount - 1u)
    {
        // power sampling: stratum li of the n + 1 is the alias bucket li
        const float f = rn[2] * lightCountF - float(li);
        const uint2 e = s.lightAlias[li];
        li = (f < __uint_as_float(e.x)) ? li : e.y;
        selectionPdf = s.lightPmf[li];
    }
    V3 point;
    float pointPdf;
    if(li < lightCount - 1u)
    {
        V3 p0, p1, p2;
        LoadLightTri(s, li, p0, p1, p2);
        float r1 = sqrtf(rn[0]);
        float r2 = rn[1];
        float a = 1 - r1;
        float b = (1 - r2) * r1;
        float c = r1 * r2;
        float area = TriArea(p0, p1, p2);
        float sPdf = 1.0f / area;
        V3 position = (p0 * a + p1 * b) + p2 * c;
        V3 normal = TriNormal(p0, p1, p2);
        V3 sd = distantPoint - position;
        float distSqr = LengthSqr(sd);
        sd = Normalize(sd);
        float NdL = Dot(normal, sd);
        bool twoSided = s.lightRad[li].w != 0.0f;
        NdL = twoSided ? fabsf(NdL) : MaxF(0.0f, NdL);
        float pdf = (NdL == 0.0f) ? 0.0f : sPdf / NdL;
        pdf *= distSqr;
        point = position;
        pointPdf = pdf;
    }
    else if(IsSkysphere(s.boundaryType))
    {
        // LightSkysphere::SampleSolidAngle (LightsDefault.hpp:340-357)
        float u = rn[0], v = rn[1], uvPdf = 1.0f, sinIncl = 1.0f;
        if(s.envW != 0) EnvSampleUV(s, rn[0], rn[1], u, v, uvPdf);
        V3 dz = (s.boundaryType == MRAY_BOUNDARY_SKYSPHERE_SPHERICAL) ? SphericalToDir(u, v, sinIncl)
                                                                       : CoOctaToDir(u, v);
        V3 dirYUp = {dz.y, dz.z, dz.x};
        point = distantPoint + dirYUp * s.sceneDiameter;
        pointPdf = SkySolidAnglePdf(s.boundaryType, uvPdf, sinIncl);
    }
    else
    {
        point = {1e10f, 1e10f, 1e10f};
        pointPdf = 0.0f;
    }
    V3 wO = Normalize(distantPoint - point);
    V4 emission;
    if(li < lightCount - 1u) emission = TriLightEmitViaPoint<SPECTRAL>(rc, s, li, wO, point, waves);
    else if(IsSkysphere(s.boundaryType))
        emission = ConvRadiance<SPECTRAL>(rc, s.envW ? EnvRadiance(s, Normalize(-wO))
                                                     : V3{s.boundaryRad[0], s.boundaryRad[1],
                                                          s.boundaryRad[2]}, waves);
    else emission = Splat4(0.0f);
    return {point, emission, selectionPdf * pointPdf};
}

__device__ __forceinline__ float PdfLightTri(const SceneView& s, uint32_t li, float ha, float hb,
                                             V3 rayPos, V3 rayDir)
{
    float selectionPDF = s.lightAlias ? s.lightPmf[li] : 1.0f / float(s.lightCount + 1u);
    V3 p0, p1, p2;
    LoadLightTri(s, li, p0, p1, p2);
    float pdf;
    if(!BaryInRange(ha, hb)) pdf = 0.0f;
    else
    {
        float hc = 1.0f - ha - hb;
        V3 position = (p0 * ha + p1 * hb) + p2 * hc;
        V3 normal = TriNormal(p0, p1, p2);
        pdf = 1.0f / TriArea(p0, p1, p2);
        float NdL = Dot(normal, -rayDir);
        bool twoSided = s.lightRad[li].w != 0.0f;
        NdL = twoSided ? fabsf(NdL) : MaxF(0.0f, NdL);
        pdf = (NdL == 0.0f) ? 0.0f : pdf / NdL;
        pdf *= LengthSqr(rayPos - position);
    }
    return pdf * selectionPDF;
}
__device__ __forceinline__ float PdfLightBoundary(const SceneView& s, V3 rayDir)
{
    float selectionPDF = 1.0f / float(s.lightCount + 1u);
    float pdf = IsSkysphere(s.boundaryType) ? SkyspherePdf(s, rayDir) : 0.0f;
    return pdf * selectionPDF;
}

// ---------------------------------------------------------------------------
// Shade: boundary / light work, NEE (work 1), scatter (work 0)
// ---------------------------------------------------------------------------
// Shading of one slot after its closest hit.  KEY < 0: any hit; KEY = 0:
// boundary or light hits only; KEY = 1 + material type: surface hits of that
// material only (the material switch folds away).
// What one shading step hands to its wave's appends: a shadow ray (written
// to the slot's sO/sD/sRad) and a radiance-cache query record.
struct ShadeOut
{
    bool shadow;
    bool query;
    uint32_t view;            // the path's view (its cache's query queue)
    float rec[kNrcRec];
};

template<bool SPECTRAL, int KEY, bool BATCH, bool INST = false>
__device__ __forceinline__ void ShadeSlot(const SlotSoA& S, const RenderConst& rc, uint32_t i,
                                          ShadeOut& out)
{
    const SceneView s = SlotScene<BATCH>(rc, S, i);
    uint4 hh = S.hit[i];
    uint32_t prim = hh.x;
    float ha = __uint_as_float(hh.z), hb = __uint_as_float(hh.w);
    uint32_t m = S.meta[i];
    uint32_t depth = MDepth(m), type = MType(m);
    bool dispersed = MDispersed(m);
    float4 ro = S.rayO[i], rd = S.rayD[i];
    V3 rpos = {ro.x, ro.y, ro.z}, rdir = {rd.x, rd.y, rd.z};
    float4 t4 = S.thr[i], r4 = S.rad[i];
    V4 thr = {t4.x, t4.y, t4.z, t4.w};
    V4 rad = {r4.x, r4.y, r4.z, r4.w};
    V4 waves = Splat4(0.0f);
    if constexpr(SPECTRAL)
    {
        float4 w4 = S.waves[i];
        waves = {w4.x, w4.y, w4.z, w4.w};
    }
    uint32_t li = kInvalid;
    const uint32_t instId = (INST && prim != kInvalid) ? S.inst[i] : 0u;
    if(prim != kInvalid)
    {
        uint32_t key = s.surfKey[prim];
        // an instance's copy of a mesh light sits at its own light offset
        if(key & MRAY_LIGHT_FLAG) li = (key & ~MRAY_LIGHT_FLAG) + (INST ? rc.inst[instId].lightOff : 0u);
    }
    bool isBoundary = (prim == kInvalid);
    if(KEY <= 0 && (isBoundary || li != kInvalid))
    {
        V3 wO = -rdir;
        V4 emission;
        if(isBoundary)
        {
            // LightSkysphere::EmitViaHit (LightsDefault.hpp:405-421)
            emission = IsSkysphere(s.boundaryType)
                ? ConvRadiance<SPECTRAL>(rc, s.envW ? EnvRadiance(s, Normalize(rdir))
                                                    : V3{s.boundaryRad[0], s.boundaryRad[1],
                                                         s.boundaryRad[2]}, waves)
                : Splat4(0.0f);
        }
        else if(!BaryInRange(ha, hb)) emission = Splat4(0.0f);
        else
        {
            V3 p0, p1, p2;
            LoadLightTri(s, li, p0, p1, p2);   // the hit triangle, in world space
            V3 nrm = TriNormal(p0, p1, p2);
            float NdL = Dot(nrm, wO);
            float4 lr = s.lightRad[li];
            emission = (lr.w == 0.0f && NdL <= 0.0f)
                ? Splat4(0.0f) : LightRadiance<SPECTRAL>(rc, s, li, lr, waves);
        }
        if(rc.mode == MRAY_SAMPLE_PURE)
        {
            depth += 1u;
            if(depth <= rc.rr1) rad = emission * thr;
        }
        else if(!(rc.mode == MRAY_SAMPLE_NEE && type != RT_CAMERA && type != RT_SPECULAR))
        {
            bool switchToMIS = (rc.mode == MRAY_SAMPLE_NEE_MIS && type == RT_PATH);
            V4 t2 = thr;
            if(switchToMIS)
            {
                float p0 = S.prevPdf[i];
                float p1 = isBoundary ? PdfLightBoundary(s, rdir)
                                      : PdfLightTri(s, li, ha, hb, rpos, rdir);
                float misPdf = Balance2(p0, 1.0f, p1, 1.0f);
                t2 = t2 * p0;
                t2 = DivideByPDF(t2, misPdf);
            }
            if((depth + 1u) <= rc.rr1) rad = rad + emission * t2;
        }
        S.rad[i] = make_float4(rad.x, rad.y, rad.z, rad.w);
        S.meta[i] = MPack(depth, ST_DEAD, type, dispersed, false);
        return;
    }

    if constexpr(KEY == 0) return;
    // Surface hit
    uint32_t st = S.rng[i];
    const MRayMaterial& M = s.mats[s.surfKey[prim]];
    Surface sf = INST ? GenSurfaceInst(s, prim, ha, hb, rdir, rc.inst[instId])
                      : GenSurface(s, prim, ha, hb, rdir);
    MatInst mat;
    mat.type = (KEY > 0) ? uint32_t(KEY - 1) : M.type;
    if constexpr(SPECTRAL)
    {
        const float4 co = s.matCoeff[s.surfKey[prim]];
        mat.albedo = EvalCoeffs(V3{co.x, co.y, co.z}, waves);
    }
    else mat.albedo = V4{M.albedo[0], M.albedo[1], M.albedo[2], 0.0f};
    mat.roughness = M.roughness;
    mat.metallic = M.metallic;
    mat.specular = M.specular;
    if(mat.type == MRAY_MAT_REFRACT)
    {
        if constexpr(SPECTRAL)
        {
            mat.frontIoR = CauchyIoR(waves.x, M.cauchyFront);
            mat.backIoR = CauchyIoR(waves.x, M.cauchyBack);
        }
        else
        {
            mat.frontIoR = M.cauchyFront[0];
            mat.backIoR = M.cauchyBack[0];
        }
    }
    else mat.frontIoR = mat.backIoR = 1.0f;
    V3 wO = Normalize(-rdir);
    bool specular = IsSpecular(Specularity(mat));
    bool terminal = false;   // radiance cache: ends after this vertex's NEE
    if(rc.nrc.on)
    {
        const NrcRenderState& nr = rc.nrc;
        const uint32_t view = BATCH ? S.view[i] : 0u;
        // path spread (Mueller et al. 2021, eq. 3-4): a0 at the first
        // hit, then the sum of sqrt(d^2 / (pdf |cos|)) over later hits
        float2 sp = S.spread[i];
        float tHit = __uint_as_float(hh.y);
        float cosT = MaxF(fabsf(Dot(rdir, sf.geoNormal)), 1e-6f);
        if(depth == 0u) sp = make_float2(tHit * tHit / (4.0f * kPi * cosT), 0.0f);
        else
        {
            float pdf = S.prevPdf[i];
            if(pdf > 0.0f && isfinite(pdf)) sp.y += sqrtf(tHit * tHit / (pdf * cosT));
        }
        S.spread[i] = sp;
        // query record: position in the view's scene box, outgoing direction
        // and normal as (theta, phi), roughness, albedo (RGB)
        const float4 bx = nr.box[view];
        float rec[kNrcRec];
        rec[0] = (sf.position.x - bx.x) * bx.w;
        rec[1] = (sf.position.y - bx.y) * bx.w;
        rec[2] = (sf.position.z - bx.z) * bx.w;
        auto sph = [](V3 d, float& th, float& ph)
        {
            th = fm::Acos(ClampF(d.z, -1.0f, 1.0f)) * kInvPi;
            ph = (fm::Atan2(d.y, d.x) + kPi) * (0.5f * kInvPi);
        };
        sph(wO, rec[3], rec[4]);
        sph(sf.geoNormal, rec[5], rec[6]);
        rec[7] = (mat.type == MRAY_MAT_UNREAL) ? M.roughness : (mat.type == MRAY_MAT_LAMBERT ? 1.0f : 0.0f);
        rec[8] = M.albedo[0];
        rec[9] = M.albedo[1];
        rec[10] = M.albedo[2];
        const bool training = (i & ((1u << nr.trainShift) - 1u)) == 0u;
        if(!training && depth >= nr.minDepth && sp.y * sp.y > nr.spreadC * sp.x)
        {
            // terminate into the cache: a query; the path ends here, or
            // after this vertex's NEE with neeAtEnd
            out.query = true;
            out.view = view;
            for(uint32_t k = 0; k < kNrcRec; k++) out.rec[k] = rec[k];
            if(!nr.neeAtEnd)
            {
                S.rng[i] = st;
                S.meta[i] = MPack(depth, ST_DEAD, type, dispersed, false);
                return;
            }
            terminal = true;
        }
        if(training)
        {
            uint32_t ts = i >> nr.trainShift;
            uint32_t nv = nr.trainCount[ts];
            if(nv < nr.maxVerts)
            {
                float* vx = nr.trainStore + (size_t(ts) * nr.maxVerts + nv) * kNrcVert;
                for(uint32_t k = 0; k < kNrcRec; k++) vx[k] = rec[k];
                vx[kNrcRec + 0] = thr.x; vx[kNrcRec + 1] = thr.y;
                vx[kNrcRec + 2] = thr.z; vx[kNrcRec + 3] = thr.w;
                vx[kNrcRec + 4] = rad.x; vx[kNrcRec + 5] = rad.y;
                vx[kNrcRec + 6] = rad.z; vx[kNrcRec + 7] = rad.w;
                nr.trainCount[ts] = nv + 1u;
            }
        }
    }
    if(rc.mode != MRAY_SAMPLE_PURE)
    {
        float rn[3];
        rn[0] = NextRN(st);
        rn[1] = NextRN(st);
        rn[2] = NextRN(st);
        V3 worldPos = sf.position;
        LightSample ls = SampleLight<SPECTRAL>(rc, s, rn, worldPos, waves);
        V3 sdir = Normalize(ls.position - worldPos);
        V3 npos = Nudge(ls.position, -sdir);
        float length = Length(npos - worldPos);
        bool pt, dsp;
        V4 refl = EvaluateBxDF(mat, sf, sdir, wO, pt, dsp);
        V4 t2 = thr * refl;
        float pdf;
        if(rc.mode == MRAY_SAMPLE_NEE_MIS)
            pdf = Balance2(PdfBxDF(mat, sf, sdir, wO), 1.0f, ls.pdf, 1.0f);
        else pdf = ls.pdf;
        V4 srad = DivideByPDF(t2 * ls.emission, pdf);
        if(SPECTRAL && dsp)
        {
            float first = srad.x;
            srad = Splat4(0.0f);
            srad.x = first;
        }
        if(specular) type = RT_SPECULAR;
        else
        {
            V3 nn = pt ? -sf.geoNormal : sf.geoNormal;
            V3 spos = Nudge(worldPos, nn);
            type = RT_SHADOW;
            // a visible shadow ray only adds srad: one whose contribution is
            // exactly zero (light facing away, BxDF zero below the surface)
            // cannot change the image and is not cast
            const bool zero = srad.x == 0.0f && srad.y == 0.0f && srad.z == 0.0f && srad.w == 0.0f;
            if((depth + 2u) <= rc.rr1 && !zero)
            {
                S.sO[i] = make_float4(spos.x, spos.y, spos.z, kEps);
                S.sD[i] = make_float4(sdir.x, sdir.y, sdir.z, length * (1.0f - kLargeEps));
                S.sRad[i] = make_float4(srad.x, srad.y, srad.z, srad.w);
                out.shadow = true;
            }
        }
    }
    if(terminal)
    {
        S.rng[i] = st;
        S.meta[i] = MPack(depth, ST_DEAD, type, dispersed, false);
        return;
    }
    // WorkFunction (PathTracerRendererShaders.h:201-301)
    uint32_t matRN = MatRNCount(mat.type);
    float rn[4];
    for(uint32_t k2 = 0; k2 < matRN + 1u; k2++) rn[k2] = NextRN(st);
    BxDFSample bs = SampleBxDF(mat, sf, wO, rn);
    bs.wIDir = Normalize(bs.wIDir);
    V4 t3 = thr * bs.refl;
    if(SPECTRAL && bs.dispersed) dispersed = true;
    depth += 1u;
    bool dead = (depth >= rc.rr1);
    if(!dead && depth >= rc.rr0 && !specular)
    {
        const float chInv = SPECTRAL ? 0.25f : 0.33333333f;
        float rrXi = rn[matRN];
        float prob = ClampF(Sum4(t3) * chInv, 0.1f, 1.0f);
        if(rrXi >= prob) dead = true;
        else t3 = t3 * (1.0f / prob);
    }
    type = specular ? RT_SPECULAR : RT_PATH;
    if(!dead)
    {
        V4 nt = DivideByPDF(t3, bs.pdf);
        S.thr[i] = make_float4(nt.x, nt.y, nt.z, nt.w);
        if(rc.mode == MRAY_SAMPLE_NEE_MIS || rc.nrc.on) S.prevPdf[i] = bs.pdf;
        V3 nn = bs.passedThrough ? -sf.geoNormal : sf.geoNormal;
        V3 np = Nudge(bs.wIPos, nn);
        S.rayO[i] = make_float4(np.x, np.y, np.z, kLargeEps);
        S.rayD[i] = make_float4(bs.wIDir.x, bs.wIDir.y, bs.wIDir.z, 3.402823466e+38f);
    }
    S.rng[i] = st;
    S.meta[i] = MPack(depth, dead ? ST_DEAD : ST_ALIVE, type, dispersed, false);
}

// The wave's appends after a shading step (every lane calls it): shadow
// rays into shadow-queue shard `sh` (one atomic per wave on that shard's
// counter), cache queries into the query queue.
__device__ __forceinline__ void ShadeAppend(const ShadeOut& o, uint32_t i, uint32_t sh, uint32_t cap,
                                            uint32_t* __restrict__ ctr, uint32_t* __restrict__ shadowQ,
                                            Sched* __restrict__ sched, const NrcRenderState& nr,
                                            bool queries)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    const uint64_t sm = __ballot(o.shadow);
    if(sm)
    {
        uint32_t b = 0;
        if(lane == 0) b = atomicAdd(Ctr(ctr, CTR_SHADOW + sh), uint32_t(__popcll(sm)));
        b = __shfl(b, 0);
        if(o.shadow) shadowQ[size_t(sh) * cap + b + uint32_t(__popcll(sm & lt))] = i;
    }
    if(queries)
    {
        // queries go to their view's queue: one pass per distinct view in
        // the wave (slots of a wave hold one or two views' paths)
        uint64_t pending = __ballot(o.query);
        while(pending)
        {
            const int leader = __ffsll((unsigned long long)pending) - 1;
            const uint32_t v = uint32_t(__shfl(int(o.view), leader));
            const bool mine = o.query && o.view == v;
            const uint64_t qm = __ballot(mine);
            uint32_t b = 0;
            if(lane == uint32_t(leader)) b = atomicAdd(Ctr(ctr, CTR_NRCQ + v), uint32_t(__popcll(qm)));
            b = uint32_t(__shfl(int(b), leader));
            if(mine)
            {
                const size_t q = size_t(v) * nr.queryCap + b + uint32_t(__popcll(qm & lt));
                for(uint32_t k = 0; k < kNrcRec; k++) nr.queryRec[q * kNrcRec + k] = o.rec[k];
                nr.querySlot[q] = i;
            }
            pending &= ~qm;
        }
    }
}

// Generic shading over the whole active list: static, one 64-item chunk per
// wave (shading is short per item, a fetch atomic would cost as much as the
// work it hands out); the chunk's shard picks the shadow-queue shard.  The
// grid covers N slots and has no loop: a grid-stride loop made the compiler
// hold the scene's many pointers in SGPRs for the loop's length, and spill
// them (74 SGPRs) and 40 more VGPRs (150 -> 110, 3 -> 4 waves per SIMD).
// The one-scene kernels run at 5 waves per SIMD (96 VGPRs, a few spilled in
// the spectral one); the batch kernels would spill more and stay at 4.
template<bool SPECTRAL, bool BATCH, bool INST = false>
__global__ void __launch_bounds__(kSlotTPB) __attribute__((amdgpu_waves_per_eu(BATCH ? 4 : 5, 8)))
KShade(SlotSoA S, RenderConst rc, const uint32_t* __restrict__ active,
       Sched* __restrict__ sched, uint32_t* __restrict__ shadowQ, uint32_t* __restrict__ ctr)
{
    const uint32_t n = sched->activeCount;
    const uint32_t base = (blockIdx.x * blockDim.x + threadIdx.x) & ~63u;
    if(base >= n) return;
    const uint32_t cap = ShadowShardCap(rc.N);
    const bool queries = rc.nrc.on;
    const uint32_t k = base + (threadIdx.x & 63u);
    ShadeOut o;
    o.shadow = false;
    o.query = false;
    o.view = 0;
    uint32_t i = 0;
    if(k < n)
    {
        i = active[k];
        ShadeSlot<SPECTRAL, -1, BATCH, INST>(S, rc, i, o);
    }
    ShadeAppend(o, i, ShardOfChunk(base >> 6, n), cap, ctr, shadowQ, sched, rc.nrc, queries);
}

// Material-keyed queues (replacing the reference's work-key sort and
// partition, RayPartitioner.cu:263-411): key 0 = boundary / light hits,
// 1 + material type otherwise; one atomic per key per wave, on the shard of
// the wave's active-list chunk.
__global__ void __launch_bounds__(kSlotTPB)
KClassify(SlotSoA S, RenderConst rc, const uint32_t* __restrict__ active, Sched* __restrict__ sched,
          uint32_t* __restrict__ queues, uint32_t* __restrict__ ctr)
{
    const SceneView& s = rc.scene;
    const uint32_t n = sched->activeCount;
    const uint32_t cap = ShardCap(rc.N);
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t below = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    for(uint32_t base = (blockIdx.x * blockDim.x + threadIdx.x) & ~63u; base < n;
        base += gridDim.x * blockDim.x)
    {
        const uint32_t sh = ShardOfChunk(base >> 6, n);
        uint32_t k = base + lane;
        uint32_t key = kMatQueues;           // no work
        uint32_t i = 0;
        if(k < n)
        {
            i = active[k];
            uint32_t prim = S.hit[i].x;
            if(prim == kInvalid) key = 0;
            else
            {
                uint32_t sk = s.surfKey[prim];
                key = (sk & MRAY_LIGHT_FLAG) ? 0u : 1u + s.mats[sk].type;
            }
        }
        #pragma unroll
        for(uint32_t q = 0; q < kMatQueues; q++)
        {
            uint64_t m = __ballot(key == q);
            if(!m) continue;
            uint32_t qb = 0;
            if(lane == 0) qb = atomicAdd(Ctr(ctr, CTR_QCOUNT + q * kShards + sh), uint32_t(__popcll(m)));
            qb = __shfl(qb, 0);
            if(key == q)
                queues[(size_t(q) * kShards + sh) * cap + qb + uint32_t(__popcll(m & below))] = i;
        }
    }
}

template<bool SPECTRAL, int KEY, bool BATCH>
__global__ void __launch_bounds__(kSlotTPB)
KShadeQ(SlotSoA S, RenderConst rc, const uint32_t* __restrict__ queues,
        Sched* __restrict__ sched, uint32_t* __restrict__ shadowQ, uint32_t* __restrict__ ctr)
{
    const uint32_t cap = ShardCap(rc.N), scap = ShadowShardCap(rc.N);
    const bool queries = rc.nrc.on;
    Shards g;
    LoadShards(ctr, CTR_QCOUNT + uint32_t(KEY) * kShards, g);
    const uint32_t n = g.pre[kShards];
    const uint32_t* q = queues + size_t(KEY) * kShards * cap;
    // one 64-item chunk of the queue per wave over a grid of N slots (no
    // loop: a loop keeps the scene's pointers live in SGPRs and spills them)
    const uint32_t base = (blockIdx.x * blockDim.x + threadIdx.x) & ~63u;
    if(base >= n) return;
    const uint32_t k = base + (threadIdx.x & 63u);
    ShadeOut o;
    o.shadow = false;
    o.query = false;
    o.view = 0;
    uint32_t i = 0, sh = 0;
    if(k < n)
    {
        i = q[ShardItem(g, k, cap, sh)];
        ShadeSlot<SPECTRAL, KEY, BATCH>(S, rc, i, o);
    }
    // the wave's first item names the shadow shard (a wave straddling two
    // queue shards adds at most 63 items to it: ShadowShardCap's slack)
    ShadeAppend(o, i, __builtin_amdgcn_readfirstlane(sh), scap, ctr, shadowQ, sched, rc.nrc,
                queries);
}

// ---------------------------------------------------------------------------
// Shadow rays + accumulation + film
// ---------------------------------------------------------------------------
// KCAccumulateShadowRaysPT: a shadow ray that reaches its light point adds
// its precomputed NEE radiance to the path.
template<int W, bool BATCH, int TPB = kTraceTPB, bool INST = false, bool COUNT = false>
__global__ void __launch_bounds__(TPB)
KShadow(TraceCtx ctx, SlotSoA S, const uint32_t* __restrict__ shadowQ, Sched* __restrict__ sched,
        const ViewGPU* __restrict__ views, NrcRenderState nrc, uint32_t* __restrict__ ctr,
        uint32_t N)
{
    __shared__ uint2 lstack[kLdsStack * TPB];
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t cap = ShadowShardCap(N);
    Shards g;
    LoadShards(ctr, CTR_SHADOW, g);
    const uint32_t n = g.pre[kShards];
    if(gtid == 0) sched->shadowRays += n;
    uint32_t nodes = 0, tris = 0;
    FetchLoop(Ctr(ctr, CTR_SHADOW_FETCH), n, [&](uint32_t k)
    {
        uint32_t sh;
        const uint32_t i = shadowQ[ShardItem(g, k, cap, sh)];
        const float4 o = S.sO[i], d = S.sD[i];
        HitRec h;
        if constexpr(INST)
        {
            uint32_t instId;
            h = TraceInstFlat<true, COUNT>(ctx, lstack, gtid, V3{o.x, o.y, o.z}, V3{d.x, d.y, d.z},
                                       o.w, d.w, instId, &nodes, &tris);
        }
        else
        {
            const uint32_t root = BATCH ? views[S.view[i]].root : 0u;
            h = TraceW<W, true, COUNT>(ctx, lstack, gtid, V3{o.x, o.y, o.z}, V3{d.x, d.y, d.z},
                                       o.w, d.w, root, &nodes, &tris);
        }
        if(h.prim == kInvalid)
        {
            float4 a = S.rad[i], b = S.sRad[i];
            S.rad[i] = make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w);
            // a training path's vertex recorded this bounce (its last) counts
            // its visible NEE as radiance gathered before the cache's part
            if(nrc.on && nrc.neeAtEnd && (i & ((1u << nrc.trainShift) - 1u)) == 0u)
            {
                const uint32_t ts = i >> nrc.trainShift;
                const uint32_t nv = nrc.trainCount[ts];
                if(nv > 0)
                {
                    float* vx = nrc.trainStore + (size_t(ts) * nrc.maxVerts + nv - 1u) * kNrcVert;
                    vx[kNrcRec + 4] += b.x;
                    vx[kNrcRec + 5] += b.y;
                    vx[kNrcRec + 6] += b.z;
                    vx[kNrcRec + 7] += b.w;
                }
            }
        }
    });
    if(COUNT)
    {
        unsigned long long nv = nodes, nt = tris;
        for(int o = 32; o > 0; o >>= 1)
        {
            nv += __shfl_xor(nv, o);
            nt += __shfl_xor(nt, o);
        }
        if((threadIdx.x & 63u) == 0 && (nv | nt))
        {
            atomicAdd(&sched->shadowNodeVisits, nv);
            atomicAdd(&sched->shadowTriTests, nt);
        }
    }
}

// Radiance-cache answers: a terminated path adds throughput x max(L, 0);
// in spectral mode the cache's RGB becomes the path's four wavelengths as an
// emitter's RGB radiance does (ConvRadiance).  Every view's queue in turn.
template<bool SPECTRAL>
__global__ void __launch_bounds__(256)
KNrcApply(SlotSoA S, RenderConst rc, Sched* __restrict__ sched, uint32_t* __restrict__ ctr)
{
    const NrcRenderState& nr = rc.nrc;
    const uint32_t t0 = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long total = 0;
    for(uint32_t v = 0; v < nr.views; v++)
    {
        const uint32_t n = *Ctr(ctr, CTR_NRCQ + v);
        total += n;
        for(uint32_t q = t0; q < n; q += gridDim.x * blockDim.x)
        {
            const size_t e = size_t(v) * nr.queryCap + q;
            const uint32_t i = nr.querySlot[e];
            const float* L = nr.queryOut + e * 3u;
            const V3 l = {MaxF(L[0], 0.0f), MaxF(L[1], 0.0f), MaxF(L[2], 0.0f)};
            float4 t = S.thr[i], r = S.rad[i];
            if constexpr(SPECTRAL)
            {
                const float4 w4 = S.waves[i];
                const V4 ls = ConvRadiance<true>(rc, l, V4{w4.x, w4.y, w4.z, w4.w});
                r.x += t.x * ls.x;
                r.y += t.y * ls.y;
                r.z += t.z * ls.z;
                r.w += t.w * ls.w;
            }
            else
            {
                r.x += t.x * l.x;
                r.y += t.y * l.y;
                r.z += t.z * l.z;
            }
            S.rad[i] = r;
        }
    }
    if(t0 == 0) sched->queries += total;
}

__global__ void KZeroFloats(float* __restrict__ p, size_t n)
{
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    for(; i < n; i += size_t(gridDim.x) * blockDim.x) p[i] = 0.0f;
}

__global__ void KResetSlots(uint32_t* __restrict__ meta, uint32_t* __restrict__ rng,
                            const uint32_t* __restrict__ seeds, uint32_t N,
                            Sched* __restrict__ sched, unsigned long long limit)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < N)
    {
        meta[i] = MPack(0, ST_INVALID, 0, false, false);
        rng[i] = seeds[i];
    }
    if(i == 0)
    {
        sched->counter = 0;
        sched->limit = limit;
        sched->base = 0;
        sched->deadTotal = 0;
        sched->extRays = 0;
        sched->shadowRays = 0;
        sched->nodeVisits = 0;
        sched->triTests = 0;
        sched->shadowNodeVisits = 0;
        sched->shadowTriTests = 0;
        sched->fill = 0;
        sched->activeCount = 0;
        sched->iterations = 0;
        sched->queries = 0;
        sched->startClock = 0;
        sched->viewsDone = 0;
        sched->viewBase = 0;
    }
}

} // namespace mray

// ===========================================================================
// Renderer object + C-ABI
// ===========================================================================
using namespace mray;

struct MRayRenderer
{
    const MRayAccel* accel = nullptr;
    MRayRenderParams params{};
    DeviceArena arena;
    SlotSoA S{};
    RenderConst rc{};
    Sched* dSched = nullptr;
    Sched* hSched = nullptr;      // pinned mirror
    Sched* hRing = nullptr;       // pinned scheduler snapshots, one per burst in flight
    hipEvent_t ringEv[kRingBursts] = {};
    uint32_t hostWaits = 0;       // event waits of the last RunToLimit (none drains the GPU)
    uint4* dBlockCnt = nullptr;
    uint32_t* dShadowQ = nullptr;
    uint32_t* dActive = nullptr;
    uint32_t* dSeeds = nullptr;
    uint32_t nBlocks = 0;
    uint32_t traceBlocks = 0;     // resident blocks (occupancy) per launch
    uint32_t shadeBlocks = 0;
    uint32_t shadeQBlocks[5] = {0, 0, 0, 0, 0};
    uint32_t classifyBlocks = 0;
    uint32_t* dQueues = nullptr;  // [5][kShards][ShardCap(N)] material queues (NULL: generic shading)
    uint32_t* dCtr = nullptr;     // [CTR_COUNT] work and queue counters, kCtrStride apart
    uint32_t shadowBlocks = 0;
    unsigned long long limit = 0;
    Sched last{};                 // stats at the previous mray_renderer_stats()
    hipStream_t stream = nullptr;     // the renderer's own (blocking) stream
    hipGraphExec_t graph = nullptr;   // kGraphIters captured iterations
    bool graphBroken = false;
    SpectralTables spectral;
    // measurement
    bool profiling = false;
    bool counting = false;
    std::vector<hipEvent_t> events;
    size_t evUsed = 0;
    double phaseMs[4] = {0, 0, 0, 0};
    uint64_t phaseLaunches[4] = {0, 0, 0, 0};
    // radiance cache (mray_renderer_attach_nrc)
    MRayNrcRenderParams nrcParams{};
    DeviceArena nrcArena;
    uint32_t nrcLog2G = 0;        // lanes per training path in KNrcEmit (log2)
    // frame schedule (mray_renderer_render_frame)
    unsigned long long frameLimit = 0;
    // traversal context: the accelerator's, or the merged batch scene's
    TraceCtx ctx{};
    // instanced scene (mray_renderer_create_instanced)
    const MRayInstScene* iscene = nullptr;
    // light selection by power (MRAY_RENDER_LIGHT_POWER): the tables
    uint2* dLightAlias = nullptr;
    float* dLightPmf = nullptr;
    uint32_t lightTableCount = 0;
    // batch (mray_batch_create, mray_instanced_batch_create): per-view
    // cameras and films; the scene is a merged batch scene or an instanced one
    bool batch = false;
    const MRayBatchScene* bscene = nullptr;
    std::vector<ViewGPU> hViews;
    std::vector<uint32_t> hViewScene;   // merged batch: each view's scene part
    ViewGPU* dViews = nullptr;
    unsigned long long* dViewBegin = nullptr;
    unsigned long long* dViewPix = nullptr;
    // radiance caches: one, or one per batch view (mray_renderer_attach_nrcs)
    std::vector<MRayNrc*> caches;
    float4* dNrcBox = nullptr;
    unsigned long long* dViewClock = nullptr;
    uint32_t traceTPB = kTraceTPB;
    // tiled rendering (MRAY_RENDER_TILED)
    bool tiled = false;
    uint32_t tileSize[2] = {0, 0};
    uint32_t tileCount[2] = {0, 0};
    TileGPU* dTile = nullptr;
    uint32_t* dPixRng = nullptr;      // [P] every pixel's random state
    uint32_t* dPixSeed = nullptr;     // [P] the states at the start of a render
    std::vector<unsigned long long> tilePaths;   // paths issued per tile
    std::vector<uint32_t> tileSPP;
    uint32_t curTile = 0;
    int loadedTile = -1;              // tile whose states the slots hold
    unsigned long long pathsDone = 0; // paths of every tile call so far
    MRayTileResult lastTile{};
    TileGPU hTile{};                  // source of the current tile's upload
};

#define MRAY_TRY_IT(expr)               \
    do {                                \
        int r_ = (expr);                \
        if(r_ != MRAY_OK) return r_;    \
    } while(0)

// Phase timing with HIP events on the launch stream (refill, trace, shade,
// shadow); harvested when statistics are read.
static int Mark(MRayRenderer* r, hipStream_t s)
{
    if(!r->profiling) return MRAY_OK;
    if(r->evUsed == r->events.size())
    {
        hipEvent_t e;
        MRAY_HIP_CHECK(hipEventCreate(&e));
        r->events.push_back(e);
    }
    MRAY_HIP_CHECK(hipEventRecord(r->events[r->evUsed++], s));
    return MRAY_OK;
}

static int HarvestEvents(MRayRenderer* r)
{
    // groups of 5 marks per iteration: phase k spans marks k .. k+1
    size_t groups = r->evUsed / 5;
    for(size_t g = 0; g < groups; g++)
        for(int k = 0; k < 4; k++)
        {
            float ms = 0.0f;
            MRAY_HIP_CHECK(hipEventElapsedTime(&ms, r->events[5 * g + k], r->events[5 * g + k + 1]));
            r->phaseMs[k] += ms;
            r->phaseLaunches[k] += 1;
        }
    r->evUsed = 0;
    return MRAY_OK;
}

template<bool SPECTRAL, bool BATCH>
static void LaunchShadeQueues(MRayRenderer* r, hipStream_t s)
{
    KShadeQ<SPECTRAL, 0, BATCH><<<r->shadeQBlocks[0], kSlotTPB, 0, s>>>(r->S, r->rc, r->dQueues, r->dSched, r->dShadowQ, r->dCtr);
    KShadeQ<SPECTRAL, 1, BATCH><<<r->shadeQBlocks[1], kSlotTPB, 0, s>>>(r->S, r->rc, r->dQueues, r->dSched, r->dShadowQ, r->dCtr);
    KShadeQ<SPECTRAL, 2, BATCH><<<r->shadeQBlocks[2], kSlotTPB, 0, s>>>(r->S, r->rc, r->dQueues, r->dSched, r->dShadowQ, r->dCtr);
    KShadeQ<SPECTRAL, 3, BATCH><<<r->shadeQBlocks[3], kSlotTPB, 0, s>>>(r->S, r->rc, r->dQueues, r->dSched, r->dShadowQ, r->dCtr);
    KShadeQ<SPECTRAL, 4, BATCH><<<r->shadeQBlocks[4], kSlotTPB, 0, s>>>(r->S, r->rc, r->dQueues, r->dSched, r->dShadowQ, r->dCtr);
}

template<bool SPECTRAL, bool BATCH, typename F>
static void ShadeQueueOccupancy(MRayRenderer* r, F&& resident)
{
    // a queue holds N items at most: one wave per 64-item chunk
    for(uint32_t q = 0; q < kMatQueues; q++)
        r->shadeQBlocks[q] = uint32_t((r->rc.N + kSlotTPB - 1u) / kSlotTPB);
}

// Closest-hit launch; a batch may use 128-thread traversal workgroups.
template<bool COUNT, bool BATCH, bool INST>
static void LaunchTrace(MRayRenderer* r, hipStream_t s)
{
    const TraceCtx& ctx = r->ctx;
    const uint32_t nb = r->traceBlocks;
    if constexpr(INST)
    {
        KTrace<4, COUNT, BATCH, kTraceTPB, true><<<nb, kTraceTPB, 0, s>>>(ctx, r->S, r->dActive, r->dSched,
                                                                         r->dViews, r->dCtr);
        return;
    }
    if constexpr(BATCH)
    {
        if(r->traceTPB == 128)
        {
            if(ctx.width == 8) KTrace<8, COUNT, true, 128><<<nb, 128, 0, s>>>(ctx, r->S, r->dActive, r->dSched, r->dViews, r->dCtr);
            else KTrace<4, COUNT, true, 128><<<nb, 128, 0, s>>>(ctx, r->S, r->dActive, r->dSched, r->dViews, r->dCtr);
            return;
        }
    }
    if(ctx.width == 8) KTrace<8, COUNT, BATCH><<<nb, kTraceTPB, 0, s>>>(ctx, r->S, r->dActive, r->dSched, r->dViews, r->dCtr);
    else if(r->params.flags & MRAY_RENDER_LANE_REFILL)
        KTraceRefill<COUNT, BATCH><<<nb, kTraceTPB, 0, s>>>(ctx, r->S, r->dActive, r->dSched, r->dViews, r->dCtr);
    else KTrace<4, COUNT, BATCH><<<nb, kTraceTPB, 0, s>>>(ctx, r->S, r->dActive, r->dSched, r->dViews, r->dCtr);
}

template<bool COUNT, bool BATCH, bool INST>
static void LaunchShadow(MRayRenderer* r, hipStream_t s)
{
    const TraceCtx& ctx = r->ctx;
    const uint32_t nb = r->shadowBlocks;
    if constexpr(INST)
    {
        KShadow<4, BATCH, kTraceTPB, true, COUNT><<<nb, kTraceTPB, 0, s>>>(ctx, r->S, r->dShadowQ, r->dSched,
                                                                          r->dViews, r->rc.nrc, r->dCtr, r->rc.N);
        return;
    }
    if constexpr(BATCH)
    {
        if(r->traceTPB == 128)
        {
            if(ctx.width == 8) KShadow<8, true, 128, false, COUNT><<<nb, 128, 0, s>>>(ctx, r->S, r->dShadowQ, r->dSched, r->dViews, r->rc.nrc, r->dCtr, r->rc.N);
            else KShadow<4, true, 128, false, COUNT><<<nb, 128, 0, s>>>(ctx, r->S, r->dShadowQ, r->dSched, r->dViews, r->rc.nrc, r->dCtr, r->rc.N);
            return;
        }
    }
    if(ctx.width == 8) KShadow<8, BATCH, kTraceTPB, false, COUNT><<<nb, kTraceTPB, 0, s>>>(ctx, r->S, r->dShadowQ, r->dSched, r->dViews, r->rc.nrc, r->dCtr, r->rc.N);
    else KShadow<4, BATCH, kTraceTPB, false, COUNT><<<nb, kTraceTPB, 0, s>>>(ctx, r->S, r->dShadowQ, r->dSched, r->dViews, r->rc.nrc, r->dCtr, r->rc.N);
}

template<bool BATCH, bool INST>
static int EnqueueIterationT(MRayRenderer* r, hipStream_t s)
{
    const uint32_t N = r->rc.N;
    if(r->profiling && r->evUsed + 5 > 5 * 4096)
    {
        MRAY_HIP_CHECK(hipStreamSynchronize(s));
        int rv = HarvestEvents(r);
        if(rv != MRAY_OK) return rv;
    }
    int mk;
    if((mk = Mark(r, s)) != MRAY_OK) return mk;
    KRefillCount<BATCH><<<r->nBlocks, kSlotTPB, 0, s>>>(r->S.meta, r->S.view, N, r->dBlockCnt);
    // per-view latency needs views issued in order: not in frame order
    KRefillScan<<<1, 1024, 0, s>>>(r->dBlockCnt, r->nBlocks, r->dSched, r->rc.viewBegin,
                                   r->rc.viewCount,
                                   (BATCH && !r->rc.frameMajor) ? r->dViewClock : nullptr, r->dCtr);
    if(r->rc.nrc.on)
    {
        // training pairs of finished training paths, before the refill reuses their slots
        const NrcRenderState& nr = r->rc.nrc;
        const uint64_t lanes = (uint64_t((N - 1u) >> nr.trainShift) + 1u) << r->nrcLog2G;
        if(r->rc.spectral)
            KNrcEmit<true, BATCH><<<DivUp(lanes, 256), 256, 0, s>>>(r->S, r->rc, r->dSched, r->nrcLog2G);
        else
            KNrcEmit<false, BATCH><<<DivUp(lanes, 256), 256, 0, s>>>(r->S, r->rc, r->dSched, r->nrcLog2G);
    }
    if(r->rc.spectral)
        KRefillApply<true, BATCH><<<r->nBlocks, kSlotTPB, 0, s>>>(r->S, r->rc, r->dBlockCnt, r->dSched,
                                                                  r->dActive);
    else
        KRefillApply<false, BATCH><<<r->nBlocks, kSlotTPB, 0, s>>>(r->S, r->rc, r->dBlockCnt, r->dSched,
                                                                   r->dActive);
    if((mk = Mark(r, s)) != MRAY_OK) return mk;
    if(r->counting) LaunchTrace<true, BATCH, INST>(r, s);
    else LaunchTrace<false, BATCH, INST>(r, s);
    if((mk = Mark(r, s)) != MRAY_OK) return mk;
    if(r->dQueues)
    {
        // material-keyed queues, then one specialised shading launch per key
        KClassify<<<r->classifyBlocks, kSlotTPB, 0, s>>>(r->S, r->rc, r->dActive, r->dSched,
                                                         r->dQueues, r->dCtr);
        if(r->rc.spectral) LaunchShadeQueues<true, BATCH>(r, s);
        else LaunchShadeQueues<false, BATCH>(r, s);
    }
    else if(r->rc.spectral)
        KShade<true, BATCH, INST><<<r->shadeBlocks, kSlotTPB, 0, s>>>(r->S, r->rc, r->dActive, r->dSched,
                                                                      r->dShadowQ, r->dCtr);
    else
        KShade<false, BATCH, INST><<<r->shadeBlocks, kSlotTPB, 0, s>>>(r->S, r->rc, r->dActive, r->dSched,
                                                                       r->dShadowQ, r->dCtr);
    if((mk = Mark(r, s)) != MRAY_OK) return mk;
    if(r->counting) LaunchShadow<true, BATCH, INST>(r, s);
    else LaunchShadow<false, BATCH, INST>(r, s);
    if(r->rc.nrc.on)
    {
        // radiance-cache answers for the paths that terminated this bounce:
        // each view's queue through its own cache, counts read on the device
        const NrcRenderState& nr = r->rc.nrc;
        for(uint32_t v = 0; v < nr.views; v++)
        {
            const uint32_t* dCount = r->dCtr + size_t(CTR_NRCQ + v) * kCtrStride;
            const size_t e = size_t(v) * nr.queryCap;
            MRAY_TRY_IT(mray_nrc_infer(r->caches[v], nr.queryRec + e * kNrcRec, 0, dCount,
                                       nr.queryOut + e * 3u, s));
        }
        if(r->rc.spectral) KNrcApply<true><<<r->nBlocks, 256, 0, s>>>(r->S, r->rc, r->dSched, r->dCtr);
        else KNrcApply<false><<<r->nBlocks, 256, 0, s>>>(r->S, r->rc, r->dSched, r->dCtr);
    }
    if((mk = Mark(r, s)) != MRAY_OK) return mk;
    MRAY_HIP_CHECK(hipGetLastError());
    return MRAY_OK;
}

static int EnqueueIteration(MRayRenderer* r, hipStream_t s)
{
    if(r->batch && r->iscene) return EnqueueIterationT<true, true>(r, s);
    if(r->batch) return EnqueueIterationT<true, false>(r, s);
    if(r->iscene) return EnqueueIterationT<false, true>(r, s);
    return EnqueueIterationT<false, false>(r, s);
}

__global__ void KSetLimit(Sched* __restrict__ sched, unsigned long long limit)
{
    if(threadIdx.x == 0 && blockIdx.x == 0) sched->limit = limit;
}

// A tile call's paths: the tile's global path indices [first, limit).
__global__ void KSetTileRange(Sched* __restrict__ sched, unsigned long long first,
                              unsigned long long limit)
{
    if(threadIdx.x == 0 && blockIdx.x == 0)
    {
        sched->counter = first;
        sched->limit = limit;
    }
}

// Slot j of a tiled renderer holds the random state of the tile's pixel
// (x0 + j mod w, y0 + j / w): loaded when the tile becomes current, stored
// back when another tile replaces it (the 2-D copies of
// RNGGroupIndependent::SetupRange, Random.cu:728-774).
template<bool LOAD>
__global__ void KTileRng(uint32_t* __restrict__ slotRng, uint32_t* __restrict__ pixRng, uint32_t W,
                         uint32_t x0, uint32_t y0, uint32_t w, uint32_t n)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if(j >= n) return;
    const size_t px = size_t(y0 + j / w) * W + x0 + j % w;
    if constexpr(LOAD) slotRng[j] = pixRng[px];
    else pixRng[px] = slotRng[j];
}

// Clears a w x h rectangle of the four film planes.
__global__ void KZeroRect(float* __restrict__ film, size_t P, uint32_t W, uint32_t x0, uint32_t y0,
                          uint32_t w, uint32_t n)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if(j >= n) return;
    const size_t px = size_t(y0 + j / w) * W + x0 + j % w;
    for(int c = 0; c < 4; c++) film[size_t(c) * P + px] = 0.0f;
}

static CameraGPU SubCamera(const CameraGPU& c, uint32_t sx, uint32_t sy, uint32_t cx, uint32_t cy,
                           uint32_t w, uint32_t h);

static int ResetRenderer(MRayRenderer* r)
{
    hipStream_t s = r->stream;
    size_t filmN = size_t(r->rc.P) * 4;
    KZeroFloats<<<1024, 256, 0, s>>>(r->rc.film, filmN);
    KResetSlots<<<DivUp(r->rc.N, 256), 256, 0, s>>>(r->S.meta, r->S.rng, r->dSeeds, r->rc.N,
                                                     r->dSched, r->limit);
    if(r->batch)
    {
        const size_t V = r->hViews.size();
        KZeroFloats<<<DivUp(2 * V, 256), 256, 0, s>>>(reinterpret_cast<float*>(r->dViewClock), 2 * V);
    }
    if(r->rc.nrc.on)
    {
        // training pairs counted since the reset (u64 per view)
        const size_t V = r->rc.nrc.views;
        KZeroFloats<<<1, 256, 0, s>>>(reinterpret_cast<float*>(r->rc.nrc.ringCount), 2 * V);
    }
    if(r->tiled)
    {
        // every pixel back to its first random state, no tile rendered; no
        // path is issued until render_tile sets a tile's range, and the
        // device tile is tile 0 until then
        KSetTileRange<<<1, 1, 0, s>>>(r->dSched, 0ull, 0ull);
        MRAY_HIP_CHECK(hipMemcpyAsync(r->dPixRng, r->dPixSeed, size_t(r->rc.P) * 4,
                                      hipMemcpyDeviceToDevice, s));
        r->hTile.cam = SubCamera(r->rc.cam, 0, 0, r->tileCount[0], r->tileCount[1],
                                 std::min(r->tileSize[0], r->params.width),
                                 std::min(r->tileSize[1], r->params.height));
        r->hTile.x0 = 0;
        r->hTile.y0 = 0;
        r->hTile.w = r->hTile.cam.width;
        r->hTile.P = r->hTile.cam.width * r->hTile.cam.height;
        MRAY_HIP_CHECK(hipMemcpyAsync(r->dTile, &r->hTile, sizeof(TileGPU), hipMemcpyHostToDevice, s));
        std::fill(r->tilePaths.begin(), r->tilePaths.end(), 0ull);
        std::fill(r->tileSPP.begin(), r->tileSPP.end(), 0u);
        r->curTile = 0;
        r->loadedTile = -1;
        r->pathsDone = 0;
        r->lastTile = MRayTileResult{};
    }
    MRAY_HIP_CHECK(hipGetLastError());
    MRAY_HIP_CHECK(hipStreamSynchronize(s));
    std::memset(&r->last, 0, sizeof(Sched));
    r->frameLimit = 0;
    return MRAY_OK;
}

static void BuildCamera(const MRayCamera& c, uint32_t W, uint32_t H, float filterRadius,
                        CameraGPU& out)
{
    // CameraPinhole ctor (CamerasDefault.hpp:8-36), same float operation order
    // as the CPU restatement (glibc tanf on both sides).
    V3 pos = {c.position[0], c.position[1], c.position[2]};
    V3 up = {c.up[0], c.up[1], c.up[2]};
    V3 gaze = {c.gaze[0], c.gaze[1], c.gaze[2]};
    float widthHalf = std::tan(c.fov[0] * 0.5f) * c.nearFar[0];
    float heightHalf = std::tan(c.fov[1] * 0.5f) * c.nearFar[0];
    V3 gazeDir = gaze - pos;
    V3 right = Normalize(Cross(gazeDir, up));
    up = Normalize(Cross(right, gazeDir));
    gazeDir = Normalize(Cross(up, right));
    V3 bl = ((pos - right * widthHalf) - up * heightHalf) + gazeDir * c.nearFar[0];
    out.position[0] = pos.x; out.position[1] = pos.y; out.position[2] = pos.z;
    out.right[0] = right.x; out.right[1] = right.y; out.right[2] = right.z;
    out.up[0] = up.x; out.up[1] = up.y; out.up[2] = up.z;
    out.bottomLeft[0] = bl.x; out.bottomLeft[1] = bl.y; out.bottomLeft[2] = bl.z;
    out.planeSize[0] = widthHalf * 2.0f;
    out.planeSize[1] = heightHalf * 2.0f;
    out.nearFar[0] = c.nearFar[0];
    out.nearFar[1] = c.nearFar[1];
    out.sigma = filterRadius * 0.285714f;
    out.width = W;
    out.height = H;
}

static int SizeLaunches(MRayRenderer* r);

// RNGGroupIndependent (Random.cu:661-703): seed32 = hi ^ lo of the 64-bit
// seed, one std::mt19937 draw per pixel through PCG32's GenerateState.
static std::vector<uint32_t> PixelSeeds(uint64_t seed, uint32_t n)
{
    std::vector<uint32_t> seeds(n);
    uint32_t seed32 = uint32_t((seed >> 32) ^ (seed & 0xFFFFFFFFull));
    std::mt19937 mt(seed32);
    for(uint32_t i = 0; i < n; i++)
    {
        uint32_t x = uint32_t(mt());
        uint32_t st = 0u * 747796405u + 2891336453u;
        st += x;
        st = st * 747796405u + 2891336453u;
        seeds[i] = st;
    }
    return seeds;
}

// ImageTiler::FindOptimumTileSize (RenderImage.cpp:20-52): a tile of about
// `hint` pixels with the image's aspect ratio, then per axis the count of
// tiles nearest to the image over that size, and the tile size that covers
// the axis with that count.  A hint below one pixel per axis is taken as 1.
static void FindTileSize(uint32_t W, uint32_t H, uint32_t hint, uint32_t out[2])
{
    const float aspect = float(W) / float(H);
    const float factor = std::sqrt(float(hint) / aspect);
    const uint32_t want[2] = {std::max(1u, uint32_t(std::round(aspect * factor))),
                              std::max(1u, uint32_t(std::round(factor)))};
    const uint32_t fb[2] = {W, H};
    for(int i = 0; i < 2; i++)
    {
        if(fb[i] < want[i])
        {
            out[i] = fb[i];
            continue;
        }
        const uint32_t count = uint32_t(std::round(float(fb[i]) / float(want[i])));
        uint32_t residual = fb[i] % count;
        residual = (residual + count - 1u) / count;
        out[i] = fb[i] / count + residual;
    }
}

// CameraPinhole::GenerateSubCamera (CamerasDefault.hpp:226-247): stratum
// (sx, sy) of a (cx, cy) split of the image plane, seen at w x h pixels.
static CameraGPU SubCamera(const CameraGPU& c, uint32_t sx, uint32_t sy, uint32_t cx, uint32_t cy,
                           uint32_t w, uint32_t h)
{
    CameraGPU t = c;
    const float dx = c.planeSize[0] / float(cx), dy = c.planeSize[1] / float(cy);
    const float rx = float(sx) * dx, ry = float(sy) * dy;
    for(int k = 0; k < 3; k++) t.bottomLeft[k] = c.bottomLeft[k] + ((rx * c.right[k]) + (ry * c.up[k]));
    t.planeSize[0] = dx;
    t.planeSize[1] = dy;
    t.width = w;
    t.height = h;
    return t;
}

// Path-slot state, queues, seeds and launch sizes, shared by the one-scene
// and batch renderers (r->params, r->rc.N, r->ctx and r->batch set).
static int InitSlots(MRayRenderer* r, size_t filmPixels)
{
    const MRayRenderParams* p = &r->params;
    const uint32_t N = r->rc.N;
    if(p->spectral)
    {
        MRAY_TRY(GetSpectralTables(r->spectral));
        r->rc.spec = r->spectral.view;
    }
    DeviceArena& ar = r->arena;
    auto A = [&](void** ptr, size_t bytes) { return DevAlloc(ar, ptr, bytes); };
    MRAY_TRY(A((void**)&r->S.rayO, size_t(N) * 16));
    MRAY_TRY(A((void**)&r->S.rayD, size_t(N) * 16));
    MRAY_TRY(A((void**)&r->S.hit, size_t(N) * 16));
    MRAY_TRY(A((void**)&r->S.thr, size_t(N) * 16));
    MRAY_TRY(A((void**)&r->S.rad, size_t(N) * 16));
    MRAY_TRY(A((void**)&r->S.meta, size_t(N) * 4));
    MRAY_TRY(A((void**)&r->S.prevPdf, size_t(N) * 4));
    MRAY_TRY(A((void**)&r->S.pixel, size_t(N) * 4));
    MRAY_TRY(A((void**)&r->S.filterW, size_t(N) * 4));
    MRAY_TRY(A((void**)&r->S.rng, size_t(N) * 4));
    MRAY_TRY(A((void**)&r->S.sO, size_t(N) * 16));
    MRAY_TRY(A((void**)&r->S.sD, size_t(N) * 16));
    MRAY_TRY(A((void**)&r->S.sRad, size_t(N) * 16));
    if(p->spectral)
    {
        MRAY_TRY(A((void**)&r->S.waves, size_t(N) * 16));
        MRAY_TRY(A((void**)&r->S.wpdf, size_t(N) * 16));
    }
    if(r->batch) MRAY_TRY(A((void**)&r->S.view, size_t(N) * 4));
    if(r->iscene) MRAY_TRY(A((void**)&r->S.inst, size_t(N) * 4));
    MRAY_TRY(A((void**)&r->rc.film, filmPixels * 4 * sizeof(float)));
    r->nBlocks = DivUp(N, kSlotTPB);
    MRAY_TRY(A((void**)&r->dBlockCnt, size_t(r->nBlocks) * sizeof(uint4)));
    MRAY_TRY(A((void**)&r->dShadowQ, size_t(kShards) * ShadowShardCap(N) * 4));
    MRAY_TRY(A((void**)&r->dCtr, size_t(CTR_COUNT) * kCtrStride * 4));
    MRAY_HIP_CHECK(hipMemset(r->dCtr, 0, size_t(CTR_COUNT) * kCtrStride * 4));
    MRAY_TRY(A((void**)&r->dActive, size_t(N) * 4));
    MRAY_TRY(A((void**)&r->dSched, sizeof(Sched)));
    MRAY_TRY(A((void**)&r->dSeeds, size_t(N) * 4));
    MRAY_HIP_CHECK(hipHostMalloc((void**)&r->hSched, sizeof(Sched), hipHostMallocDefault));
    // RNGGroupIndependent (Random.cu:661-703): seed32 = hi ^ lo of the 64-bit
    // seed, one std::mt19937 draw per pixel through GenerateState.
    {
        std::vector<uint32_t> seeds = PixelSeeds(p->seed, N);
        MRAY_HIP_CHECK(hipMemcpy(r->dSeeds, seeds.data(), size_t(N) * 4, hipMemcpyHostToDevice));
    }
    return SizeLaunches(r);
}

// Persistent launches are sized to what is resident at once (occupancy x
// CUs): a block that is not resident would start only after a resident one
// finished, stretching the tail.  The traversal grid never exceeds the
// spill-stack allocation (ctx.gridThreads).
static int SizeLaunches(MRayRenderer* r)
{
    const MRayRenderParams* p = &r->params;
    const uint32_t N = r->rc.N;
    int dev = 0;
    MRAY_HIP_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    MRAY_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
    const uint32_t cus = uint32_t(std::max(1, prop.multiProcessorCount));
    auto resident = [&](auto kernel, uint32_t tpb, uint32_t cap) -> int
    {
        int nb = 0;
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, int(tpb), 0);
        if(e != hipSuccess || nb < 1) nb = 1;
        return int(std::min<uint64_t>(uint64_t(cus) * uint32_t(nb), cap));
    };
    const uint32_t tpb = r->traceTPB;
    const uint32_t maxTraceBlocks = r->ctx.gridThreads / tpb;
    const bool w8 = r->ctx.width == 8;
    const bool refill = (p->flags & MRAY_RENDER_LANE_REFILL) != 0u;
    if(r->iscene)
    {
        r->traceBlocks = uint32_t(resident(KTrace<4, false, false, kTraceTPB, true>, tpb, maxTraceBlocks));
        r->shadowBlocks = uint32_t(resident(KShadow<4, false, kTraceTPB, true>, tpb, maxTraceBlocks));
    }
    else if(!r->batch)
    {
        r->traceBlocks = uint32_t(w8 ? resident(KTrace<8, false, false>, tpb, maxTraceBlocks)
                                : refill ? resident(KTraceRefill<false, false>, tpb, maxTraceBlocks)
                                         : resident(KTrace<4, false, false>, tpb, maxTraceBlocks));
        r->shadowBlocks = uint32_t(w8 ? resident(KShadow<8, false>, tpb, maxTraceBlocks)
                                      : resident(KShadow<4, false>, tpb, maxTraceBlocks));
    }
    else if(tpb == 128)
    {
        r->traceBlocks = uint32_t(w8 ? resident(KTrace<8, false, true, 128>, tpb, maxTraceBlocks)
                                     : resident(KTrace<4, false, true, 128>, tpb, maxTraceBlocks));
        r->shadowBlocks = uint32_t(w8 ? resident(KShadow<8, true, 128>, tpb, maxTraceBlocks)
                                      : resident(KShadow<4, true, 128>, tpb, maxTraceBlocks));
    }
    else
    {
        r->traceBlocks = uint32_t(w8 ? resident(KTrace<8, false, true>, tpb, maxTraceBlocks)
                                : refill ? resident(KTraceRefill<false, true>, tpb, maxTraceBlocks)
                                         : resident(KTrace<4, false, true>, tpb, maxTraceBlocks));
        r->shadowBlocks = uint32_t(w8 ? resident(KShadow<8, true>, tpb, maxTraceBlocks)
                                      : resident(KShadow<4, true>, tpb, maxTraceBlocks));
    }
    if(p->flags & MRAY_RENDER_MATERIAL_QUEUES)
    {
        if(!r->dQueues)
            MRAY_TRY(DevAlloc(r->arena, (void**)&r->dQueues, size_t(kMatQueues) * kShards * ShardCap(N) * 4));
        r->classifyBlocks = uint32_t(resident(KClassify, kSlotTPB, 1u << 20));
        if(r->bscene)
        {
            if(p->spectral) ShadeQueueOccupancy<true, true>(r, resident);
            else ShadeQueueOccupancy<false, true>(r, resident);
        }
        else if(p->spectral) ShadeQueueOccupancy<true, false>(r, resident);
        else ShadeQueueOccupancy<false, false>(r, resident);
    }
    // one wave per 64-item chunk of the active list, N slots at most
    r->shadeBlocks = uint32_t((r->rc.N + kSlotTPB - 1u) / kSlotTPB);
    return MRAY_OK;
}

static int SetupLightPower(MRayRenderer* r, const std::vector<uint32_t>& offs,
                           const std::vector<uint32_t>& counts, uint32_t total);

// Spectral renderers: convert the scene's constant colours (material
// albedos, light radiances) to sigmoid coefficients once.
static int SetupSpectralCoeffs(MRayRenderer* r, uint32_t nm, uint32_t nl)
{
    r->rc.scene.matCoeff = nullptr;
    r->rc.scene.lightCoeff = nullptr;
    if(!r->rc.spectral) return MRAY_OK;
    float4 *dm = nullptr, *dl = nullptr;
    MRAY_TRY(DevAlloc(r->arena, (void**)&dm, std::max<size_t>(nm, 1) * sizeof(float4)));
    MRAY_TRY(DevAlloc(r->arena, (void**)&dl, std::max<size_t>(nl, 1) * sizeof(float4)));
    const uint32_t n = std::max(nm, nl);
    if(n)
        KSpectralCoeffs<<<DivUp(n, 256), 256, 0, r->stream>>>(r->rc.spec, r->rc.scene.mats, nm,
                                                              r->rc.scene.lightRad, nl, dm, dl);
    MRAY_HIP_CHECK(hipGetLastError());
    MRAY_HIP_CHECK(hipStreamSynchronize(r->stream));
    r->rc.scene.matCoeff = dm;
    r->rc.scene.lightCoeff = dl;
    return MRAY_OK;
}

static int CreateRenderer(MRayRenderer* r, const MRayAccel* a, const MRaySceneDesc* scene,
                          const MRayRenderParams* p)
{
    r->accel = a;
    r->ctx = a->ctx;
    r->params = *p;
    // a blocking stream of default priority: orders with the null stream
    MRAY_HIP_CHECK(hipStreamCreate(&r->stream));
    const uint64_t P = uint64_t(p->width) * p->height;
    uint32_t N = p->maxRayCount ? p->maxRayCount : uint32_t(P);
    if(N > P) N = uint32_t(P);
    if(p->flags & MRAY_RENDER_TILED)
    {
        // one slot per pixel of a (covering) tile
        r->tiled = true;
        FindTileSize(p->width, p->height, p->maxRayCount ? p->maxRayCount : (1u << 21), r->tileSize);
        r->tileCount[0] = DivUp(p->width, r->tileSize[0]);
        r->tileCount[1] = DivUp(p->height, r->tileSize[1]);
        N = r->tileSize[0] * r->tileSize[1];
    }
    r->rc.N = N;
    r->rc.P = uint32_t(P);
    r->rc.W = p->width;
    r->rc.rr0 = p->rrRange[0];
    r->rc.rr1 = p->rrRange[1];
    r->rc.mode = p->sampleMode;
    r->rc.spectral = p->spectral;
    r->rc.scene = a->view;
    float fr = (p->filterRadius > 0.0f) ? p->filterRadius : 1.0f;
    BuildCamera(scene ? scene->camera : a->camera, p->width, p->height, fr, r->rc.cam);
    r->limit = (unsigned long long)p->totalSPP * P;
    MRAY_TRY(SetupLightPower(r, {0u}, {a->view.lightCount}, a->view.lightCount));
    MRAY_TRY(InitSlots(r, size_t(P)));
    MRAY_TRY(SetupSpectralCoeffs(r, uint32_t(a->hMats.size()), a->view.lightCount));
    if(r->tiled)
    {
        const uint32_t T = r->tileCount[0] * r->tileCount[1];
        MRAY_TRY(DevAlloc(r->arena, (void**)&r->dTile, sizeof(TileGPU)));
        MRAY_TRY(DevAlloc(r->arena, (void**)&r->dPixRng, size_t(P) * 4));
        MRAY_TRY(DevAlloc(r->arena, (void**)&r->dPixSeed, size_t(P) * 4));
        std::vector<uint32_t> seeds = PixelSeeds(p->seed, uint32_t(P));
        MRAY_HIP_CHECK(hipMemcpy(r->dPixSeed, seeds.data(), size_t(P) * 4, hipMemcpyHostToDevice));
        r->tilePaths.assign(T, 0ull);
        r->tileSPP.assign(T, 0u);
        r->rc.tile = r->dTile;
    }
    return ResetRenderer(r);
}

// MRAY_RENDER_LIGHT_POWER: alias tables over the given light ranges of the
// renderer's scene view (one range per scene of a batch).
static int SetupLightPower(MRayRenderer* r, const std::vector<uint32_t>& offs,
                           const std::vector<uint32_t>& counts, uint32_t total)
{
    r->rc.scene.lightAlias = nullptr;
    r->rc.scene.lightPmf = nullptr;
    if(!(r->params.flags & MRAY_RENDER_LIGHT_POWER)) return MRAY_OK;
    MRAY_TRY(BuildLightPower(r->arena, r->rc.scene, offs, counts, total, &r->dLightAlias,
                             &r->dLightPmf));
    r->lightTableCount = total;
    r->rc.scene.lightAlias = r->dLightAlias;
    r->rc.scene.lightPmf = r->dLightPmf;
    return MRAY_OK;
}

static int CreateInstanced(MRayRenderer* r, const MRayInstScene* sc, const MRayRenderParams* p)
{
    if(p->flags & MRAY_RENDER_TILED)
    {
        SetError("renderer_create_instanced: tiled rendering takes one flattened scene");
        return MRAY_ERR_UNSUPPORTED;
    }
    r->iscene = sc;
    r->ctx = sc->ctx;
    r->params = *p;
    MRAY_HIP_CHECK(hipStreamCreate(&r->stream));
    const uint64_t P = uint64_t(p->width) * p->height;
    uint32_t N = p->maxRayCount ? p->maxRayCount : uint32_t(P);
    if(N > P) N = uint32_t(P);
    r->rc.N = N;
    r->rc.P = uint32_t(P);
    r->rc.W = p->width;
    r->rc.rr0 = p->rrRange[0];
    r->rc.rr1 = p->rrRange[1];
    r->rc.mode = p->sampleMode;
    r->rc.spectral = p->spectral;
    r->rc.scene = sc->view;
    r->rc.inst = sc->ctx.inst;
    const float fr = (p->filterRadius > 0.0f) ? p->filterRadius : 1.0f;
    BuildCamera(sc->camera, p->width, p->height, fr, r->rc.cam);
    r->limit = (unsigned long long)p->totalSPP * P;
    MRAY_TRY(SetupLightPower(r, {0u}, {sc->view.lightCount}, sc->view.lightCount));
    MRAY_TRY(InitSlots(r, size_t(P)));
    MRAY_TRY(SetupSpectralCoeffs(r, sc->meshes.mats, sc->view.lightCount));
    return ResetRenderer(r);
}

static int CheckSampling(const MRayRenderParams& p, const char* who)
{
    if(p.sampleMode > MRAY_SAMPLE_NEE_MIS)
    {
        // PathTracerRendererT::PushAttribute throws on unknown values
        SetError(std::string(who) + ": unknown sampleMode");
        return MRAY_ERR_INVALID;
    }
    if(p.rrRange[1] == 0 || p.rrRange[0] > p.rrRange[1])
    {
        SetError(std::string(who) + ": rrRange must satisfy 0 < max and start <= max");
        return MRAY_ERR_INVALID;
    }
    return MRAY_OK;
}

// The one scene an instanced batch's views look at, as a batch part.
static MRayBatchScene::Part InstancedPart(const MRayInstScene* is)
{
    MRayBatchScene::Part part;
    part.lightOff = 0;
    part.lightCount = is->view.lightCount;
    part.boundaryType = is->view.boundaryType;
    for(int c = 0; c < 3; c++) part.boundaryRad[c] = is->view.boundaryRad[c];
    part.sceneDiameter = is->view.sceneDiameter;
    part.camera = is->camera;
    part.env = is->view;
    return part;
}

// Views over a merged batch scene (b) or over one instanced scene (is).
static int CreateBatch(MRayRenderer* r, const MRayBatchScene* b, const MRayInstScene* is,
                       const MRayBatchView* views, uint32_t viewCount, const MRayRenderParams* p)
{
    if(p->flags & MRAY_RENDER_TILED)
    {
        SetError("batch_create: a batch renders whole views; tiling is for one scene");
        return MRAY_ERR_UNSUPPORTED;
    }
    r->batch = true;
    r->bscene = b;
    r->iscene = is;
    r->ctx = is ? is->ctx : b->ctx;
    r->params = *p;
    MRAY_HIP_CHECK(hipStreamCreate(&r->stream));
    r->rc.rr0 = p->rrRange[0];
    r->rc.rr1 = p->rrRange[1];
    r->rc.mode = p->sampleMode;
    r->rc.spectral = p->spectral;
    r->rc.scene = is ? is->view : b->view;
    if(is) r->rc.inst = is->ctx.inst;
    const float fr = (p->filterRadius > 0.0f) ? p->filterRadius : 1.0f;
    uint64_t pixels = 0, paths = 0;
    const MRayBatchScene::Part instPart = is ? InstancedPart(is) : MRayBatchScene::Part{};
    r->hViews.resize(viewCount);
    for(uint32_t k = 0; k < viewCount; k++)
    {
        const MRayBatchView& v = views[k];
        const MRayBatchScene::Part& part = is ? instPart : b->parts[v.scene];
        r->hViewScene.push_back(is ? 0u : v.scene);
        ViewGPU& g = r->hViews[k];
        g = ViewGPU{};
        BuildCamera(v.useCamera ? v.camera : part.camera, v.width, v.height, fr, g.cam);
        g.P = v.width * v.height;
        g.W = v.width;
        g.film = reinterpret_cast<float*>(uintptr_t(pixels * 4 * sizeof(float)));  // offset, fixed below
        g.pathBegin = paths;
        g.pathEnd = paths + uint64_t(g.P) * v.spp;
        g.root = part.root;
        g.lightOff = part.lightOff;
        g.lightCount = part.lightCount;
        g.boundaryType = part.boundaryType;
        for(int c = 0; c < 3; c++) g.boundaryRad[c] = part.boundaryRad[c];
        g.sceneDiameter = part.sceneDiameter;
        g.envTex = part.env.envTex;
        g.envCdfX = part.env.envCdfX;
        g.envCdfY = part.env.envCdfY;
        g.envW = part.env.envW;
        g.envH = part.env.envH;
        pixels += g.P;
        paths = g.pathEnd;
    }
    uint32_t N = p->maxRayCount ? p->maxRayCount : uint32_t(std::min<uint64_t>(pixels, 1u << 21));
    if(N > pixels) N = uint32_t(pixels);
    r->rc.N = N;
    r->rc.P = uint32_t(pixels);
    r->rc.W = 0;
    r->limit = paths;
    MRAY_TRY(InitSlots(r, size_t(pixels)));
    for(ViewGPU& g : r->hViews)
        g.film = r->rc.film + (reinterpret_cast<uintptr_t>(g.film) / sizeof(float));
    DeviceArena& ar = r->arena;
    MRAY_TRY(DevAlloc(ar, (void**)&r->dViews, viewCount * sizeof(ViewGPU)));
    MRAY_TRY(DevAlloc(ar, (void**)&r->dViewBegin, (viewCount + 1) * sizeof(unsigned long long)));
    MRAY_TRY(DevAlloc(ar, (void**)&r->dViewClock, viewCount * sizeof(unsigned long long)));
    MRAY_HIP_CHECK(hipMemcpy(r->dViews, r->hViews.data(), viewCount * sizeof(ViewGPU),
                             hipMemcpyHostToDevice));
    r->rc.views = r->dViews;
    r->rc.viewCount = viewCount;
    {
        std::vector<unsigned long long> begin(viewCount + 1);
        for(uint32_t k = 0; k < viewCount; k++) begin[k] = r->hViews[k].pathBegin;
        begin[viewCount] = r->hViews[viewCount - 1].pathEnd;
        MRAY_HIP_CHECK(hipMemcpy(r->dViewBegin, begin.data(), begin.size() * sizeof(unsigned long long),
                                 hipMemcpyHostToDevice));
    }
    r->rc.viewBegin = r->dViewBegin;
    r->rc.viewClock = r->dViewClock;
    {
        std::vector<unsigned long long> pix(viewCount + 1, 0ull);
        for(uint32_t k = 0; k < viewCount; k++) pix[k + 1] = pix[k] + r->hViews[k].P;
        MRAY_TRY(DevAlloc(ar, (void**)&r->dViewPix, (viewCount + 1) * sizeof(unsigned long long)));
        MRAY_HIP_CHECK(hipMemcpy(r->dViewPix, pix.data(), pix.size() * sizeof(unsigned long long),
                                 hipMemcpyHostToDevice));
        r->rc.viewPix = r->dViewPix;
        r->rc.pixTotal = pix[viewCount];
        r->rc.frameMajor = 0;
    }
    if(is) MRAY_TRY(SetupLightPower(r, {0u}, {is->view.lightCount}, is->view.lightCount));
    else
    {
        std::vector<uint32_t> offs, counts;
        for(const MRayBatchScene::Part& part : b->parts)
        {
            offs.push_back(part.lightOff);
            counts.push_back(part.lightCount);
        }
        MRAY_TRY(SetupLightPower(r, offs, counts, b->lights));
    }
    MRAY_TRY(SetupSpectralCoeffs(r, is ? is->meshes.mats : b->mats, is ? is->view.lightCount : b->lights));
    return ResetRenderer(r);
}

static void FreeRenderer(MRayRenderer* r)
{
    if(r->graph) (void)hipGraphExecDestroy(r->graph);
    if(r->stream) (void)hipStreamDestroy(r->stream);
    for(hipEvent_t e : r->events) (void)hipEventDestroy(e);
    if(r->hSched) (void)hipHostFree(r->hSched);
    if(r->hRing) (void)hipHostFree(r->hRing);
    for(hipEvent_t e : r->ringEv)
        if(e) (void)hipEventDestroy(e);
    r->nrcArena.Release();
    r->arena.Release();
    delete r;
}

extern "C" {

int mray_renderer_create(const MRayAccel* accel, const MRaySceneDesc* hScene,
                         const MRayRenderParams* hParams, MRayRenderer** out)
{
    if(!accel || !hParams || !out)
    {
        SetError("renderer_create: NULL argument");
        return MRAY_ERR_INVALID;
    }
    *out = nullptr;
    const MRayRenderParams& p = *hParams;
    if(p.width == 0 || p.height == 0 || p.totalSPP == 0)
    {
        SetError("renderer_create: width, height and totalSPP must be non-zero");
        return MRAY_ERR_INVALID;
    }
    if(uint64_t(p.width) * p.height > (1ull << 31))
    {
        SetError("renderer_create: image too large");
        return MRAY_ERR_INVALID;
    }
    MRAY_TRY(CheckSampling(p, "renderer_create"));
    MRayRenderer* r = new MRayRenderer();
    int rv = CreateRenderer(r, accel, hScene, hParams);
    if(rv != MRAY_OK)
    {
        (void)hipDeviceSynchronize();
        FreeRenderer(r);
        return rv;
    }
    *out = r;
    return MRAY_OK;
}

int mray_renderer_create_instanced(const MRayInstScene* scene, const MRayRenderParams* hParams,
                                   MRayRenderer** out)
{
    if(!scene || !hParams || !out)
    {
        SetError("renderer_create_instanced: NULL argument");
        return MRAY_ERR_INVALID;
    }
    *out = nullptr;
    const MRayRenderParams& p = *hParams;
    if(p.width == 0 || p.height == 0 || p.totalSPP == 0 || uint64_t(p.width) * p.height > (1ull << 31))
    {
        SetError("renderer_create_instanced: bad image size or totalSPP");
        return MRAY_ERR_INVALID;
    }
    if(p.flags & MRAY_RENDER_MATERIAL_QUEUES)
    {
        SetError("renderer_create_instanced: material queues are not available for instanced scenes");
        return MRAY_ERR_UNSUPPORTED;
    }
    MRAY_TRY(CheckSampling(p, "renderer_create_instanced"));
    MRayRenderer* r = new MRayRenderer();
    int rv = CreateInstanced(r, scene, hParams);
    if(rv != MRAY_OK)
    {
        (void)hipDeviceSynchronize();
        FreeRenderer(r);
        return rv;
    }
    *out = r;
    return MRAY_OK;
}

int mray_renderer_destroy(MRayRenderer* r)
{
    if(!r) return MRAY_OK;
    (void)hipDeviceSynchronize();
    FreeRenderer(r);
    return MRAY_OK;
}

static int SetFrameOrder(MRayRenderer* r, bool on);

int mray_renderer_reset(MRayRenderer* r)
{
    if(!r) { SetError("renderer is NULL"); return MRAY_ERR_INVALID; }
    // a batch renderer returns to view order (every view's spp, views in turn)
    if(r->batch) MRAY_TRY(SetFrameOrder(r, false));
    return ResetRenderer(r);
}

int mray_renderer_iterate(MRayRenderer* r, uint32_t iterations, void* stream)
{
    if(!r) { SetError("renderer is NULL"); return MRAY_ERR_INVALID; }
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : r->stream;
    // Bursts of kGraphIters bounces replay one captured hipGraph on the
    // renderer's stream (every kernel reads its counts from device memory, so
    // the same graph serves every burst).  Measurement, a caller's stream or
    // MRAY_RENDER_NO_GRAPH run the launches directly.
    const bool useGraph = !stream && !r->profiling && !r->counting && !r->graphBroken &&
                          !(r->params.flags & MRAY_RENDER_NO_GRAPH);
    if(useGraph && iterations >= kGraphIters)
    {
        if(!r->graph)
        {
            hipGraph_t g = nullptr;
            MRAY_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            int rv = MRAY_OK;
            for(uint32_t k = 0; k < kGraphIters && rv == MRAY_OK; k++) rv = EnqueueIteration(r, s);
            hipError_t ec = hipStreamEndCapture(s, &g);
            if(rv != MRAY_OK || ec != hipSuccess || !g ||
               hipGraphInstantiate(&r->graph, g, nullptr, nullptr, 0) != hipSuccess)
            {
                // fall back to direct launches for this renderer
                (void)hipGetLastError();
                r->graph = nullptr;
                r->graphBroken = true;
            }
            if(g) (void)hipGraphDestroy(g);
        }
        while(r->graph && iterations >= kGraphIters)
        {
            MRAY_HIP_CHECK(hipGraphLaunch(r->graph, s));
            iterations -= kGraphIters;
        }
    }
    for(uint32_t k = 0; k < iterations; k++) MRAY_TRY(EnqueueIteration(r, s));
    return MRAY_OK;
}

// Queue bursts of kGraphIters bounces until deadTotal reaches `limit`.
// After each burst the scheduler record is copied asynchronously into a
// pinned ring slot; the host reads the snapshot of the burst kLookahead
// bursts back, so the GPU never waits on the host and the device drains once,
// when the caller reads the statistics.  Bounces queued after the frame has
// finished find no work and are cheap (every kernel reads its counts from
// device memory).
static int RunToLimit(MRayRenderer* r, unsigned long long limit, const char* who)
{
    hipStream_t s = r->stream;
    if(!r->hRing)
    {
        MRAY_HIP_CHECK(hipHostMalloc((void**)&r->hRing, sizeof(Sched) * kRingBursts,
                                     hipHostMallocDefault));
        for(hipEvent_t& e : r->ringEv)
            MRAY_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    r->hostWaits = 0;
    unsigned long long prevDead = 0, prevExt = 0;
    bool havePrev = false;
    for(uint64_t b = 0;; b++)
    {
        MRAY_TRY(mray_renderer_iterate(r, kGraphIters, nullptr));
        const uint32_t slot = uint32_t(b % kRingBursts);
        MRAY_HIP_CHECK(hipMemcpyAsync(&r->hRing[slot], r->dSched, sizeof(Sched),
                                      hipMemcpyDeviceToHost, s));
        MRAY_HIP_CHECK(hipEventRecord(r->ringEv[slot], s));
        if(b < kLookahead) continue;
        const uint32_t old = uint32_t((b - kLookahead) % kRingBursts);
        MRAY_HIP_CHECK(hipEventSynchronize(r->ringEv[old]));
        r->hostWaits++;
        const Sched& c = r->hRing[old];
        if(c.deadTotal >= limit) break;
        if(havePrev && c.deadTotal == prevDead && c.extRays == prevExt)
        {
            SetError(std::string(who) + ": no progress");
            return MRAY_ERR_INVALID;
        }
        prevDead = c.deadTotal;
        prevExt = c.extRays;
        havePrev = true;
    }
    return MRAY_OK;
}

int mray_renderer_stats(MRayRenderer* r, MRayRenderStats* out)
{
    if(!r || !out) { SetError("renderer_stats: NULL argument"); return MRAY_ERR_INVALID; }
    MRAY_HIP_CHECK(hipDeviceSynchronize());
    MRAY_HIP_CHECK(hipMemcpy(r->hSched, r->dSched, sizeof(Sched), hipMemcpyDeviceToHost));
    const Sched& c = *r->hSched;
    out->deadPaths = c.deadTotal - r->last.deadTotal;
    out->totalDeadPaths = c.deadTotal;
    out->extensionRays = c.extRays - r->last.extRays;
    out->shadowRays = c.shadowRays - r->last.shadowRays;
    out->iterations = c.iterations - r->last.iterations;
    out->finished = (c.deadTotal >= r->limit) ? 1u : 0u;
    out->nodeVisits = c.nodeVisits - r->last.nodeVisits;
    out->triTests = c.triTests - r->last.triTests;
    out->shadowNodeVisits = c.shadowNodeVisits - r->last.shadowNodeVisits;
    out->shadowTriTests = c.shadowTriTests - r->last.shadowTriTests;
    r->last = c;
    if(r->profiling) return HarvestEvents(r);
    return MRAY_OK;
}

int mray_renderer_render_all(MRayRenderer* r, MRayRenderStats* hStats)
{
    if(!r) { SetError("renderer is NULL"); return MRAY_ERR_INVALID; }
    if(r->rc.frameMajor)
    {
        SetError("render_all: this batch renderer renders by frames; reset it first");
        return MRAY_ERR_INVALID;
    }
    if(r->tiled)
    {
        // every tile up to totalSPP, in tile order from the current one
        uint32_t waits = 0;
        MRayRenderStats sum{};
        for(uint32_t k = 0; k < r->tileCount[0] * r->tileCount[1]; k++)
        {
            const uint32_t have = r->tileSPP[r->curTile];
            MRayRenderStats st{};
            MRAY_TRY(mray_renderer_render_tile(r, have < r->params.totalSPP ? r->params.totalSPP - have : 0u,
                                               nullptr, &st));
            waits += st.hostWaits;
            sum.deadPaths += st.deadPaths;
            sum.extensionRays += st.extensionRays;
            sum.shadowRays += st.shadowRays;
            sum.iterations += st.iterations;
            sum.nodeVisits += st.nodeVisits;
            sum.triTests += st.triTests;
            sum.shadowNodeVisits += st.shadowNodeVisits;
            sum.shadowTriTests += st.shadowTriTests;
            sum.totalDeadPaths = st.totalDeadPaths;
        }
        sum.finished = 1;
        sum.hostWaits = waits;
        sum.deviceDrains = r->tileCount[0] * r->tileCount[1];
        if(hStats) *hStats = sum;
        return MRAY_OK;
    }
    MRAY_TRY(RunToLimit(r, r->limit, "render_all"));
    MRayRenderStats total{};
    MRAY_TRY(mray_renderer_stats(r, &total));
    total.finished = 1;
    total.hostWaits = r->hostWaits;
    total.deviceDrains = 1;
    if(hStats) *hStats = total;
    return MRAY_OK;
}

int mray_renderer_read_film(MRayRenderer* r, float* hOut)
{
    if(!r || !hOut) { SetError("read_film: NULL argument"); return MRAY_ERR_INVALID; }
    MRAY_HIP_CHECK(hipDeviceSynchronize());
    MRAY_HIP_CHECK(hipMemcpy(hOut, r->rc.film, size_t(r->rc.P) * 4 * sizeof(float),
                             hipMemcpyDeviceToHost));
    return MRAY_OK;
}

int mray_renderer_take_film(MRayRenderer* r, float* hOut)
{
    if(!r || !hOut) { SetError("take_film: NULL argument"); return MRAY_ERR_INVALID; }
    const size_t n = size_t(r->rc.P) * 4;
    MRAY_HIP_CHECK(hipStreamSynchronize(r->stream));
    MRAY_HIP_CHECK(hipMemcpyAsync(hOut, r->rc.film, n * sizeof(float), hipMemcpyDeviceToHost,
                                  r->stream));
    KZeroFloats<<<DivUp(n, 256 * 8), 256, 0, r->stream>>>(r->rc.film, n);
    MRAY_HIP_CHECK(hipGetLastError());
    MRAY_HIP_CHECK(hipStreamSynchronize(r->stream));
    return MRAY_OK;
}

int mray_renderer_tiling(const MRayRenderer* r, MRayTileInfo* hOut)
{
    if(!r || !hOut) { SetError("tiling: NULL argument"); return MRAY_ERR_INVALID; }
    if(!r->tiled) { SetError("tiling: the renderer was not created with MRAY_RENDER_TILED"); return MRAY_ERR_INVALID; }
    for(int k = 0; k < 2; k++)
    {
        hOut->tileSize[k] = r->tileSize[k];
        hOut->tileCount[k] = r->tileCount[k];
    }
    return MRAY_OK;
}

// PathTracerRendererT::DoLatencyRender (PathTracerRenderer.cu:1076-1180):
// the tile's states in the slots, the tile's sub-camera, its next
// passCount x pixels path indices, bursts until all of them are in the film.
int mray_renderer_render_tile(MRayRenderer* r, uint32_t passCount, MRayTileResult* hTile,
                              MRayRenderStats* hStats)
{
    if(!r) { SetError("render_tile: renderer is NULL"); return MRAY_ERR_INVALID; }
    if(!r->tiled) { SetError("render_tile: the renderer was not created with MRAY_RENDER_TILED"); return MRAY_ERR_INVALID; }
    hipStream_t s = r->stream;
    const uint32_t W = r->params.width, H = r->params.height;
    auto rect = [&](uint32_t t, uint32_t& x0, uint32_t& y0, uint32_t& w, uint32_t& h)
    {
        x0 = (t % r->tileCount[0]) * r->tileSize[0];
        y0 = (t / r->tileCount[0]) * r->tileSize[1];
        w = std::min(r->tileSize[0], W - x0);
        h = std::min(r->tileSize[1], H - y0);
    };
    const uint32_t t = r->curTile;
    uint32_t x0, y0, w, h;
    rect(t, x0, y0, w, h);
    if(r->loadedTile != int(t))
    {
        if(r->loadedTile >= 0)
        {
            uint32_t px0, py0, pw, ph;
            rect(uint32_t(r->loadedTile), px0, py0, pw, ph);
            KTileRng<false><<<DivUp(pw * ph, 256), 256, 0, s>>>(r->S.rng, r->dPixRng, W, px0, py0, pw, pw * ph);
        }
        KTileRng<true><<<DivUp(w * h, 256), 256, 0, s>>>(r->S.rng, r->dPixRng, W, x0, y0, w, w * h);
        MRAY_HIP_CHECK(hipGetLastError());
        r->loadedTile = int(t);
    }
    MRAY_HIP_CHECK(hipStreamSynchronize(s));        // the previous upload has been read
    r->hTile.cam = SubCamera(r->rc.cam, t % r->tileCount[0], t / r->tileCount[0], r->tileCount[0],
                             r->tileCount[1], w, h);
    r->hTile.x0 = x0;
    r->hTile.y0 = y0;
    r->hTile.w = w;
    r->hTile.P = w * h;
    MRAY_HIP_CHECK(hipMemcpyAsync(r->dTile, &r->hTile, sizeof(TileGPU), hipMemcpyHostToDevice, s));
    const unsigned long long first = r->tilePaths[t];
    const uint32_t spp = r->tileSPP[t] + passCount;
    const unsigned long long lim = (unsigned long long)spp * (w * h);
    KSetTileRange<<<1, 1, 0, s>>>(r->dSched, first, lim);
    MRAY_HIP_CHECK(hipGetLastError());
    r->hostWaits = 0;
    if(lim > first)
    {
        r->pathsDone += lim - first;
        MRAY_TRY_IT(RunToLimit(r, r->pathsDone, "render_tile"));
    }
    r->tilePaths[t] = lim;
    r->tileSPP[t] = spp;
    r->lastTile.tileIndex = t;
    r->lastTile.pixelMin[0] = x0;
    r->lastTile.pixelMin[1] = y0;
    r->lastTile.pixelMax[0] = x0 + w;
    r->lastTile.pixelMax[1] = y0 + h;
    r->lastTile.tileSPP = spp;
    r->curTile = (t + 1u) % (r->tileCount[0] * r->tileCount[1]);
    if(hTile) *hTile = r->lastTile;
    MRayRenderStats total{};
    MRAY_TRY_IT(mray_renderer_stats(r, &total));
    total.hostWaits = r->hostWaits;
    total.deviceDrains = 1;
    total.finished = 1;
    for(uint32_t k : r->tileSPP) total.finished &= (k >= r->params.totalSPP) ? 1u : 0u;
    if(hStats) *hStats = total;
    return MRAY_OK;
}

int mray_renderer_take_tile(MRayRenderer* r, float* hOut)
{
    if(!r || !hOut) { SetError("take_tile: NULL argument"); return MRAY_ERR_INVALID; }
    if(!r->tiled) { SetError("take_tile: the renderer was not created with MRAY_RENDER_TILED"); return MRAY_ERR_INVALID; }
    const MRayTileResult& t = r->lastTile;
    const uint32_t w = t.pixelMax[0] - t.pixelMin[0], h = t.pixelMax[1] - t.pixelMin[1];
    if(w * h == 0) { SetError("take_tile: no tile rendered yet"); return MRAY_ERR_INVALID; }
    hipStream_t s = r->stream;
    const size_t P = r->rc.P, W = r->params.width;
    MRAY_HIP_CHECK(hipStreamSynchronize(s));
    for(size_t c = 0; c < 4; c++)
    {
        const float* src = r->rc.film + c * P + size_t(t.pixelMin[1]) * W + t.pixelMin[0];
        MRAY_HIP_CHECK(hipMemcpy2DAsync(hOut + c * size_t(w) * h, size_t(w) * 4, src, W * 4,
                                        size_t(w) * 4, h, hipMemcpyDeviceToHost, s));
    }
    KZeroRect<<<DivUp(w * h, 256), 256, 0, s>>>(r->rc.film, P, uint32_t(W), t.pixelMin[0], t.pixelMin[1],
                                                w, w * h);
    MRAY_HIP_CHECK(hipGetLastError());
    MRAY_HIP_CHECK(hipStreamSynchronize(s));
    return MRAY_OK;
}

int mray_renderer_set_measurement(MRayRenderer* r, uint32_t profiling, uint32_t counting)
{
    if(!r) { SetError("renderer is NULL"); return MRAY_ERR_INVALID; }
    MRAY_HIP_CHECK(hipDeviceSynchronize());
    r->evUsed = 0;
    r->profiling = profiling != 0;
    r->counting = counting != 0;
    for(int k = 0; k < 4; k++) { r->phaseMs[k] = 0.0; r->phaseLaunches[k] = 0; }
    return MRAY_OK;
}

int mray_renderer_phase_times(MRayRenderer* r, double* hMs, uint64_t* hLaunches)
{
    if(!r || !hMs || !hLaunches) { SetError("phase_times: NULL argument"); return MRAY_ERR_INVALID; }
    MRAY_HIP_CHECK(hipDeviceSynchronize());
    MRAY_TRY(HarvestEvents(r));
    for(int k = 0; k < 4; k++) { hMs[k] = r->phaseMs[k]; hLaunches[k] = r->phaseLaunches[k]; }
    return MRAY_OK;
}

int mray_renderer_film_device(MRayRenderer* r, float** dFilm)
{
    if(!r || !dFilm) { SetError("film_device: NULL argument"); return MRAY_ERR_INVALID; }
    *dFilm = r->rc.film;
    return MRAY_OK;
}

// Scene box of the view's scene (a one-scene renderer: its scene), mapped to
// [0,1]^3 with one scale (keeps the aspect): (min xyz, 1 / extent).
static float4 NrcBox(const MRayRenderer* r, uint32_t view, int& rv)
{
    float box[6] = {0, 0, 0, 1, 1, 1};
    rv = MRAY_OK;
    if(r->iscene) std::memcpy(box, r->iscene->box, sizeof(box));
    else if(r->bscene)
    {
        const uint32_t part = r->hViewScene[view];
        std::memcpy(box, r->bscene->parts[part].box, sizeof(box));
    }
    else if(hipMemcpy(box, r->accel->dSceneBox, sizeof(box), hipMemcpyDeviceToHost) != hipSuccess)
    {
        SetError("attach_nrc: scene box read failed");
        rv = MRAY_ERR_HIP;
    }
    float ext = std::max({box[3] - box[0], box[4] - box[1], box[5] - box[2], 1e-6f});
    return make_float4(box[0] - 1e-3f * ext, box[1] - 1e-3f * ext, box[2] - 1e-3f * ext,
                       1.0f / (ext * 1.002f));
}

int mray_renderer_attach_nrcs(MRayRenderer* r, MRayNrc* const* caches, uint32_t count,
                              const MRayNrcRenderParams* p)
{
    if(!r || !p) { SetError("attach_nrc: NULL argument"); return MRAY_ERR_INVALID; }
    MRAY_HIP_CHECK(hipDeviceSynchronize());
    if(r->graph) { (void)hipGraphExecDestroy(r->graph); r->graph = nullptr; }
    if(count == 0 || !caches)
    {
        r->caches.clear();
        r->rc.nrc = NrcRenderState{};
        r->nrcArena.Release();
        return MRAY_OK;
    }
    if(r->tiled)
    {
        SetError("attach_nrc: a tiled renderer takes no radiance cache");
        return MRAY_ERR_UNSUPPORTED;
    }
    const uint32_t views = r->batch ? uint32_t(r->hViews.size()) : 1u;
    if(count != views)
    {
        SetError("attach_nrc: " + std::to_string(count) + " caches for " + std::to_string(views) +
                 " view(s): one cache per view");
        return MRAY_ERR_INVALID;
    }
    if(views > kMaxNrcViews)
    {
        SetError("attach_nrc: at most " + std::to_string(kMaxNrcViews) + " views carry caches");
        return MRAY_ERR_UNSUPPORTED;
    }
    for(uint32_t v = 0; v < count; v++)
        if(!caches[v]) { SetError("attach_nrc: NULL cache"); return MRAY_ERR_INVALID; }
    if(r->dQueues)
    {
        SetError("attach_nrc: the radiance cache runs with the generic shading kernel "
                 "(no material queues)");
        return MRAY_ERR_UNSUPPORTED;
    }
    if(p->trainShift > 16 || p->trainBatch == 0 || (p->trainBatch % 32u) != 0 || p->trainSteps == 0 ||
       !(p->spreadThreshold > 0.0f))
    {
        SetError("attach_nrc: trainShift <= 16, trainBatch a non-zero multiple of 32, "
                 "trainSteps > 0, spreadThreshold > 0");
        return MRAY_ERR_INVALID;
    }
    uint64_t ring = uint64_t(p->trainBatch) * p->trainSteps;
    if((ring & (ring - 1)) != 0 || ring > (1u << 24))
    {
        SetError("attach_nrc: trainBatch * trainSteps must be a power of two <= 2^24");
        return MRAY_ERR_INVALID;
    }
    r->nrcArena.Release();
    NrcRenderState& nr = r->rc.nrc;
    nr = NrcRenderState{};
    const uint32_t N = r->rc.N;
    const uint32_t slots = (N >> p->trainShift) + 1u;
    nr.trainShift = p->trainShift;
    nr.maxVerts = std::min(64u, std::max(1u, r->rc.rr1));
    r->nrcLog2G = 0;
    while((1u << r->nrcLog2G) < nr.maxVerts) r->nrcLog2G++;
    nr.ringSize = uint32_t(ring);
    nr.spreadC = p->spreadThreshold;
    nr.minDepth = std::max(1u, p->minDepth);
    nr.neeAtEnd = (p->flags & MRAY_NRC_NEE_AT_END) ? 1u : 0u;
    nr.views = views;
    nr.queryCap = N;
    auto A = [&](void** ptr, size_t bytes) { return DevAlloc(r->nrcArena, ptr, bytes); };
    MRAY_TRY_IT(A((void**)&nr.trainStore, size_t(slots) * nr.maxVerts * kNrcVert * 4));
    MRAY_TRY_IT(A((void**)&nr.trainCount, size_t(slots) * 4));
    MRAY_TRY_IT(A((void**)&nr.queryRec, size_t(views) * N * kNrcRec * 4));
    MRAY_TRY_IT(A((void**)&nr.querySlot, size_t(views) * N * 4));
    MRAY_TRY_IT(A((void**)&nr.queryOut, size_t(views) * N * 3 * 4));
    MRAY_TRY_IT(A((void**)&nr.ringRec, size_t(views) * ring * kNrcRec * 4));
    MRAY_TRY_IT(A((void**)&nr.ringTgt, size_t(views) * ring * 3 * 4));
    MRAY_TRY_IT(A((void**)&nr.ringCount, size_t(views) * sizeof(unsigned long long)));
    MRAY_TRY_IT(A((void**)&r->dNrcBox, size_t(views) * sizeof(float4)));
    MRAY_TRY_IT(A((void**)&r->S.spread, size_t(N) * sizeof(float2)));
    MRAY_HIP_CHECK(hipMemset(nr.trainCount, 0, size_t(slots) * 4));
    MRAY_HIP_CHECK(hipMemset(nr.ringTgt, 0, size_t(views) * ring * 3 * 4));
    MRAY_HIP_CHECK(hipMemset(nr.ringRec, 0, size_t(views) * ring * kNrcRec * 4));
    MRAY_HIP_CHECK(hipMemset(nr.ringCount, 0, size_t(views) * sizeof(unsigned long long)));
    std::vector<float4> boxes(views);
    for(uint32_t v = 0; v < views; v++)
    {
        int rv;
        boxes[v] = NrcBox(r, v, rv);
        if(rv != MRAY_OK) return rv;
    }
    MRAY_HIP_CHECK(hipMemcpy(r->dNrcBox, boxes.data(), views * sizeof(float4), hipMemcpyHostToDevice));
    nr.box = r->dNrcBox;
    nr.on = 1;
    r->caches.assign(caches, caches + count);
    r->nrcParams = *p;
    return MRAY_OK;
}

int mray_renderer_attach_nrc(MRayRenderer* r, MRayNrc* nrc, const MRayNrcRenderParams* p)
{
    return mray_renderer_attach_nrcs(r, nrc ? &nrc : nullptr, nrc ? 1u : 0u, p);
}

// A batch renderer switches to frame order at its first frame (and back to
// view order at reset): the captured bursts hold the old constants.
static int SetFrameOrder(MRayRenderer* r, bool on)
{
    if(bool(r->rc.frameMajor) == on) return MRAY_OK;
    MRAY_HIP_CHECK(hipDeviceSynchronize());
    if(r->graph) { (void)hipGraphExecDestroy(r->graph); r->graph = nullptr; }
    r->rc.frameMajor = on ? 1u : 0u;
    return MRAY_OK;
}

int mray_renderer_render_frame(MRayRenderer* r, uint32_t spp, MRayRenderStats* hStats)
{
    if(!r || spp == 0) { SetError("render_frame: NULL renderer or spp == 0"); return MRAY_ERR_INVALID; }
    if(r->tiled) { SetError("render_frame: a tiled renderer renders with mray_renderer_render_tile"); return MRAY_ERR_INVALID; }
    hipStream_t s = r->stream;
    if(r->batch && !r->rc.frameMajor)
    {
        // every frame renders `spp` samples of every view: restart in frame order
        MRAY_TRY_IT(SetFrameOrder(r, true));
        MRAY_TRY_IT(ResetRenderer(r));
    }
    r->frameLimit += (unsigned long long)spp * r->rc.P;
    KSetLimit<<<1, 1, 0, s>>>(r->dSched, r->frameLimit);
    MRAY_HIP_CHECK(hipGetLastError());
    MRAY_TRY_IT(RunToLimit(r, r->frameLimit, "render_frame"));
    MRayRenderStats total{};
    MRAY_TRY_IT(mray_renderer_stats(r, &total));
    total.hostWaits = r->hostWaits;
    total.deviceDrains = 1;
    if(r->rc.nrc.on)
    {
        // online training of every view's cache on its most recent pairs
        const NrcRenderState& nr = r->rc.nrc;
        std::vector<unsigned long long> made(nr.views);
        MRAY_HIP_CHECK(hipMemcpy(made.data(), nr.ringCount, nr.views * sizeof(unsigned long long),
                                 hipMemcpyDeviceToHost));
        const uint32_t B = r->nrcParams.trainBatch;
        for(uint32_t v = 0; v < nr.views; v++)
        {
            unsigned long long have = std::min<unsigned long long>(made[v], nr.ringSize);
            uint32_t steps = uint32_t(std::min<unsigned long long>(r->nrcParams.trainSteps, have / B));
            const size_t base = size_t(v) * nr.ringSize;
            for(uint32_t k = 0; k < steps; k++)
                MRAY_TRY_IT(mray_nrc_train(r->caches[v], nr.ringRec + (base + size_t(k) * B) * kNrcRec,
                                           nr.ringTgt + (base + size_t(k) * B) * 3, B, s));
        }
        MRAY_HIP_CHECK(hipStreamSynchronize(s));
    }
    total.finished = 1;
    if(hStats) *hStats = total;
    return MRAY_OK;
}

int mray_renderer_cache_stats(MRayRenderer* r, uint64_t* hQueries, uint64_t* hTrainingPairs)
{
    if(!r || !hQueries || !hTrainingPairs) { SetError("cache_stats: NULL argument"); return MRAY_ERR_INVALID; }
    MRAY_HIP_CHECK(hipDeviceSynchronize());
    MRAY_HIP_CHECK(hipMemcpy(r->hSched, r->dSched, sizeof(Sched), hipMemcpyDeviceToHost));
    *hQueries = r->hSched->queries;
    *hTrainingPairs = 0;
    const NrcRenderState& nr = r->rc.nrc;
    if(nr.on)
    {
        std::vector<unsigned long long> made(nr.views);
        MRAY_HIP_CHECK(hipMemcpy(made.data(), nr.ringCount, nr.views * sizeof(unsigned long long),
                                 hipMemcpyDeviceToHost));
        for(unsigned long long m : made) *hTrainingPairs += m;
    }
    return MRAY_OK;
}

int mray_batch_create(const MRayBatchScene* scene, const MRayBatchView* hViews, uint32_t viewCount,
                      const MRayRenderParams* hParams, MRayRenderer** out)
{
    if(!scene || !hViews || !hParams || !out || viewCount == 0)
    {
        SetError("batch_create: NULL argument or no views");
        return MRAY_ERR_INVALID;
    }
    *out = nullptr;
    MRAY_TRY(CheckSampling(*hParams, "batch_create"));
    uint64_t pixels = 0;
    for(uint32_t k = 0; k < viewCount; k++)
    {
        const MRayBatchView& v = hViews[k];
        if(v.scene >= scene->parts.size())
        {
            SetError("batch_create: view " + std::to_string(k) + " names scene " +
                     std::to_string(v.scene) + " of " + std::to_string(scene->parts.size()));
            return MRAY_ERR_INVALID;
        }
        if(v.width == 0 || v.height == 0 || v.spp == 0)
        {
            SetError("batch_create: view width, height and spp must be non-zero");
            return MRAY_ERR_INVALID;
        }
        pixels += uint64_t(v.width) * v.height;
    }
    if(pixels > (1ull << 31))
    {
        SetError("batch_create: the views hold more than 2^31 pixels");
        return MRAY_ERR_INVALID;
    }
    MRayRenderer* r = new MRayRenderer();
    int rv = CreateBatch(r, scene, nullptr, hViews, viewCount, hParams);
    if(rv != MRAY_OK)
    {
        (void)hipDeviceSynchronize();
        FreeRenderer(r);
        return rv;
    }
    *out = r;
    return MRAY_OK;
}

int mray_instanced_batch_create(const MRayInstScene* scene, const MRayBatchView* hViews,
                                uint32_t viewCount, const MRayRenderParams* hParams, MRayRenderer** out)
{
    if(!scene || !hViews || !hParams || !out || viewCount == 0)
    {
        SetError("instanced_batch_create: NULL argument or no views");
        return MRAY_ERR_INVALID;
    }
    *out = nullptr;
    MRAY_TRY(CheckSampling(*hParams, "instanced_batch_create"));
    if(hParams->flags & MRAY_RENDER_MATERIAL_QUEUES)
    {
        SetError("instanced_batch_create: material queues are not available for instanced scenes");
        return MRAY_ERR_UNSUPPORTED;
    }
    uint64_t pixels = 0;
    for(uint32_t k = 0; k < viewCount; k++)
    {
        const MRayBatchView& v = hViews[k];
        if(v.scene != 0)
        {
            SetError("instanced_batch_create: view " + std::to_string(k) + " names scene " +
                     std::to_string(v.scene) + "; an instanced scene is scene 0");
            return MRAY_ERR_INVALID;
        }
        if(v.width == 0 || v.height == 0 || v.spp == 0)
        {
            SetError("instanced_batch_create: view width, height and spp must be non-zero");
            return MRAY_ERR_INVALID;
        }
        pixels += uint64_t(v.width) * v.height;
    }
    if(pixels > (1ull << 31))
    {
        SetError("instanced_batch_create: the views hold more than 2^31 pixels");
        return MRAY_ERR_INVALID;
    }
    MRayRenderer* r = new MRayRenderer();
    int rv = CreateBatch(r, nullptr, scene, hViews, viewCount, hParams);
    if(rv != MRAY_OK)
    {
        (void)hipDeviceSynchronize();
        FreeRenderer(r);
        return rv;
    }
    *out = r;
    return MRAY_OK;
}

int mray_renderer_light_table(const MRayRenderer* r, float* hProb, uint32_t* hAlias, float* hPmf)
{
    if(!r || !hProb || !hAlias || !hPmf) { SetError("renderer_light_table: NULL argument"); return MRAY_ERR_INVALID; }
    if(!r->dLightAlias)
    {
        SetError("renderer_light_table: the renderer samples lights uniformly");
        return MRAY_ERR_INVALID;
    }
    std::vector<uint2> packed(r->lightTableCount);
    MRAY_HIP_CHECK(hipMemcpy(packed.data(), r->dLightAlias, packed.size() * sizeof(uint2),
                             hipMemcpyDeviceToHost));
    MRAY_HIP_CHECK(hipMemcpy(hPmf, r->dLightPmf, size_t(r->lightTableCount) * sizeof(float),
                             hipMemcpyDeviceToHost));
    for(uint32_t i = 0; i < r->lightTableCount; i++)
    {
        std::memcpy(&hProb[i], &packed[i].x, 4);
        hAlias[i] = packed[i].y;
    }
    return MRAY_OK;
}

int mray_batch_read_film(MRayRenderer* r, uint32_t view, float* hOut)
{
    if(!r || !hOut) { SetError("batch_read_film: NULL argument"); return MRAY_ERR_INVALID; }
    if(!r->batch || view >= r->hViews.size())
    {
        SetError("batch_read_film: not a batch renderer or view out of range");
        return MRAY_ERR_INVALID;
    }
    const ViewGPU& v = r->hViews[view];
    MRAY_HIP_CHECK(hipDeviceSynchronize());
    MRAY_HIP_CHECK(hipMemcpy(hOut, v.film, size_t(v.P) * 4 * sizeof(float), hipMemcpyDeviceToHost));
    return MRAY_OK;
}

int mray_batch_view_times(MRayRenderer* r, double* hMs)
{
    if(!r || !hMs) { SetError("batch_view_times: NULL argument"); return MRAY_ERR_INVALID; }
    if(!r->batch) { SetError("batch_view_times: not a batch renderer"); return MRAY_ERR_INVALID; }
    const size_t V = r->hViews.size();
    std::vector<unsigned long long> clk(V);
    MRAY_HIP_CHECK(hipDeviceSynchronize());
    MRAY_HIP_CHECK(hipMemcpy(clk.data(), r->dViewClock, V * sizeof(unsigned long long),
                             hipMemcpyDeviceToHost));
    MRAY_HIP_CHECK(hipMemcpy(r->hSched, r->dSched, sizeof(Sched), hipMemcpyDeviceToHost));
    int dev = 0, khz = 0;
    MRAY_HIP_CHECK(hipGetDevice(&dev));
    MRAY_HIP_CHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    const unsigned long long t0 = r->hSched->startClock;
    for(size_t k = 0; k < V; k++)
        hMs[k] = (clk[k] == 0 || khz <= 0) ? -1.0 : double(clk[k] - t0) / double(khz);
    return MRAY_OK;
}

int mray_batch_set_trace_tpb(MRayRenderer* r, uint32_t tpb)
{
    if(!r || !r->bscene || (tpb != 128 && tpb != 256))
    {
        SetError("batch_set_trace_tpb: a batch renderer over a merged scene and tpb 128 or 256");
        return MRAY_ERR_INVALID;
    }
    MRAY_HIP_CHECK(hipDeviceSynchronize());
    if(r->graph) { (void)hipGraphExecDestroy(r->graph); r->graph = nullptr; }
    r->traceTPB = tpb;
    return SizeLaunches(r);
}

int mray_batch_autotune(const MRayBatchScene* scene4, const MRayBatchScene* scene8,
                        const MRayBatchView* hViews, uint32_t viewCount,
                        const MRayRenderParams* hParams, MRayBatchTune* hBest,
                        MRayBatchTune* hTrials, uint32_t maxTrials, uint32_t* hTrialCount)
{
    if((!scene4 && !scene8) || !hViews || !hParams || !hBest || viewCount == 0)
    {
        SetError("batch_autotune: NULL argument");
        return MRAY_ERR_INVALID;
    }
    if((scene4 && scene4->width != 4) || (scene8 && scene8->width != 8))
    {
        SetError("batch_autotune: scene4 / scene8 must hold 4-wide / 8-wide trees");
        return MRAY_ERR_INVALID;
    }
    // trials: every view at 1 spp, so each scene weighs in as in the real batch
    std::vector<MRayBatchView> tv(hViews, hViews + viewCount);
    uint64_t pixels = 0;
    for(MRayBatchView& v : tv)
    {
        v.spp = 1;
        pixels += uint64_t(v.width) * v.height;
    }
    std::vector<uint32_t> slotCands;
    for(uint32_t c : {1u << 20, 1u << 21, 1u << 22})
    {
        uint32_t n = uint32_t(std::min<uint64_t>(c, pixels));
        if(std::find(slotCands.begin(), slotCands.end(), n) == slotCands.end()) slotCands.push_back(n);
    }
    MRayBatchTune best{};
    uint32_t nTrials = 0;
    const MRayBatchScene* scenes[2] = {scene4, scene8};
    for(const MRayBatchScene* sc : scenes)
    {
        if(!sc) continue;
        for(uint32_t queues = 0; queues < 2; queues++)
            for(uint32_t slots : slotCands)
            {
                MRayRenderParams p = *hParams;
                p.maxRayCount = slots;
                p.flags = (p.flags & ~MRAY_RENDER_MATERIAL_QUEUES) | (queues ? MRAY_RENDER_MATERIAL_QUEUES : 0u);
                MRayRenderer* r = nullptr;
                MRAY_TRY(mray_batch_create(sc, tv.data(), viewCount, &p, &r));
                for(uint32_t tpb : {256u, 128u})
                {
                    int rv = mray_batch_set_trace_tpb(r, tpb);
                    MRayRenderStats st{};
                    // untimed pass (graph capture, caches), then the timed one
                    if(rv == MRAY_OK) rv = mray_renderer_render_all(r, &st);
                    if(rv == MRAY_OK) rv = mray_renderer_reset(r);
                    auto t0 = std::chrono::steady_clock::now();
                    if(rv == MRAY_OK) rv = mray_renderer_render_all(r, &st);
                    auto t1 = std::chrono::steady_clock::now();
                    if(rv != MRAY_OK)
                    {
                        mray_renderer_destroy(r);
                        return rv;
                    }
                    double sec = std::chrono::duration<double>(t1 - t0).count();
                    MRayBatchTune t{sc->width, queues, slots, tpb, double(st.totalDeadPaths) / sec * 1e-6};
                    if(hTrials && nTrials < maxTrials) hTrials[nTrials] = t;
                    nTrials++;
                    if(t.mpathsPerS > best.mpathsPerS) best = t;
                }
                mray_renderer_destroy(r);
            }
    }
    *hBest = best;
    if(hTrialCount) *hTrialCount = nTrials;
    return MRAY_OK;
}

} // extern "C"
