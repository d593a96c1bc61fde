// Ray-casting kernels behind mray_accel_cast_closest / _visibility
// (BaseAcceleratorI::CastRays / CastVisibilityRays, Tracer/AcceleratorC.h:375-430).
#include "traverse.h"

namespace mray
{

template<int W, bool COUNT>
__global__ void __launch_bounds__(kTraceTPB)
KCCastClosest(TraceCtx ctx, const float4* __restrict__ rays, uint32_t n,
              uint32_t* __restrict__ prim, float* __restrict__ tOut, float* __restrict__ bary,
              unsigned long long* __restrict__ counters)
{
    __shared__ uint2 lstack[kLdsStack * kTraceTPB];
    uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t nodeVisits = 0, triTests = 0;
    for(uint32_t i = gtid; i < n; i += ctx.gridThreads)
    {
        float4 a = rays[2 * size_t(i)];
        float4 b = rays[2 * size_t(i) + 1];
        HitRec h = TraceW<W, false, COUNT>(ctx, lstack, gtid, V3{a.x, a.y, a.z}, V3{b.x, b.y, b.z},
                                       a.w, b.w, 0u, &nodeVisits, &triTests);
        prim[i] = h.prim;
        if(tOut) tOut[i] = (h.prim == kInvalid) ? b.w : h.t;
        if(bary)
        {
            bary[2 * size_t(i)] = (h.prim == kInvalid) ? 0.0f : h.a;
            bary[2 * size_t(i) + 1] = (h.prim == kInvalid) ? 0.0f : h.b;
        }
    }
    if(COUNT)
    {
        atomicAdd(&counters[0], (unsigned long long)nodeVisits);
        atomicAdd(&counters[1], (unsigned long long)triTests);
    }
}

template<int W>
__global__ void __launch_bounds__(kTraceTPB)
KCCastVisibility(TraceCtx ctx, const float4* __restrict__ rays, uint32_t n,
                 uint32_t* __restrict__ visible)
{
    __shared__ uint2 lstack[kLdsStack * kTraceTPB];
    uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    for(uint32_t i = gtid; i < n; i += ctx.gridThreads)
    {
        float4 a = rays[2 * size_t(i)];
        float4 b = rays[2 * size_t(i) + 1];
        HitRec h = TraceW<W, true, false>(ctx, lstack, gtid, V3{a.x, a.y, a.z}, V3{b.x, b.y, b.z},
                                      a.w, b.w, 0u);
        visible[i] = (h.prim == kInvalid) ? 1u : 0u;
    }
}

int LaunchCastClosest(const TraceCtx& ctx, const float* dRays, uint32_t n, uint32_t* dPrim,
                      float* dT, float* dBary, unsigned long long* dCounters, hipStream_t s)
{
    if(n == 0) return MRAY_OK;
    uint32_t blocks = std::min(ctx.gridThreads / kTraceTPB, DivUp(n, kTraceTPB));
    const float4* r = reinterpret_cast<const float4*>(dRays);
    const bool w8 = ctx.width == 8;
    if(dCounters)
    {
        if(w8) KCCastClosest<8, true><<<blocks, kTraceTPB, 0, s>>>(ctx, r, n, dPrim, dT, dBary, dCounters);
        else KCCastClosest<4, true><<<blocks, kTraceTPB, 0, s>>>(ctx, r, n, dPrim, dT, dBary, dCounters);
    }
    else
    {
        if(w8) KCCastClosest<8, false><<<blocks, kTraceTPB, 0, s>>>(ctx, r, n, dPrim, dT, dBary, nullptr);
        else KCCastClosest<4, false><<<blocks, kTraceTPB, 0, s>>>(ctx, r, n, dPrim, dT, dBary, nullptr);
    }
    MRAY_HIP_CHECK(hipGetLastError());
    return MRAY_OK;
}

int LaunchCastVisibility(const TraceCtx& ctx, const float* dRays, uint32_t n, uint32_t* dVisible,
                         hipStream_t s)
{
    if(n == 0) return MRAY_OK;
    uint32_t blocks = std::min(ctx.gridThreads / kTraceTPB, DivUp(n, kTraceTPB));
    const float4* r = reinterpret_cast<const float4*>(dRays);
    if(ctx.width == 8) KCCastVisibility<8><<<blocks, kTraceTPB, 0, s>>>(ctx, r, n, dVisible);
    else KCCastVisibility<4><<<blocks, kTraceTPB, 0, s>>>(ctx, r, n, dVisible);
    MRAY_HIP_CHECK(hipGetLastError());
    return MRAY_OK;
}

} // namespace mray
