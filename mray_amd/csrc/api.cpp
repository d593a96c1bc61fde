// C-ABI of libmray_amd.so: library utilities and the accelerator.
// Declarations and the reference interfaces they replace: include/mray_amd.h.
#include "common.h"
#include "lbvh.h"
#include "scene.h"
#include "traverse.h"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

namespace mray
{

static thread_local std::string gLastError;

void SetError(const std::string& msg) { gLastError = msg; }

int LaunchCastClosest(const TraceCtx& ctx, const float* dRays, uint32_t n, uint32_t* dPrim,
                      float* dT, float* dBary, unsigned long long* dCounters, hipStream_t s);
int LaunchCastVisibility(const TraceCtx& ctx, const float* dRays, uint32_t n, uint32_t* dVisible,
                         hipStream_t s);

// Device allocation helper: records every pointer for release.
int DevAlloc(DeviceArena& arena, void** p, size_t bytes)
{
    *p = nullptr;
    if(bytes == 0) return MRAY_OK;
    hipError_t e = hipMalloc(p, bytes);
    if(e != hipSuccess)
    {
        SetError(std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
        return MRAY_ERR_OOM;
    }
    arena.ptrs.push_back(*p);
    return MRAY_OK;
}

void DeviceArena::Release()
{
    for(void* p : ptrs) (void)hipFree(p);
    ptrs.clear();
}

template<class T>
int Upload(DeviceArena& a, T** d, const T* h, size_t count)
{
    MRAY_TRY(DevAlloc(a, reinterpret_cast<void**>(d), count * sizeof(T)));
    if(count) MRAY_HIP_CHECK(hipMemcpy(*d, h, count * sizeof(T), hipMemcpyHostToDevice));
    return MRAY_OK;
}

static int ValidateScene(const MRaySceneDesc* s)
{
    if(!s) { SetError("scene description is NULL"); return MRAY_ERR_INVALID; }
    if(s->triangleCount == 0) { SetError("scene has no triangles"); return MRAY_ERR_INVALID; }
    if(s->triangleCount >= (1u << kLeafCountShift))
    {
        SetError("scene has too many triangles (limit 2^26)");
        return MRAY_ERR_INVALID;
    }
    if(!s->hPositions || !s->hIndices)
    {
        SetError("scene positions / indices are NULL");
        return MRAY_ERR_INVALID;
    }
    for(size_t i = 0; i < size_t(s->triangleCount) * 3; i++)
        if(s->hIndices[i] >= s->vertexCount)
        {
            SetError("triangle index out of range of vertexCount");
            return MRAY_ERR_INVALID;
        }
    for(uint32_t t = 0; s->hSurfaceKeys && t < s->triangleCount; t++)
    {
        uint32_t k = s->hSurfaceKeys[t];
        if(k & MRAY_LIGHT_FLAG)
        {
            uint32_t li = k & ~MRAY_LIGHT_FLAG;
            if(li >= s->lightCount || s->hLightTriangles[li] != t)
            {
                SetError("light surface key does not match hLightTriangles");
                return MRAY_ERR_INVALID;
            }
        }
        else if(k >= s->materialCount)
        {
            SetError("material index out of range");
            return MRAY_ERR_INVALID;
        }
    }
    for(uint32_t m = 0; m < s->materialCount; m++)
        if(s->hMaterials[m].type > MRAY_MAT_UNREAL)
        {
            SetError("unknown material type");
            return MRAY_ERR_INVALID;
        }
    if(s->boundaryType > MRAY_BOUNDARY_SKYSPHERE_SPHERICAL)
    {
        SetError("unknown boundary light type");
        return MRAY_ERR_INVALID;
    }
    return MRAY_OK;
}

// AcceleratorGroupLBVH::Construct + BaseAcceleratorLBVH::InternalConstruct
// (AcceleratorLBVH.hpp:433-582, AcceleratorLBVH.cu:537-740): K1..K7 on the
// GPU, then the 4-wide collapse and the triangle soup.
static int BuildAccel(MRayAccel* a, const MRaySceneDesc* s, uint32_t builder, uint32_t width,
                      bool hostCollapse)
{
    hipStream_t st = nullptr;
    const uint32_t n = s->triangleCount;
    DeviceArena& ar = a->arena;
    MRAY_TRY(Upload(ar, &a->dPos, s->hPositions, size_t(s->vertexCount) * 3));
    MRAY_TRY(Upload(ar, &a->dIdx, s->hIndices, size_t(n) * 3));
    MRAY_TRY(DevAlloc(ar, (void**)&a->dCenters, size_t(n) * 3 * sizeof(float)));
    MRAY_TRY(DevAlloc(ar, (void**)&a->dAabbs, size_t(n) * 6 * sizeof(float)));
    MRAY_TRY(DevAlloc(ar, (void**)&a->dSceneBox, 6 * sizeof(float)));
    MRAY_TRY(DevAlloc(ar, (void**)&a->dCodes, size_t(n) * sizeof(uint64_t)));
    MRAY_TRY(DevAlloc(ar, (void**)&a->dSortedCodes, size_t(n) * sizeof(uint64_t)));
    MRAY_TRY(DevAlloc(ar, (void**)&a->dLeafOrder, size_t(n) * sizeof(uint32_t)));
    uint32_t nodeCount = (n > 1) ? n - 1 : 1;
    MRAY_TRY(DevAlloc(ar, (void**)&a->dBinNodes, size_t(nodeCount) * 3 * sizeof(uint32_t)));
    MRAY_TRY(DevAlloc(ar, (void**)&a->dBinBoxes, size_t(nodeCount) * 6 * sizeof(float)));
    MRAY_TRY(DevAlloc(ar, (void**)&a->dBinRange, size_t(nodeCount) * 2 * sizeof(uint32_t)));
    size_t tempBytes = std::max({RadixSortTempBytes(n), LbvhTempBytes(n), ReduceAabbTempBytes()});
    void* dTemp = nullptr;
    MRAY_TRY(DevAlloc(ar, &dTemp, tempBytes));

    MRAY_TRY(LaunchPrimCentersAabbs(a->dPos, a->dIdx, n, a->dCenters, a->dAabbs, st));
    MRAY_TRY(LaunchReduceAabb(a->dAabbs, n, a->dSceneBox, static_cast<float*>(dTemp), st));
    MRAY_TRY(LaunchMorton63(a->dCenters, n, a->dSceneBox, a->dCodes, st));
    MRAY_HIP_CHECK(hipMemcpyAsync(a->dSortedCodes, a->dCodes, size_t(n) * 8,
                                  hipMemcpyDeviceToDevice, st));
    MRAY_TRY(LaunchIota(a->dLeafOrder, n, st));
    MRAY_TRY(LaunchRadixSort(a->dSortedCodes, a->dLeafOrder, n, dTemp, tempBytes, st));
    MRAY_TRY(LaunchLbvhBuild(a->dSortedCodes, a->dLeafOrder, a->dAabbs, n, a->dBinNodes,
                             a->dBinBoxes, a->dBinRange, dTemp, tempBytes, st));
    MRAY_HIP_CHECK(hipStreamSynchronize(st));

    // The binary tree's topology, checked on the host: every node and leaf
    // reached once from the root, each split at one leaf-order point.
    std::vector<uint32_t> binNodes(size_t(nodeCount) * 3), binRange(size_t(nodeCount) * 2);
    MRAY_HIP_CHECK(hipMemcpy(binNodes.data(), a->dBinNodes, binNodes.size() * 4,
                             hipMemcpyDeviceToHost));
    MRAY_HIP_CHECK(hipMemcpy(binRange.data(), a->dBinRange, binRange.size() * 4,
                             hipMemcpyDeviceToHost));
    if(!ValidateBinaryTree(binNodes, binRange, n))
    {
        // The reference's equal-code Delta broke the Karras prefix order for
        // this code set; rebuild with the (code, index) tie-break.
        MRAY_TRY(LaunchLbvhBuild(a->dSortedCodes, a->dLeafOrder, a->dAabbs, n, a->dBinNodes,
                                 a->dBinBoxes, a->dBinRange, dTemp, tempBytes, st, 1u));
        MRAY_HIP_CHECK(hipStreamSynchronize(st));
        MRAY_HIP_CHECK(hipMemcpy(binNodes.data(), a->dBinNodes, binNodes.size() * 4,
                                 hipMemcpyDeviceToHost));
        MRAY_HIP_CHECK(hipMemcpy(binRange.data(), a->dBinRange, binRange.size() * 4,
                                 hipMemcpyDeviceToHost));
        a->strictTie = 1;
        if(!ValidateBinaryTree(binNodes, binRange, n))
        {
            SetError("LBVH construction produced a malformed tree");
            return MRAY_ERR_INVALID;
        }
    }
    a->builder = builder;
    a->width = width;
    if(builder == MRAY_BUILDER_LBVH && width == 4 && n > 1 && !hostCollapse)
    {
        // the whole build on the device: collapse, then the soup in leaf order
        MRAY_TRY(CollapseToBvh4Device(a->dBinNodes, a->dBinBoxes, a->dBinRange, a->dAabbs, n, ar,
                                      &a->dNodes, &a->wideNodes, &a->leafSlots, &a->maxDepth, st));
        MRAY_TRY(DevAlloc(ar, (void**)&a->dSoup, size_t(n) * sizeof(SoupTri)));
        uint8_t* dCull = nullptr;
        if(s->hCullBackFace)
        {
            MRAY_HIP_CHECK(hipMalloc((void**)&dCull, n));
            MRAY_HIP_CHECK(hipMemcpy(dCull, s->hCullBackFace, n, hipMemcpyHostToDevice));
        }
        int rg = LaunchGatherSoup(a->dPos, a->dIdx, dCull, a->dLeafOrder, n, a->dSoup, st);
        MRAY_HIP_CHECK(hipStreamSynchronize(st));
        if(dCull) MRAY_HIP_CHECK(hipFree(dCull));
        MRAY_TRY(rg);
    }
    else
    {
        // Collapse on the host (SAH tree, 8-wide nodes, or on request) and
        // gather the triangle soup there.
        std::vector<float> binBoxes(size_t(nodeCount) * 6), aabbs(size_t(n) * 6);
        std::vector<uint32_t> leafOrder(n);
        MRAY_HIP_CHECK(hipMemcpy(binBoxes.data(), a->dBinBoxes, binBoxes.size() * 4,
                                 hipMemcpyDeviceToHost));
        MRAY_HIP_CHECK(hipMemcpy(aabbs.data(), a->dAabbs, aabbs.size() * 4, hipMemcpyDeviceToHost));
        MRAY_HIP_CHECK(hipMemcpy(leafOrder.data(), a->dLeafOrder, size_t(n) * 4,
                                 hipMemcpyDeviceToHost));
        if(builder == MRAY_BUILDER_SAH && n > 1)
        {
            // traversal tree: binned SAH over the same triangle boxes (the LBVH
            // above stays on the device for the reference-layout products)
            std::vector<uint32_t> sNodes, sRange, sOrder;
            std::vector<float> sBoxes;
            if(!BuildSahBinary(aabbs, n, sNodes, sBoxes, sRange, sOrder))
            {
                SetError("SAH construction produced a malformed tree");
                return MRAY_ERR_INVALID;
            }
            binNodes.swap(sNodes);
            binBoxes.swap(sBoxes);
            binRange.swap(sRange);
            leafOrder.swap(sOrder);
        }
        if(width == 8)
        {
            CBvh8 w;
            if(!CollapseToCBvh8(binNodes, binBoxes, binRange, aabbs, leafOrder, w))
            {
                SetError("compressed 8-wide BVH: a child box has no conservative 8-bit encoding");
                return MRAY_ERR_INVALID;
            }
            a->wideNodes = uint32_t(w.nodes.size());
            a->leafSlots = w.leafSlots;
            a->maxDepth = w.maxDepth;
            MRAY_TRY(Upload(ar, &a->dCNodes, w.nodes.data(), w.nodes.size()));
        }
        else
        {
            WideBvh w = CollapseToBvh4(binNodes, binBoxes, binRange, aabbs, leafOrder);
            a->wideNodes = uint32_t(w.nodes.size());
            a->leafSlots = w.leafSlots;
            a->maxDepth = w.maxDepth;
            MRAY_TRY(Upload(ar, &a->dNodes, w.nodes.data(), w.nodes.size()));
        }
        std::vector<SoupTri> soup(n);
        for(uint32_t p = 0; p < n; p++)
        {
            uint32_t t = leafOrder[p];
            SoupTri& o = soup[p];
            for(int v = 0; v < 3; v++)
            {
                const float* src = s->hPositions + 3 * size_t(s->hIndices[3 * size_t(t) + v]);
                float* dst = (v == 0) ? o.v0 : (v == 1) ? o.v1 : o.v2;
                dst[0] = src[0]; dst[1] = src[1]; dst[2] = src[2];
            }
            o.prim = t;
            o.cull = s->hCullBackFace ? s->hCullBackFace[t] : 0u;
            o.pad = 0;
        }
        MRAY_TRY(Upload(ar, &a->dSoup, soup.data(), soup.size()));
    }

    // Traversal stack spill area: <= width-1 entries per level below the root.
    int dev = 0;
    MRAY_HIP_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    MRAY_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
    uint32_t cus = uint32_t(std::max(1, prop.multiProcessorCount));
    a->ctx.gridThreads = cus * 8u * kTraceTPB;
    uint32_t need = (width - 1u) * a->maxDepth + 4u;
    a->ctx.gstackDepth = (need > kLdsStack) ? need - kLdsStack : 1u;
    MRAY_TRY(DevAlloc(ar, (void**)&a->ctx.gstack,
                      size_t(a->ctx.gstackDepth) * a->ctx.gridThreads * sizeof(uint2)));
    a->ctx.maxVisits = a->wideNodes + 1u;
    a->ctx.nodes = a->dNodes;
    a->ctx.cnodes = a->dCNodes;
    a->ctx.soup = a->dSoup;
    a->ctx.width = width;
    return MRAY_OK;
}

} // namespace mray

using namespace mray;

extern "C" {

const char* mray_last_error(void) { return gLastError.c_str(); }
const char* mray_version(void) { return "mray_amd 0.1 (gfx950)"; }

int mray_device_count(int* hCount)
{
    if(!hCount) return MRAY_ERR_INVALID;
    MRAY_HIP_CHECK(hipGetDeviceCount(hCount));
    return MRAY_OK;
}

int mray_device_memory(size_t* hTotal, size_t* hUsed)
{
    if(!hTotal || !hUsed) return MRAY_ERR_INVALID;
    size_t freeB = 0, total = 0;
    MRAY_HIP_CHECK(hipMemGetInfo(&freeB, &total));
    *hTotal = total;
    *hUsed = total - freeB;
    return MRAY_OK;
}

int mray_prim_centers_aabbs(const float* dPositions, const uint32_t* dIndices,
                            uint32_t triangleCount, float* dCenters, float* dAabbs, void* stream)
{
    return LaunchPrimCentersAabbs(dPositions, dIndices, triangleCount, dCenters, dAabbs,
                                  static_cast<hipStream_t>(stream));
}

int mray_reduce_aabb(const float* dAabbs, uint32_t count, float* dOut, void* stream)
{
    if(count == 0) { SetError("reduce of zero boxes"); return MRAY_ERR_INVALID; }
    hipStream_t s = static_cast<hipStream_t>(stream);
    float* dPartial = nullptr;
    MRAY_HIP_CHECK(hipMallocAsync((void**)&dPartial, ReduceAabbTempBytes(), s));
    int r = LaunchReduceAabb(dAabbs, count, dOut, dPartial, s);
    MRAY_HIP_CHECK(hipFreeAsync(dPartial, s));
    return r;
}

int mray_morton63(const float* dCenters, uint32_t count, const float* dSceneAabb,
                  uint64_t* dCodes, void* stream)
{
    return LaunchMorton63(dCenters, count, dSceneAabb, dCodes, static_cast<hipStream_t>(stream));
}

int mray_morton30(const float* dCenters, uint32_t count, const float* dSceneAabb,
                  uint32_t* dCodes, void* stream)
{
    return LaunchMorton30(dCenters, count, dSceneAabb, dCodes, static_cast<hipStream_t>(stream));
}

size_t mray_radix_sort_temp_bytes(uint32_t count) { return RadixSortTempBytes(count); }

int mray_radix_sort_u64_u32(uint64_t* dKeys, uint32_t* dValues, uint32_t count, void* dTemp,
                            size_t tempBytes, void* stream)
{
    return LaunchRadixSort(dKeys, dValues, count, dTemp, tempBytes,
                           static_cast<hipStream_t>(stream));
}

size_t mray_lbvh_temp_bytes(uint32_t count) { return LbvhTempBytes(count); }

int mray_lbvh_build(const uint64_t* dSortedCodes, const uint32_t* dLeafOrder,
                    const float* dLeafAabbs, uint32_t count, uint32_t* dNodes, float* dNodeAabbs,
                    void* dTemp, size_t tempBytes, void* stream)
{
    return LaunchLbvhBuild(dSortedCodes, dLeafOrder, dLeafAabbs, count, dNodes, dNodeAabbs,
                           nullptr, dTemp, tempBytes, static_cast<hipStream_t>(stream));
}

int mray_accel_create(const MRaySceneDesc* hScene, MRayAccel** out)
{
    MRayAccelOptions o{};
    o.builder = MRAY_BUILDER_SAH;
    return mray_accel_create_ex(hScene, &o, out);
}

int mray_accel_create_ex(const MRaySceneDesc* hScene, const MRayAccelOptions* hOpt, MRayAccel** out)
{
    if(!out || !hOpt) { SetError("out or options is NULL"); return MRAY_ERR_INVALID; }
    if(hOpt->builder > MRAY_BUILDER_LBVH) { SetError("unknown builder"); return MRAY_ERR_INVALID; }
    const uint32_t width = hOpt->width ? hOpt->width : 4u;
    if(width != 4u && width != 8u) { SetError("BVH width must be 4 or 8"); return MRAY_ERR_INVALID; }
    *out = nullptr;
    MRAY_TRY(ValidateScene(hScene));
    MRayAccel* a = new MRayAccel();
    a->nTri = hScene->triangleCount;
    a->nVert = hScene->vertexCount;
    int r = BuildAccel(a, hScene, hOpt->builder, width, (hOpt->flags & MRAY_ACCEL_HOST_COLLAPSE) != 0u);
    if(r == MRAY_OK) r = UploadSceneData(a, hScene);
    if(r != MRAY_OK)
    {
        a->arena.Release();
        delete a;
        return r;
    }
    *out = a;
    return MRAY_OK;
}

int mray_accel_destroy(MRayAccel* accel)
{
    if(!accel) return MRAY_OK;
    (void)hipDeviceSynchronize();
    accel->arena.Release();
    delete accel;
    return MRAY_OK;
}

int mray_accel_read_wide(const MRayAccel* a, uint32_t* hNodes, uint32_t* hSoupPrims)
{
    if(!a) { SetError("accel is NULL"); return MRAY_ERR_INVALID; }
    if(a->width != 4 || !a->dNodes) { SetError("read_wide: not a 4-wide accelerator"); return MRAY_ERR_INVALID; }
    MRAY_HIP_CHECK(hipDeviceSynchronize());
    if(hNodes)
        MRAY_HIP_CHECK(hipMemcpy(hNodes, a->dNodes, size_t(a->wideNodes) * sizeof(BNode4),
                                 hipMemcpyDeviceToHost));
    if(hSoupPrims && a->nTri)
        MRAY_HIP_CHECK(hipMemcpy2D(hSoupPrims, 4, reinterpret_cast<const char*>(a->dSoup) + 12,
                                   sizeof(SoupTri), 4, a->nTri, hipMemcpyDeviceToHost));
    return MRAY_OK;
}

int mray_accel_info(const MRayAccel* a, uint32_t* hBinaryNodes, uint32_t* hWideNodes,
                    uint32_t* hLeafTriangles)
{
    if(!a) { SetError("accel is NULL"); return MRAY_ERR_INVALID; }
    if(hBinaryNodes) *hBinaryNodes = (a->nTri > 1) ? a->nTri - 1 : 1;
    if(hWideNodes) *hWideNodes = a->wideNodes;
    if(hLeafTriangles) *hLeafTriangles = a->nTri;
    return MRAY_OK;
}

int mray_accel_read_lbvh(const MRayAccel* a, uint64_t* hCodes, uint32_t* hLeafOrder,
                         uint32_t* hNodes, float* hNodeAabbs)
{
    if(!a) { SetError("accel is NULL"); return MRAY_ERR_INVALID; }
    size_t n = a->nTri, m = (n > 1) ? n - 1 : 1;
    if(hCodes) MRAY_HIP_CHECK(hipMemcpy(hCodes, a->dCodes, n * 8, hipMemcpyDeviceToHost));
    if(hLeafOrder)
        MRAY_HIP_CHECK(hipMemcpy(hLeafOrder, a->dLeafOrder, n * 4, hipMemcpyDeviceToHost));
    if(hNodes) MRAY_HIP_CHECK(hipMemcpy(hNodes, a->dBinNodes, m * 12, hipMemcpyDeviceToHost));
    if(hNodeAabbs)
        MRAY_HIP_CHECK(hipMemcpy(hNodeAabbs, a->dBinBoxes, m * 24, hipMemcpyDeviceToHost));
    return MRAY_OK;
}

int mray_accel_cast_closest(const MRayAccel* a, const float* dRays, uint32_t count,
                            uint32_t* dPrim, float* dT, float* dBary, void* stream)
{
    if(!a || (!dRays && count) || (!dPrim && count))
    {
        SetError("cast_closest: NULL argument");
        return MRAY_ERR_INVALID;
    }
    return LaunchCastClosest(a->ctx, dRays, count, dPrim, dT, dBary, nullptr,
                             static_cast<hipStream_t>(stream));
}

// Same cast, also accumulating {nodes visited, triangles tested} into
// dCounters[2] (u64); used to compute algorithmic traversal bytes.
int mray_accel_cast_closest_counted(const MRayAccel* a, const float* dRays, uint32_t count,
                                    uint32_t* dPrim, float* dT, float* dBary,
                                    unsigned long long* dCounters, void* stream)
{
    if(!a || !dCounters) { SetError("cast_closest_counted: NULL argument"); return MRAY_ERR_INVALID; }
    return LaunchCastClosest(a->ctx, dRays, count, dPrim, dT, dBary, dCounters,
                             static_cast<hipStream_t>(stream));
}

int mray_accel_cast_visibility(const MRayAccel* a, const float* dRays, uint32_t count,
                               uint32_t* dVisible, void* stream)
{
    if(!a || (!dRays && count) || (!dVisible && count))
    {
        SetError("cast_visibility: NULL argument");
        return MRAY_ERR_INVALID;
    }
    return LaunchCastVisibility(a->ctx, dRays, count, dVisible, static_cast<hipStream_t>(stream));
}

} // extern "C"
