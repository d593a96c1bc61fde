// Merged scene of a batch (mray_batch_scene_create): device-to-device copies
// of each scene's accelerator and shading tables into one set of arrays,
// then relocation kernels that shift every index by its scene's offsets.
#include "common.h"
#include "lbvh.h"
#include "scene.h"
#include "traverse.h"

#include <algorithm>
#include <vector>

namespace mray
{

// Child reference of a wide node: internal node index -> + nodeOff, leaf
// range start (soup position) -> + triOff, empty slot unchanged.
__device__ __forceinline__ uint32_t RelocChild(uint32_t c, uint32_t nodeOff, uint32_t triOff)
{
    if(c == kInvalid) return c;
    if(c & kLeafFlag) return (c & ~kLeafStartMask) | ((c & kLeafStartMask) + triOff);
    return c + nodeOff;
}

__global__ void KRelocNodes4(BNode4* __restrict__ n, uint32_t count, uint32_t nodeOff, uint32_t triOff)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= count) return;
    uint4 c = *reinterpret_cast<const uint4*>(n[i].child);
    c.x = RelocChild(c.x, nodeOff, triOff);
    c.y = RelocChild(c.y, nodeOff, triOff);
    c.z = RelocChild(c.z, nodeOff, triOff);
    c.w = RelocChild(c.w, nodeOff, triOff);
    *reinterpret_cast<uint4*>(n[i].child) = c;
}

__global__ void KRelocNodes8(CNode8* __restrict__ n, uint32_t count, uint32_t nodeOff, uint32_t triOff)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i >= count) return;
    for(int k = 0; k < 8; k++) n[i].child[k] = RelocChild(n[i].child[k], nodeOff, triOff);
}

__global__ void KRelocSoup(SoupTri* __restrict__ t, uint32_t count, uint32_t triOff)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < count) t[i].prim += triOff;
}

__global__ void KAddU32(uint32_t* __restrict__ p, size_t count, uint32_t add)
{
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if(i < count) p[i] += add;
}

// Surface keys: material index -> + matOff; light keys keep their
// scene-local light index (the view's light range is applied at shading).
__global__ void KRelocKeys(uint32_t* __restrict__ k, uint32_t count, uint32_t matOff)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if(i < count && !(k[i] & MRAY_LIGHT_FLAG)) k[i] += matOff;
}

int BuildBatchScene(MRayBatchScene* b, const MRayAccel* const* accels, uint32_t count)
{
    b->width = accels[0]->width;
    uint64_t nodes = 0, tris = 0, verts = 0, mats = 0, lights = 0;
    uint32_t maxDepth = 0;
    for(uint32_t k = 0; k < count; k++)
    {
        const MRayAccel* a = accels[k];
        if(a->width != b->width)
        {
            SetError("batch_scene_create: every accelerator must have the same BVH width");
            return MRAY_ERR_INVALID;
        }
        if(a->nTri == 0 || a->wideNodes == 0)
        {
            SetError("batch_scene_create: a scene has no triangles");
            return MRAY_ERR_INVALID;
        }
        MRayBatchScene::Part part;
        part.root = uint32_t(nodes);
        part.triOff = uint32_t(tris);
        part.nTri = a->nTri;
        part.vertOff = uint32_t(verts);
        part.matOff = uint32_t(mats);
        part.lightOff = uint32_t(lights);
        part.lightCount = a->view.lightCount;
        part.boundaryType = a->view.boundaryType;
        for(int c = 0; c < 3; c++) part.boundaryRad[c] = a->view.boundaryRad[c];
        part.sceneDiameter = a->view.sceneDiameter;
        part.camera = a->camera;
        MRAY_HIP_CHECK(hipMemcpy(part.box, a->dSceneBox, sizeof(part.box), hipMemcpyDeviceToHost));
        b->parts.push_back(part);
        nodes += a->wideNodes;
        tris += a->nTri;
        verts += a->nVert;
        mats += a->hMats.size();
        lights += a->view.lightCount;
        maxDepth = std::max(maxDepth, a->maxDepth);
    }
    for(uint32_t k = 0; k < count; k++)
        MRAY_TRY(CopyEnvironment(b->arena, accels[k]->view, b->parts[k].env));
    // leaf ranges hold a soup position in 26 bits; vertex indices are 32-bit
    if(tris > kLeafStartMask || nodes >= kLeafFlag || verts >= kInvalid || mats >= MRAY_LIGHT_FLAG)
    {
        SetError("batch_scene_create: merged scene exceeds 2^26 triangles or 2^31 nodes");
        return MRAY_ERR_INVALID;
    }
    b->nodes = uint32_t(nodes);
    b->tris = uint32_t(tris);
    b->verts = uint32_t(verts);
    b->mats = uint32_t(mats);
    b->lights = uint32_t(lights);
    DeviceArena& ar = b->arena;
    auto A = [&](void** p, size_t bytes) { return DevAlloc(ar, p, std::max<size_t>(bytes, 16)); };
    BNode4* dNodes = nullptr;
    CNode8* dCNodes = nullptr;
    SoupTri* dSoup = nullptr;
    float* dPos = nullptr;
    float4* dTbn = nullptr;
    uint32_t *dIdx = nullptr, *dKey = nullptr;
    float* dLightPos = nullptr;
    MRayMaterial* dMats = nullptr;
    float4* dLightRad = nullptr;
    int rv;
    if(b->width == 8) { if((rv = A((void**)&dCNodes, nodes * sizeof(CNode8))) != MRAY_OK) return rv; }
    else if((rv = A((void**)&dNodes, nodes * sizeof(BNode4))) != MRAY_OK) return rv;
    if((rv = A((void**)&dSoup, tris * sizeof(SoupTri))) != MRAY_OK) return rv;
    if((rv = A((void**)&dPos, verts * 3 * sizeof(float))) != MRAY_OK) return rv;
    if((rv = A((void**)&dTbn, verts * sizeof(float4))) != MRAY_OK) return rv;
    if((rv = A((void**)&dIdx, tris * 3 * sizeof(uint32_t))) != MRAY_OK) return rv;
    if((rv = A((void**)&dKey, tris * sizeof(uint32_t))) != MRAY_OK) return rv;
    if((rv = A((void**)&dMats, mats * sizeof(MRayMaterial))) != MRAY_OK) return rv;
    if((rv = A((void**)&dLightPos, lights * 9 * sizeof(float))) != MRAY_OK) return rv;
    if((rv = A((void**)&dLightRad, lights * sizeof(float4))) != MRAY_OK) return rv;
    const hipMemcpyKind D2D = hipMemcpyDeviceToDevice;
    for(uint32_t k = 0; k < count; k++)
    {
        const MRayAccel* a = accels[k];
        const MRayBatchScene::Part& p = b->parts[k];
        const uint32_t T = a->nTri, V = a->nVert, M = uint32_t(a->hMats.size()), L = p.lightCount;
        if(b->width == 8)
        {
            MRAY_HIP_CHECK(hipMemcpy(dCNodes + p.root, a->dCNodes, size_t(a->wideNodes) * sizeof(CNode8), D2D));
            KRelocNodes8<<<DivUp(a->wideNodes, 256), 256>>>(dCNodes + p.root, a->wideNodes, p.root, p.triOff);
        }
        else
        {
            MRAY_HIP_CHECK(hipMemcpy(dNodes + p.root, a->dNodes, size_t(a->wideNodes) * sizeof(BNode4), D2D));
            KRelocNodes4<<<DivUp(a->wideNodes, 256), 256>>>(dNodes + p.root, a->wideNodes, p.root, p.triOff);
        }
        MRAY_HIP_CHECK(hipMemcpy(dSoup + p.triOff, a->dSoup, size_t(T) * sizeof(SoupTri), D2D));
        KRelocSoup<<<DivUp(T, 256), 256>>>(dSoup + p.triOff, T, p.triOff);
        MRAY_HIP_CHECK(hipMemcpy(dPos + size_t(p.vertOff) * 3, a->dPos, size_t(V) * 3 * sizeof(float), D2D));
        MRAY_HIP_CHECK(hipMemcpy(dTbn + p.vertOff, a->dTbn, size_t(V) * sizeof(float4), D2D));
        MRAY_HIP_CHECK(hipMemcpy(dIdx + size_t(p.triOff) * 3, a->dIdx, size_t(T) * 3 * sizeof(uint32_t), D2D));
        KAddU32<<<DivUp(size_t(T) * 3, 256), 256>>>(dIdx + size_t(p.triOff) * 3, size_t(T) * 3, p.vertOff);
        MRAY_HIP_CHECK(hipMemcpy(dKey + p.triOff, a->dSurfKey, size_t(T) * sizeof(uint32_t), D2D));
        KRelocKeys<<<DivUp(T, 256), 256>>>(dKey + p.triOff, T, p.matOff);
        MRAY_HIP_CHECK(hipMemcpy(dMats + p.matOff, a->dMats, size_t(M) * sizeof(MRayMaterial), D2D));
        if(L)
        {
            MRAY_HIP_CHECK(hipMemcpy(dLightPos + size_t(p.lightOff) * 9, a->dLightPos,
                                     size_t(L) * 9 * sizeof(float), D2D));
            MRAY_HIP_CHECK(hipMemcpy(dLightRad + p.lightOff, a->dLightRad, size_t(L) * sizeof(float4), D2D));
        }
        MRAY_HIP_CHECK(hipGetLastError());
    }
    MRAY_HIP_CHECK(hipDeviceSynchronize());
    SceneView& v = b->view;
    v.pos = dPos;
    v.tbn = dTbn;
    v.idx = dIdx;
    v.surfKey = dKey;
    v.mats = dMats;
    v.lightPos = dLightPos;
    v.lightRad = dLightRad;
    v.lightCount = 0;           // per view
    v.boundaryType = MRAY_BOUNDARY_NULL;
    v.sceneDiameter = 0.0f;
    // traversal context over the merged tree; spill stack for the deepest scene
    int dev = 0;
    MRAY_HIP_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    MRAY_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
    const uint32_t cus = uint32_t(std::max(1, prop.multiProcessorCount));
    TraceCtx& c = b->ctx;
    c.gridThreads = cus * 8u * kTraceTPB;
    const uint32_t need = (b->width - 1u) * maxDepth + 4u;
    c.gstackDepth = (need > kLdsStack) ? need - kLdsStack : 1u;
    if((rv = DevAlloc(ar, (void**)&c.gstack, size_t(c.gstackDepth) * c.gridThreads * sizeof(uint2))) != MRAY_OK)
        return rv;
    c.maxVisits = b->nodes + 1u;
    c.nodes = dNodes;
    c.cnodes = dCNodes;
    c.soup = dSoup;
    c.width = b->width;
    return MRAY_OK;
}

} // namespace mray

using namespace mray;

extern "C" {

int mray_batch_scene_create(const MRayAccel* const* accels, uint32_t count, MRayBatchScene** out)
{
    if(!accels || !out || count == 0)
    {
        SetError("batch_scene_create: NULL argument or no scenes");
        return MRAY_ERR_INVALID;
    }
    *out = nullptr;
    for(uint32_t k = 0; k < count; k++)
        if(!accels[k])
        {
            SetError("batch_scene_create: NULL accelerator");
            return MRAY_ERR_INVALID;
        }
    MRayBatchScene* b = new MRayBatchScene();
    int rv = BuildBatchScene(b, accels, count);
    if(rv != MRAY_OK)
    {
        (void)hipDeviceSynchronize();
        b->arena.Release();
        delete b;
        return rv;
    }
    *out = b;
    return MRAY_OK;
}

int mray_batch_scene_destroy(MRayBatchScene* b)
{
    if(!b) return MRAY_OK;
    (void)hipDeviceSynchronize();
    b->arena.Release();
    delete b;
    return MRAY_OK;
}

int mray_batch_scene_info(const MRayBatchScene* b, uint32_t* hScenes, uint32_t* hTriangles,
                          uint32_t* hNodes, uint32_t* hWidth)
{
    if(!b) { SetError("batch scene is NULL"); return MRAY_ERR_INVALID; }
    if(hScenes) *hScenes = uint32_t(b->parts.size());
    if(hTriangles) *hTriangles = b->tris;
    if(hNodes) *hNodes = b->nodes;
    if(hWidth) *hWidth = b->width;
    return MRAY_OK;
}

} // extern "C"
