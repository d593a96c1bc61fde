// Upload of the shading-side scene tables (materials, vertex TBN quaternions,
// surface keys, emissive triangles, boundary light).
#include "scene.h"

#include <algorithm>

#include <cstring>
#include <vector>

namespace mray
{

int UploadSceneData(MRayAccel* a, const MRaySceneDesc* s)
{
    DeviceArena& ar = a->arena;
    const size_t V = s->vertexCount, T = s->triangleCount;
    std::vector<float4> tbn(V);
    for(size_t v = 0; v < V; v++)
    {
        if(s->hTbnQuats)
            tbn[v] = make_float4(s->hTbnQuats[4 * v], s->hTbnQuats[4 * v + 1],
                                 s->hTbnQuats[4 * v + 2], s->hTbnQuats[4 * v + 3]);
        else
            tbn[v] = make_float4(1, 0, 0, 0);
    }
    std::vector<uint32_t> keys(T, 0u);
    if(s->hSurfaceKeys) std::memcpy(keys.data(), s->hSurfaceKeys, T * 4);
    std::vector<MRayMaterial> mats(s->hMaterials, s->hMaterials + s->materialCount);
    if(mats.empty()) mats.push_back(MRayMaterial{});
    std::vector<uint32_t> ltri(s->hLightTriangles, s->hLightTriangles + s->lightCount);
    if(ltri.empty()) ltri.push_back(0);
    // light triangles' world-space vertices, read directly by light sampling
    std::vector<float> lpos(size_t(std::max<uint32_t>(1u, s->lightCount)) * 9, 0.0f);
    for(uint32_t l = 0; l < s->lightCount; l++)
        for(int v = 0; v < 3; v++)
        {
            const uint32_t vi = s->hIndices[3 * size_t(s->hLightTriangles[l]) + v];
            for(int c = 0; c < 3; c++) lpos[size_t(l) * 9 + v * 3 + c] = s->hPositions[3 * size_t(vi) + c];
        }
    std::vector<float4> lrad(std::max<uint32_t>(1u, s->lightCount));
    for(uint32_t l = 0; l < s->lightCount; l++)
    {
        bool two = s->hLightTwoSided ? (s->hLightTwoSided[l] != 0) : false;
        lrad[l] = make_float4(s->hLightRadiance[3 * l], s->hLightRadiance[3 * l + 1],
                              s->hLightRadiance[3 * l + 2], two ? 1.0f : 0.0f);
    }
    auto up = [&](void** d, const void* h, size_t bytes) -> int
    {
        int r = DevAlloc(ar, d, bytes);
        if(r != MRAY_OK) return r;
        MRAY_HIP_CHECK(hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice));
        return MRAY_OK;
    };
    int r;
    if((r = up((void**)&a->dTbn, tbn.data(), V * sizeof(float4))) != MRAY_OK) return r;
    if((r = up((void**)&a->dSurfKey, keys.data(), T * 4)) != MRAY_OK) return r;
    if((r = up((void**)&a->dMats, mats.data(), mats.size() * sizeof(MRayMaterial))) != MRAY_OK)
        return r;
    if((r = up((void**)&a->dLightTri, ltri.data(), ltri.size() * 4)) != MRAY_OK) return r;
    if((r = up((void**)&a->dLightRad, lrad.data(), lrad.size() * sizeof(float4))) != MRAY_OK)
        return r;
    if((r = up((void**)&a->dLightPos, lpos.data(), lpos.size() * sizeof(float))) != MRAY_OK)
        return r;
    a->hMats = mats;
    a->camera = s->camera;
    SceneView& v = a->view;
    v.pos = a->dPos;
    v.tbn = a->dTbn;
    v.idx = a->dIdx;
    v.surfKey = a->dSurfKey;
    v.mats = a->dMats;
    v.lightPos = a->dLightPos;
    v.lightRad = a->dLightRad;
    v.lightCount = s->lightCount;
    v.boundaryType = s->boundaryType;
    v.boundaryRad[0] = s->boundaryRadiance[0];
    v.boundaryRad[1] = s->boundaryRadiance[1];
    v.boundaryRad[2] = s->boundaryRadiance[2];
    v.sceneDiameter = s->sceneDiameter;
    return MRAY_OK;
}

} // namespace mray
