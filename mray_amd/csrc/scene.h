// Device-resident scene: geometry + LBVH/4-wide BVH + shading tables.
#pragma once

#include "common.h"
#include "lbvh.h"
#include "traverse.h"

#include <vector>

namespace mray
{

struct DeviceArena
{
    std::vector<void*> ptrs;
    void Release();
};

int DevAlloc(DeviceArena& arena, void** p, size_t bytes);

// Shading-side view of the scene, passed by value to kernels.
struct SceneView
{
    const float* __restrict__ pos;        // [V][3]
    const float4* __restrict__ tbn;       // [V] (w,x,y,z)
    const uint32_t* __restrict__ idx;     // [T][3]
    const uint32_t* __restrict__ surfKey; // [T]
    const MRayMaterial* __restrict__ mats;
    const float* __restrict__ lightPos;   // [lightCount][9] world-space vertices
    const float4* __restrict__ lightRad;  // rgb, w = two-sided flag
    uint32_t lightCount;                  // emissive triangles (boundary excluded)
    uint32_t boundaryType;
    float boundaryRad[3];
    float sceneDiameter;
    // textured Skysphere (envW != 0; else boundaryRad is constant): radiance
    // texels in concentric-octahedral uv, row y = v, and its luminance
    // distribution as normalized row CDFs plus the row-marginal CDF
    const float4* __restrict__ envTex;    // [envH][envW] rgb, w unused
    const float* __restrict__ envCdfX;    // [envH][envW]
    const float* __restrict__ envCdfY;    // [envH]
    uint32_t envW, envH;
    // light selection by power (null: uniform): per light {prob bits, alias}
    // of its range's alias table, and its selection probability
    const uint2* __restrict__ lightAlias;
    const float* __restrict__ lightPmf;
    // spectral renderers: per material its albedo's sigmoid coefficients,
    // per light its radiance coefficients and scale (null in RGB mode)
    const float4* __restrict__ matCoeff;
    const float4* __restrict__ lightCoeff;
};

// Alias table of n lights (pos9 [n][9] world vertices, rad [n] rgb + two-sided
// flag) by emitted power (lightpower.cpp).
void LightPowerTable(const float* pos9, const float4* rad, uint32_t n, float* prob, uint32_t* alias,
                     float* pmf);
// Tables for the light ranges [rangeOff, +rangeCount) of v's light arrays
// (total lights), uploaded into `ar`.
int BuildLightPower(DeviceArena& ar, const SceneView& v, const std::vector<uint32_t>& rangeOff,
                    const std::vector<uint32_t>& rangeCount, uint32_t total, uint2** outAlias,
                    float** outPmf);

// Builds the texture and luminance CDFs of a textured Skysphere into `ar`
// and points `v` at them (envmap.cpp).
int BuildEnvironment(DeviceArena& ar, SceneView& v, uint32_t W, uint32_t H, const float* hRgb);
// Device copy of v's environment into `ar` (a merged scene's own copy).
int CopyEnvironment(DeviceArena& ar, const SceneView& from, SceneView& to);

} // namespace mray

struct MRayAccel
{
    mray::DeviceArena arena;
    uint32_t nTri = 0, nVert = 0;
    uint32_t wideNodes = 0, leafSlots = 0, maxDepth = 0;
    uint32_t strictTie = 0;       // 1: rebuilt with the (code, index) tie-break
    uint32_t builder = 0;         // MRAY_BUILDER_* of the traversal tree
    uint32_t width = 4;           // 4 (BNode4) or 8 (CNode8)
    float* dPos = nullptr;
    uint32_t* dIdx = nullptr;
    float* dCenters = nullptr;
    float* dAabbs = nullptr;
    float* dSceneBox = nullptr;
    uint64_t* dCodes = nullptr;
    uint64_t* dSortedCodes = nullptr;
    uint32_t* dLeafOrder = nullptr;
    uint32_t* dBinNodes = nullptr;
    float* dBinBoxes = nullptr;
    uint32_t* dBinRange = nullptr;
    mray::BNode4* dNodes = nullptr;
    mray::CNode8* dCNodes = nullptr;
    mray::SoupTri* dSoup = nullptr;
    mray::TraceCtx ctx{};
    // shading tables
    float4* dTbn = nullptr;
    uint32_t* dSurfKey = nullptr;
    MRayMaterial* dMats = nullptr;
    uint32_t* dLightTri = nullptr;
    float4* dLightRad = nullptr;
    float* dLightPos = nullptr;
    mray::SceneView view{};
    MRayCamera camera{};
    std::vector<MRayMaterial> hMats;
};

// Several scenes merged for one batch of viewports: every scene's triangle
// soup, traversal tree, vertices and shading tables concatenated on the device
// with their indices relocated (tree children, leaf ranges, soup triangle
// ids, vertex indices, material keys; light vertices are world-space), so one traversal
// context and one SceneView serve all of them; a scene is entered at its
// own root node.
struct MRayBatchScene
{
    struct Part
    {
        uint32_t root = 0;          // first node of the scene's tree
        uint32_t triOff = 0, nTri = 0;
        uint32_t vertOff = 0, matOff = 0;
        uint32_t lightOff = 0, lightCount = 0;
        uint32_t boundaryType = 0;
        float boundaryRad[3] = {0, 0, 0};
        float sceneDiameter = 0;
        MRayCamera camera{};
        float box[6] = {0, 0, 0, 0, 0, 0};   // scene box (min xyz, max xyz)
        // textured Skysphere of the scene (the merged scene's own copy)
        mray::SceneView env{};
    };
    mray::DeviceArena arena;
    std::vector<Part> parts;
    uint32_t width = 4;
    uint32_t nodes = 0, tris = 0, verts = 0, mats = 0, lights = 0;
    mray::SceneView view{};
    mray::TraceCtx ctx{};
};

// Instances of meshes (mray_instanced_scene_create): the meshes merged as
// for a batch (each keeps its own tree and root), a 4-wide tree over the
// instances' world boxes, the instance table and the world-space light
// table (every emissive triangle of every instance).
struct MRayInstScene
{
    MRayBatchScene meshes;
    mray::DeviceArena arena;
    uint32_t instanceCount = 0;
    uint32_t tlasNodes = 0, tlasDepth = 0;
    uint64_t effectiveTriangles = 0;
    float box[6] = {0, 0, 0, 0, 0, 0};   // world bounds
    MRayCamera camera{};
    mray::SceneView view{};
    mray::TraceCtx ctx{};
};

namespace mray
{
int UploadSceneData(MRayAccel* a, const MRaySceneDesc* s);
int BuildBatchScene(MRayBatchScene* b, const MRayAccel* const* accels, uint32_t count);
}
