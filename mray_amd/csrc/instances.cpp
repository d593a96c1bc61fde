// Instanced scenes (mray_instanced_scene_create): meshes merged once, a
// 4-wide tree over the instances' world boxes built with the binned-SAH
// builder, the instance table and the world-space light table.
#include "common.h"
#include "lbvh.h"
#include "scene.h"
#include "traverse.h"

#include <algorithm>
#include <cmath>
#include <vector>


namespace mray
{
namespace
{
struct Quat { double w, x, y, z; };

void Rotate(const Quat& q, const double v[3], double out[3])
{
    // q v q* for a unit quaternion
    const double ux = q.x, uy = q.y, uz = q.z;
    const double tx = 2.0 * (uy * v[2] - uz * v[1]);
    const double ty = 2.0 * (uz * v[0] - ux * v[2]);
    const double tz = 2.0 * (ux * v[1] - uy * v[0]);
    out[0] = v[0] + q.w * tx + (uy * tz - uz * ty);
    out[1] = v[1] + q.w * ty + (uz * tx - ux * tz);
    out[2] = v[2] + q.w * tz + (ux * ty - uy * tx);
}

int BuildInstanced(MRayInstScene* sc, const MRayAccel* const* meshes, uint32_t meshCount,
                   const MRayInstance* inst, uint32_t n, const MRayInstEnv& env)
{
    MRayBatchScene& mb = sc->meshes;
    MRAY_TRY(BuildBatchScene(&mb, meshes, meshCount));
    if(mb.width != 4)
    {
        SetError("instanced_scene_create: meshes need 4-wide trees");
        return MRAY_ERR_INVALID;
    }
    // mesh bounds and light vertices (host copies)
    std::vector<float> meshBox(size_t(meshCount) * 6);
    std::vector<std::vector<float>> meshLightPos(meshCount);
    std::vector<std::vector<float4>> meshLightRad(meshCount);
    uint32_t maxMeshDepth = 0;
    for(uint32_t m = 0; m < meshCount; m++)
    {
        MRAY_HIP_CHECK(hipMemcpy(&meshBox[size_t(m) * 6], meshes[m]->dSceneBox, 6 * sizeof(float),
                                 hipMemcpyDeviceToHost));
        const uint32_t L = meshes[m]->view.lightCount;
        meshLightPos[m].resize(size_t(L) * 9);
        meshLightRad[m].resize(L);
        if(L)
        {
            MRAY_HIP_CHECK(hipMemcpy(meshLightPos[m].data(), meshes[m]->dLightPos, size_t(L) * 9 * 4,
                                     hipMemcpyDeviceToHost));
            MRAY_HIP_CHECK(hipMemcpy(meshLightRad[m].data(), meshes[m]->dLightRad, size_t(L) * 16,
                                     hipMemcpyDeviceToHost));
        }
        maxMeshDepth = std::max(maxMeshDepth, meshes[m]->maxDepth);
    }
    // instance table, world boxes, world lights
    std::vector<InstGPU> table(n);
    std::vector<float> boxes(size_t(n) * 6);
    std::vector<float> lightPos;
    std::vector<float4> lightRad;
    float world[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    uint64_t effective = 0;
    for(uint32_t i = 0; i < n; i++)
    {
        const MRayInstance& in = inst[i];
        if(in.mesh >= meshCount || !(in.scale[0] > 0.0f) || !(in.scale[1] > 0.0f) ||
           !(in.scale[2] > 0.0f))
        {
            SetError("instanced_scene_create: instance " + std::to_string(i) +
                     " names a missing mesh or has a non-positive scale");
            return MRAY_ERR_INVALID;
        }
        Quat q{in.rotation[0], in.rotation[1], in.rotation[2], in.rotation[3]};
        const double len = std::sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
        if(!(len > 0.0))
        {
            SetError("instanced_scene_create: zero rotation quaternion");
            return MRAY_ERR_INVALID;
        }
        q = {q.w / len, q.x / len, q.y / len, q.z / len};
        const double sc_[3] = {in.scale[0], in.scale[1], in.scale[2]};
        const bool uniform = (in.scale[0] == in.scale[1]) && (in.scale[1] == in.scale[2]);
        const double t[3] = {in.translation[0], in.translation[1], in.translation[2]};
        InstGPU& g = table[i];
        g.rot = make_float4(float(q.w), float(q.x), float(q.y), float(q.z));
        g.transScale = make_float4(float(t[0]), float(t[1]), float(t[2]), uniform ? 1.0f : 0.0f);
        g.scale = make_float4(in.scale[0], in.scale[1], in.scale[2], 0.0f);
        // world -> object: x_o = R^-1 (x_w - t) / s
        const Quat qi{q.w, -q.x, -q.y, -q.z};
        double rows[3][4];
        for(int c = 0; c < 3; c++)
        {
            double e[3] = {0, 0, 0};
            e[c] = 1.0;
            double col[3];
            Rotate(qi, e, col);
            for(int r = 0; r < 3; r++) rows[r][c] = col[r] / sc_[r];
        }
        for(int r = 0; r < 3; r++)
            rows[r][3] = -(rows[r][0] * t[0] + rows[r][1] * t[1] + rows[r][2] * t[2]);
        for(int r = 0; r < 3; r++)
            g.inv[r] = make_float4(float(rows[r][0]), float(rows[r][1]), float(rows[r][2]),
                                   float(rows[r][3]));
        g.root = mb.parts[in.mesh].root;
        g.lightOff = uint32_t(lightRad.size());
        g.mesh = in.mesh;
        g.flags = env.flags & MRAY_INST_ROTATE_FRAMES;
        // world box of the mesh box's corners
        const float* mbx = &meshBox[size_t(in.mesh) * 6];
        float* wb = &boxes[size_t(i) * 6];
        for(int k = 0; k < 3; k++) { wb[k] = INFINITY; wb[3 + k] = -INFINITY; }
        for(int corner = 0; corner < 8; corner++)
        {
            const double p[3] = {mbx[(corner & 1) ? 3 : 0] * sc_[0], mbx[(corner & 2) ? 4 : 1] * sc_[1],
                                  mbx[(corner & 4) ? 5 : 2] * sc_[2]};
            double w[3];
            Rotate(q, p, w);
            for(int k = 0; k < 3; k++)
            {
                const float v = float(w[k] + t[k]);
                // widen by one float ulp each way: the box stays conservative
                wb[k] = std::min(wb[k], std::nextafter(v, -INFINITY));
                wb[3 + k] = std::max(wb[3 + k], std::nextafter(v, INFINITY));
            }
        }
        for(int k = 0; k < 3; k++)
        {
            world[k] = std::min(world[k], wb[k]);
            world[3 + k] = std::max(world[3 + k], wb[3 + k]);
        }
        // this instance's lights in world space
        const std::vector<float>& lp = meshLightPos[in.mesh];
        for(size_t l = 0; l < meshLightRad[in.mesh].size(); l++)
        {
            for(int v = 0; v < 3; v++)
            {
                const double p[3] = {lp[l * 9 + v * 3] * sc_[0], lp[l * 9 + v * 3 + 1] * sc_[1],
                                     lp[l * 9 + v * 3 + 2] * sc_[2]};
                double w[3];
                Rotate(q, p, w);
                for(int k = 0; k < 3; k++) lightPos.push_back(float(w[k] + t[k]));
            }
            lightRad.push_back(meshLightRad[in.mesh][l]);
        }
        effective += meshes[in.mesh]->nTri;
    }
    // tree over the instances
    std::vector<uint32_t> order;
    WideBvh w;
    if(n > 1)
    {
        std::vector<uint32_t> bn, br;
        std::vector<float> bb;
        if(!BuildSahBinary(boxes, n, bn, bb, br, order))
        {
            SetError("instanced_scene_create: instance tree construction failed");
            return MRAY_ERR_INVALID;
        }
        // one instance per leaf: each instance's own world box is tested in
        // its parent before its mesh tree is entered
        w = CollapseToBvh4(bn, bb, br, boxes, order, 1u);
    }
    else
    {
        order.assign(1, 0u);
        BNode4 root{};
        for(int k = 0; k < 4; k++)
        {
            root.lox[k] = root.loy[k] = root.loz[k] = INFINITY;
            root.hix[k] = root.hiy[k] = root.hiz[k] = -INFINITY;
            root.child[k] = kInvalid;
        }
        root.lox[0] = boxes[0]; root.loy[0] = boxes[1]; root.loz[0] = boxes[2];
        root.hix[0] = boxes[3]; root.hiy[0] = boxes[4]; root.hiz[0] = boxes[5];
        root.child[0] = kLeafFlag | 0u;     // one instance at leaf position 0
        w.nodes.push_back(root);
        w.maxDepth = 1;
    }
    DeviceArena& ar = sc->arena;
    BNode4* dTlas = nullptr;
    uint32_t *dOrder = nullptr;
    InstGPU* dInst = nullptr;
    float* dLightPos = nullptr;
    float4* dLightRad = nullptr;
    auto up = [&](void** d, const void* h, size_t bytes) -> int
    {
        int r = DevAlloc(ar, d, std::max<size_t>(bytes, 16));
        if(r != MRAY_OK) return r;
        if(bytes) MRAY_HIP_CHECK(hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice));
        return MRAY_OK;
    };
    MRAY_TRY(up((void**)&dTlas, w.nodes.data(), w.nodes.size() * sizeof(BNode4)));
    MRAY_TRY(up((void**)&dOrder, order.data(), order.size() * 4));
    MRAY_TRY(up((void**)&dInst, table.data(), table.size() * sizeof(InstGPU)));
    MRAY_TRY(up((void**)&dLightPos, lightPos.data(), lightPos.size() * 4));
    MRAY_TRY(up((void**)&dLightRad, lightRad.data(), lightRad.size() * sizeof(float4)));
    sc->instanceCount = n;
    sc->tlasNodes = uint32_t(w.nodes.size());
    sc->tlasDepth = w.maxDepth;
    sc->effectiveTriangles = effective;
    for(int k = 0; k < 6; k++) sc->box[k] = world[k];
    sc->camera = env.camera;
    // shading view: the merged mesh arrays with the world light table
    SceneView& v = sc->view;
    v = mb.view;
    v.lightPos = dLightPos;
    v.lightRad = dLightRad;
    v.lightCount = uint32_t(lightRad.size());
    v.boundaryType = env.boundaryType;
    for(int k = 0; k < 3; k++) v.boundaryRad[k] = env.boundaryRadiance[k];
    // the boundary light's diameter: the world box's XZ span, as the
    // reference derives it (TracerBase.cpp:1651-1664), unless given
    const float dx = world[3] - world[0], dz = world[5] - world[2];
    v.sceneDiameter = (env.sceneDiameter > 0.0f) ? env.sceneDiameter : std::sqrt(dx * dx + dz * dz);
    // traversal: the mesh trees plus the instance tree on one stack
    TraceCtx& c = sc->ctx;
    c = mb.ctx;
    const uint32_t need = 3u * (w.maxDepth + maxMeshDepth) + 4u;
    c.gstackDepth = (need > kLdsStack) ? need - kLdsStack : 1u;
    MRAY_TRY(DevAlloc(ar, (void**)&c.gstack, size_t(c.gstackDepth) * c.gridThreads * sizeof(uint2)));
    c.tlas = dTlas;
    c.tlasOrder = dOrder;
    c.inst = dInst;
    c.tlasVisits = sc->tlasNodes + 1u;
    // a walk enters each instance at most once and visits each node of its
    // tree at most once while inside
    c.instWalkVisits = uint32_t(std::min<uint64_t>(0xFFFFFFFFull,
                                                   uint64_t(c.tlasVisits) + uint64_t(n) * c.maxVisits));
    return MRAY_OK;
}
} // namespace
} // namespace mray

using namespace mray;

extern "C" {

int mray_instanced_scene_create(const MRayAccel* const* meshes, uint32_t meshCount,
                                const MRayInstance* hInstances, uint32_t instanceCount,
                                const MRayInstEnv* hEnv, MRayInstScene** out)
{
    if(!meshes || !hInstances || !hEnv || !out || meshCount == 0 || instanceCount == 0)
    {
        SetError("instanced_scene_create: NULL argument, no meshes or no instances");
        return MRAY_ERR_INVALID;
    }
    *out = nullptr;
    if(hEnv->boundaryType > MRAY_BOUNDARY_SKYSPHERE_SPHERICAL)
    {
        SetError("instanced_scene_create: unknown boundaryType");
        return MRAY_ERR_INVALID;
    }
    MRayInstScene* sc = new MRayInstScene();
    int rv = BuildInstanced(sc, meshes, meshCount, hInstances, instanceCount, *hEnv);
    if(rv != MRAY_OK)
    {
        (void)hipDeviceSynchronize();
        sc->arena.Release();
        sc->meshes.arena.Release();
        delete sc;
        return rv;
    }
    *out = sc;
    return MRAY_OK;
}

int mray_instanced_scene_destroy(MRayInstScene* sc)
{
    if(!sc) return MRAY_OK;
    (void)hipDeviceSynchronize();
    sc->arena.Release();
    sc->meshes.arena.Release();
    delete sc;
    return MRAY_OK;
}

int mray_instanced_scene_info(const MRayInstScene* sc, uint32_t* hInstances,
                              uint64_t* hEffectiveTriangles, uint32_t* hLights, uint32_t* hTlasNodes)
{
    if(!sc) { SetError("instanced scene is NULL"); return MRAY_ERR_INVALID; }
    if(hInstances) *hInstances = sc->instanceCount;
    if(hEffectiveTriangles) *hEffectiveTriangles = sc->effectiveTriangles;
    if(hLights) *hLights = sc->view.lightCount;
    if(hTlasNodes) *hTlasNodes = sc->tlasNodes;
    return MRAY_OK;
}

} // extern "C"
