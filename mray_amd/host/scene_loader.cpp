// MRay JSON scene files -> TracerI calls (the role of the reference's
// SceneLoaderMRay, SceneLoaderMRay/SceneLoaderMRay.cpp; format:
// Docs/markdown/scene/mrayScene.md, field names SceneLoaderMRay/NodeNames.h).
//
// Loads what the path tracer supports: inline "Triangle" primitives (tags
// nodeTriangle / nodeTriangleIndexed), Lambert / Reflect / Refract / Unreal
// materials with constant parameters, Null / Primitive / Skysphere_CoOcta /
// Skysphere_Spherical lights with constant radiance, Pinhole cameras, Identity / Single
// transforms (layout "matrix" or "trs"), and the Surfaces / LightSurfaces /
// CameraSurfaces / Boundary definitions.  Only entities the definitions use
// are loaded, as in the reference.  Anything else throws MRayError.
//
// Vertex normals become TBN quaternions q with R(q) n = +Z (the shortest arc;
// a normal facing -Z takes a half turn about X); the tangent frame is
// therefore fixed by the normal alone.
#include "mray_tracer.hpp"
#include "json.h"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <map>
#include <set>
#include <sstream>

namespace mray_amd
{
namespace
{

using json::Value;

[[noreturn]] void Unsupported_(const std::string& what)
{
    throw MRayError("scene: " + what + " is not supported by the MI355X path tracer");
}

std::vector<uint32_t> Ids(const Value& v)
{
    std::vector<uint32_t> out;
    if(v.IsArr())
        for(const Value& e : v.Arr()) out.push_back(e.U32());
    else out.push_back(v.U32());
    return out;
}

// i-th value of a field that is per-id when the struct holds several ids
const Value& PerId(const Value& node, std::string_view key, size_t i, size_t count)
{
    const Value& v = node.At(key);
    if(count > 1)
    {
        if(!v.IsArr() || v.Arr().size() != count)
            throw MRayError("scene: field \"" + std::string(key) + "\" must hold one entry per id");
        return v.Arr()[i];
    }
    return v;
}

Vector3 Vec3(const Value& v)
{
    if(v.IsObj()) throw MRayError("scene: texture-backed values are not supported");
    const auto& a = v.Arr();
    if(a.size() != 3) throw MRayError("scene: expected a 3-vector");
    return Vector3{{a[0].F32(), a[1].F32(), a[2].F32()}};
}

Vector2 Vec2(const Value& v)
{
    const auto& a = v.Arr();
    if(a.size() != 2) throw MRayError("scene: expected a 2-vector");
    return Vector2{{a[0].F32(), a[1].F32()}};
}

Quaternion TbnFromNormal(Vector3 n)
{
    double x = n.v[0], y = n.v[1], z = n.v[2];
    double len = std::sqrt(x * x + y * y + z * z);
    if(len > 0) { x /= len; y /= len; z /= len; }
    double q[4] = {1.0 + z, y, -x, 0.0};
    if(z < -1.0 + 1e-7) { q[0] = 0; q[1] = 1; q[2] = 0; q[3] = 0; }
    double l = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    return Quaternion{{float(q[0] / l), float(q[1] / l), float(q[2] / l), float(q[3] / l)}};
}

// row-major 4x4 of T * Rz * Ry * Rx * S (rotation angles in degrees)
Matrix4x4 Trs(Vector3 t, Vector3 r, Vector3 s)
{
    const double k = 3.14159265358979323846 / 180.0;
    double cx = std::cos(r.v[0] * k), sx = std::sin(r.v[0] * k);
    double cy = std::cos(r.v[1] * k), sy = std::sin(r.v[1] * k);
    double cz = std::cos(r.v[2] * k), sz = std::sin(r.v[2] * k);
    double Rx[3][3] = {{1, 0, 0}, {0, cx, -sx}, {0, sx, cx}};
    double Ry[3][3] = {{cy, 0, sy}, {0, 1, 0}, {-sy, 0, cy}};
    double Rz[3][3] = {{cz, -sz, 0}, {sz, cz, 0}, {0, 0, 1}};
    double A[3][3], R[3][3];
    for(int i = 0; i < 3; i++)
        for(int j = 0; j < 3; j++)
        {
            A[i][j] = 0;
            for(int m = 0; m < 3; m++) A[i][j] += Ry[i][m] * Rx[m][j];
        }
    for(int i = 0; i < 3; i++)
        for(int j = 0; j < 3; j++)
        {
            R[i][j] = 0;
            for(int m = 0; m < 3; m++) R[i][j] += Rz[i][m] * A[m][j];
        }
    Matrix4x4 out;
    for(int i = 0; i < 3; i++)
    {
        for(int j = 0; j < 3; j++) out.m[i * 4 + j] = float(R[i][j] * s.v[j]);
        out.m[i * 4 + 3] = t.v[i];
    }
    out.m[12] = out.m[13] = out.m[14] = 0.0f;
    out.m[15] = 1.0f;
    return out;
}

template<class T>
TransientData Data(const std::vector<T>& v)
{
    TransientData d(std::in_place_type_t<T>{}, v.size());
    d.Push(std::span<const T>(v));
    return d;
}

struct Indexed
{
    // id -> (struct node, position inside the struct's id list, id count)
    std::map<uint32_t, std::tuple<const Value*, size_t, size_t>> byId;
    explicit Indexed(const Value* list)
    {
        if(!list) return;
        for(const Value& node : list->Arr())
        {
            std::vector<uint32_t> ids = Ids(node.At("id"));
            for(size_t i = 0; i < ids.size(); i++)
            {
                if(byId.count(ids[i])) throw MRayError("scene: duplicate id " + std::to_string(ids[i]));
                byId[ids[i]] = {&node, i, ids.size()};
            }
        }
    }
    const std::tuple<const Value*, size_t, size_t>& Get(uint32_t id, const char* what) const
    {
        auto it = byId.find(id);
        if(it == byId.end()) throw MRayError(std::string("scene: unknown ") + what + " id " + std::to_string(id));
        return it->second;
    }
};

} // namespace

SceneIds LoadSceneMRay(TracerI& tracer, const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    if(!f) throw MRayError("scene: cannot open " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    std::string text = ss.str();
    Value root = json::Parse(text);

    Indexed prims(root.Find("Primitives")), mats(root.Find("Materials")), trans(root.Find("Transforms")),
        lights(root.Find("Lights")), cams(root.Find("Cameras"));

    const Value* surfList = root.Find("Surfaces");
    const Value* lightSurfList = root.Find("LightSurfaces");
    const Value* camSurfList = root.Find("CameraSurfaces");
    const Value& boundary = root.At("Boundary");

    // ---- what the definitions use ------------------------------------------
    std::set<uint32_t> usedPrims, usedMats, usedTrans, usedLights, usedCams;
    auto transOf = [&](const Value& n) -> uint32_t
    {
        const Value* t = n.Find("transform");
        return t ? t->U32() : UINT32_MAX;
    };
    if(surfList)
        for(const Value& sn : surfList->Arr())
        {
            for(uint32_t p : Ids(sn.At("primitive"))) usedPrims.insert(p);
            for(uint32_t m : Ids(sn.At("material"))) usedMats.insert(m);
            if(transOf(sn) != UINT32_MAX) usedTrans.insert(transOf(sn));
        }
    if(lightSurfList)
        for(const Value& ln : lightSurfList->Arr())
        {
            usedLights.insert(ln.At("light").U32());
            if(transOf(ln) != UINT32_MAX) usedTrans.insert(transOf(ln));
        }
    if(camSurfList)
        for(const Value& cn : camSurfList->Arr())
        {
            usedCams.insert(cn.At("camera").U32());
            if(transOf(cn) != UINT32_MAX) usedTrans.insert(transOf(cn));
        }
    usedLights.insert(boundary.At("light").U32());
    if(transOf(boundary) != UINT32_MAX) usedTrans.insert(transOf(boundary));
    for(uint32_t l : usedLights)
    {
        auto [node, i, cnt] = lights.Get(l, "light");
        if(node->At("type").Str() == "Primitive")
            usedPrims.insert(PerId(*node, "primitive", i, cnt).U32());
    }

    // ---- primitives ------------------------------------------------------------
    PrimGroupId pg{0};
    std::map<uint32_t, PrimBatchId> primBatch;
    struct Mesh { std::vector<Vector3> pos; std::vector<Quaternion> tbn; std::vector<Vector2> uv; std::vector<Vector3ui> idx; };
    std::vector<std::pair<uint32_t, Mesh>> meshes;
    for(uint32_t p : usedPrims)
    {
        auto [node, i, cnt] = prims.Get(p, "primitive");
        if(node->At("type").Str() != "Triangle") Unsupported_("primitive type " + node->At("type").Str());
        const std::string& tag = node->At("tag").Str();
        if(tag != "nodeTriangle" && tag != "nodeTriangleIndexed") Unsupported_("primitive tag " + tag);
        Mesh m;
        for(const Value& v : PerId(*node, "position", i, cnt).Arr()) m.pos.push_back(Vec3(v));
        if(tag == "nodeTriangleIndexed")
            for(const Value& v : PerId(*node, "index", i, cnt).Arr())
            {
                const auto& a = v.Arr();
                if(a.size() != 3) throw MRayError("scene: index entries are 3-vectors");
                m.idx.push_back(Vector3ui{{a[0].U32(), a[1].U32(), a[2].U32()}});
            }
        else
        {
            if(m.pos.size() % 3) throw MRayError("scene: nodeTriangle positions must come in threes");
            for(uint32_t t = 0; t < m.pos.size() / 3; t++) m.idx.push_back(Vector3ui{{3 * t, 3 * t + 1, 3 * t + 2}});
        }
        for(const Vector3ui& t : m.idx)
            for(int c = 0; c < 3; c++)
                if(t.v[c] >= m.pos.size()) throw MRayError("scene: triangle index out of range");
        std::vector<Vector3> nrm(m.pos.size(), Vector3{{0, 0, 0}});
        const Value* nv = (cnt > 1) ? (node->Find("normal") ? &node->At("normal").Arr()[i] : nullptr)
                                    : node->Find("normal");
        if(nv)
        {
            const auto& a = nv->Arr();
            if(a.size() != m.pos.size()) throw MRayError("scene: one normal per position");
            for(size_t k = 0; k < a.size(); k++) nrm[k] = Vec3(a[k]);
        }
        else
        {
            // area-weighted face normals (as the reference does without normals)
            for(const Vector3ui& t : m.idx)
            {
                const float* p0 = m.pos[t.v[0]].v.data(); const float* p1 = m.pos[t.v[1]].v.data(); const float* p2 = m.pos[t.v[2]].v.data();
                float e0[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
                float e1[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
                float c[3] = {e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2], e0[0] * e1[1] - e0[1] * e1[0]};
                for(int k = 0; k < 3; k++)
                    for(int d = 0; d < 3; d++) nrm[t.v[k]].v[d] += c[d];
            }
        }
        for(const Vector3& n : nrm) m.tbn.push_back(TbnFromNormal(n));
        const Value* uv = (cnt > 1) ? (node->Find("uv") ? &node->At("uv").Arr()[i] : nullptr) : node->Find("uv");
        if(uv)
            for(const Value& v : uv->Arr()) m.uv.push_back(Vec2(v));
        m.uv.resize(m.pos.size(), Vector2{{0, 0}});
        meshes.emplace_back(p, std::move(m));
    }
    if(!meshes.empty())
    {
        pg = tracer.CreatePrimitiveGroup("(P)Triangle");
        std::vector<PrimCount> counts;
        for(const auto& [p, m] : meshes) counts.push_back(PrimCount{uint32_t(m.idx.size()), uint32_t(m.pos.size())});
        PrimBatchIdList b = tracer.ReservePrimitiveBatches(pg, counts);
        tracer.CommitPrimReservations(pg);
        for(size_t k = 0; k < meshes.size(); k++)
        {
            const Mesh& m = meshes[k].second;
            tracer.PushPrimAttribute(pg, b[k], 0, Data(m.pos));
            tracer.PushPrimAttribute(pg, b[k], 1, Data(m.tbn));
            tracer.PushPrimAttribute(pg, b[k], 2, Data(m.uv));
            tracer.PushPrimAttribute(pg, b[k], 3, Data(m.idx));
            primBatch[meshes[k].first] = b[k];
        }
    }

    // ---- materials -------------------------------------------------------------
    std::map<std::string, MatGroupId> matGroups;
    std::map<uint32_t, MaterialId> matId;
    std::vector<std::tuple<uint32_t, MatGroupId, MaterialId>> pending;
    for(uint32_t m : usedMats)
    {
        auto [node, i, cnt] = mats.Get(m, "material");
        std::string type = "(Mt)" + node->At("type").Str();
        if(!matGroups.count(type)) matGroups[type] = tracer.CreateMaterialGroup(type);
        MaterialId id = tracer.ReserveMaterial(matGroups[type], {});
        matId[m] = id;
        pending.emplace_back(m, matGroups[type], id);
    }
    for(auto& kv : matGroups) tracer.CommitMatReservations(kv.second);
    const auto texNone = std::vector<Optional<TextureId>>{std::nullopt};
    for(auto& [m, g, id] : pending)
    {
        auto [node, i, cnt] = mats.Get(m, "material");
        MatAttributeInfoList info = tracer.AttributeInfo(g);
        CommonIdRange r = {uint32_t(id), uint32_t(id)};
        for(uint32_t a = 0; a < info.size(); a++)
        {
            const Value* v = node->Find(info[a].name);
            if(!v)
            {
                if(info[a].isOptional == AttributeOptionality::MR_OPTIONAL) continue;
                throw MRayError("scene: material " + std::to_string(m) + " lacks \"" + info[a].name + "\"");
            }
            const Value& e = (cnt > 1) ? PerId(*node, info[a].name, i, cnt) : *v;
            if(e.IsStr() && e.Str() == "-") continue;
            if(info[a].isTexturable == AttributeTexturable::MR_TEXTURE_ONLY)
                throw MRayError("scene: texture-backed material attributes are not supported");
            TransientData d = (info[a].dataType == MRayDataEnum::MR_FLOAT)
                ? Data(std::vector<float>{e.F32()}) : Data(std::vector<Vector3>{Vec3(e)});
            if(info[a].isTexturable == AttributeTexturable::MR_CONSTANT_ONLY)
                tracer.PushMatAttribute(g, r, a, std::move(d));
            else
                tracer.PushMatAttribute(g, r, a, std::move(d), texNone);
        }
    }

    // ---- transforms -----------------------------------------------------------------
    std::map<uint32_t, TransformId> transId;
    TransGroupId single{0};
    bool haveSingle = false;
    std::vector<std::pair<TransformId, Matrix4x4>> matrices;
    for(uint32_t t : usedTrans)
    {
        auto [node, i, cnt] = trans.Get(t, "transform");
        const std::string& type = node->At("type").Str();
        if(type == "Identity") { transId[t] = TransformId(0); continue; }
        if(type != "Single") Unsupported_("transform type " + type);
        if(!haveSingle) { single = tracer.CreateTransformGroup("(T)Single"); haveSingle = true; }
        Matrix4x4 mtx;
        const std::string& layout = node->At("layout").Str();
        if(layout == "matrix")
        {
            const auto& a = PerId(*node, "matrix", i, cnt).Arr();
            if(a.size() != 16) throw MRayError("scene: matrix needs 16 values");
            for(int k = 0; k < 16; k++) mtx.m[k] = a[k].F32();
        }
        else if(layout == "trs")
        {
            auto opt = [&](const char* key, Vector3 def)
            {
                return node->Find(key) ? Vec3(PerId(*node, key, i, cnt)) : def;
            };
            mtx = Trs(opt("translate", Vector3{{0, 0, 0}}), opt("rotate", Vector3{{0, 0, 0}}),
                      opt("scale", Vector3{{1, 1, 1}}));
        }
        else Unsupported_("transform layout " + layout);
        TransformId id = tracer.ReserveTransformation(single, {});
        transId[t] = id;
        matrices.emplace_back(id, mtx);
    }
    if(haveSingle)
    {
        tracer.CommitTransReservations(single);
        for(auto& [id, mtx] : matrices)
            tracer.PushTransAttribute(single, {uint32_t(id), uint32_t(id)}, 0,
                                      Data(std::vector<Matrix4x4>{mtx}));
    }
    auto T = [&](const Value& n) { uint32_t t = transOf(n); return t == UINT32_MAX ? TransformId(0) : transId.at(t); };

    // ---- lights -------------------------------------------------------------------
    std::map<uint32_t, LightId> lightId;
    LightGroupId primLights{0}, sky{0}, skySph{0};
    bool havePrimLights = false, haveSky = false, haveSkySph = false;
    struct LightPush { LightGroupId g; LightId id; Vector3 rad; bool twoSided; bool prim; };
    std::vector<LightPush> lpush;
    for(uint32_t l : usedLights)
    {
        auto [node, i, cnt] = lights.Get(l, "light");
        const std::string& type = node->At("type").Str();
        if(type == "Null") { lightId[l] = LightId(0); continue; }
        if(type == "Primitive")
        {
            if(!havePrimLights) { primLights = tracer.CreateLightGroup("(L)Prim(P)Triangle", pg); havePrimLights = true; }
            uint32_t p = PerId(*node, "primitive", i, cnt).U32();
            LightId id = tracer.ReserveLight(primLights, {}, primBatch.at(p));
            lightId[l] = id;
            bool two = node->Find("isTwoSided") ? PerId(*node, "isTwoSided", i, cnt).Bool() : false;
            lpush.push_back({primLights, id, Vec3(PerId(*node, "radiance", i, cnt)), two, true});
            continue;
        }
        if(type == "Skysphere_CoOcta" || type == "Skysphere")
        {
            if(!haveSky) { sky = tracer.CreateLightGroup("(L)Skysphere_CoOcta"); haveSky = true; }
            LightId id = tracer.ReserveLight(sky, {});
            lightId[l] = id;
            lpush.push_back({sky, id, Vec3(PerId(*node, "radiance", i, cnt)), false, false});
            continue;
        }
        if(type == "Skysphere_Spherical")
        {
            if(!haveSkySph) { skySph = tracer.CreateLightGroup("(L)Skysphere_Spherical"); haveSkySph = true; }
            LightId id = tracer.ReserveLight(skySph, {});
            lightId[l] = id;
            lpush.push_back({skySph, id, Vec3(PerId(*node, "radiance", i, cnt)), false, false});
            continue;
        }
        Unsupported_("light type " + type);
    }
    if(havePrimLights) tracer.CommitLightReservations(primLights);
    if(haveSky) tracer.CommitLightReservations(sky);
    if(haveSkySph) tracer.CommitLightReservations(skySph);
    for(const LightPush& lp : lpush)
    {
        CommonIdRange r = {uint32_t(lp.id), uint32_t(lp.id)};
        tracer.PushLightAttribute(lp.g, r, 0, Data(std::vector<Vector3>{lp.rad}), texNone);
        if(lp.prim)
        {
            TransientData d(std::in_place_type_t<bool>{}, 1);
            d.Push(std::span<const bool>(&lp.twoSided, 1));
            tracer.PushLightAttribute(lp.g, r, 1, std::move(d));
        }
    }

    // ---- cameras (SceneLoaderMRay.cpp:595-662) -------------------------------------------
    std::map<uint32_t, CameraId> camId;
    if(!usedCams.empty())
    {
        CameraGroupId cg = tracer.CreateCameraGroup("(C)Pinhole");
        std::vector<std::pair<uint32_t, CameraId>> order;
        for(uint32_t c : usedCams) { CameraId id = tracer.ReserveCamera(cg, {}); camId[c] = id; order.emplace_back(c, id); }
        tracer.CommitCamReservations(cg);
        for(auto& [c, id] : order)
        {
            auto [node, i, cnt] = cams.Get(c, "camera");
            if(node->At("type").Str() != "Pinhole") Unsupported_("camera type " + node->At("type").Str());
            bool isFovX = PerId(*node, "isFovX", i, cnt).Bool();
            float fov = PerId(*node, "fov", i, cnt).F32();
            float aspect = PerId(*node, "aspect", i, cnt).F32();
            Vector2 planes = Vec2(PerId(*node, "planes", i, cnt));
            const float d2r = 3.14159265358979323846f / 180.0f;
            Vector4 fnp{{fov, fov, planes.v[0], planes.v[1]}};
            if(isFovX)
            {
                fnp.v[0] *= d2r;
                fnp.v[1] = 2.0f * std::atan(std::tan(fnp.v[0] * 0.5f) / aspect);
            }
            else
            {
                fnp.v[1] *= d2r;
                fnp.v[0] = 2.0f * std::atan(std::tan(fnp.v[1] * 0.5f) * aspect);
            }
            CommonIdRange r = {uint32_t(id), uint32_t(id)};
            tracer.PushCamAttribute(cg, r, 0, Data(std::vector<Vector4>{fnp}));
            tracer.PushCamAttribute(cg, r, 1, Data(std::vector<Vector3>{Vec3(PerId(*node, "gaze", i, cnt))}));
            tracer.PushCamAttribute(cg, r, 2, Data(std::vector<Vector3>{Vec3(PerId(*node, "position", i, cnt))}));
            tracer.PushCamAttribute(cg, r, 3, Data(std::vector<Vector3>{Vec3(PerId(*node, "up", i, cnt))}));
        }
    }

    // ---- definitions --------------------------------------------------------------------
    SceneIds out;
    if(surfList)
        for(const Value& sn : surfList->Arr())
        {
            SurfaceParams sp;
            std::vector<uint32_t> p = Ids(sn.At("primitive")), m = Ids(sn.At("material"));
            if(p.size() != m.size()) throw MRayError("scene: one material per primitive in a surface");
            for(size_t k = 0; k < p.size(); k++)
            {
                sp.primBatches.push_back(primBatch.at(p[k]));
                sp.materials.push_back(matId.at(m[k]));
            }
            sp.transformId = T(sn);
            const Value* cull = sn.Find("cullBackFace");
            for(size_t k = 0; k < p.size(); k++)
                sp.cullFaceFlags.push_back(cull ? (cull->IsArr() ? cull->Arr().at(k).Bool() : cull->Bool()) : true);
            if(const Value* am = sn.Find("alphaMap"))
            {
                auto bad = [](const Value& v) { return !(v.IsStr() && v.Str() == "-"); };
                if(am->IsArr() ? std::any_of(am->Arr().begin(), am->Arr().end(), bad) : bad(*am))
                    Unsupported_("alpha maps");
            }
            out.surfaces.push_back(tracer.CreateSurface(std::move(sp)));
        }
    if(lightSurfList)
        for(const Value& ln : lightSurfList->Arr())
            out.lightSurfaces.push_back(tracer.CreateLightSurface(
                LightSurfaceParams{lightId.at(ln.At("light").U32()), T(ln), {}}));
    if(camSurfList)
        for(const Value& cn : camSurfList->Arr())
            out.cameraSurfaces.push_back(tracer.CreateCameraSurface(
                CameraSurfaceParams{camId.at(cn.At("camera").U32()), T(cn), {}}));
    out.boundary = tracer.SetBoundarySurface(lightId.at(boundary.At("light").U32()), T(boundary));
    return out;
}

} // namespace mray_amd
