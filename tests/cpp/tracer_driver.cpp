// TracerI front-end driver for the tests: loads libmray_tracer.so with
// dlopen/dlsym (as the reference's SharedLibrary does, Core/SharedLibrary.h),
// constructs the tracer and feeds it a flat scene through the TracerI calls
// a scene loader makes (SceneLoaderMRay.cpp: groups -> reservations ->
// commits -> attribute pushes -> surfaces), then either
//   flatten <out>   writes CommitSurfaces' host-side scene (no GPU), or
//   render  <out> W H spp spectral mode rr0 rr1 seed slots renderMode burstSize [exr prefix]
//                   (slots = parallelizationHint, 0: W x H; renderMode
//                   Throughput or Latency)
//                   renders through SetupRenderEnv / CreateRenderer /
//                   StartRender / DoRenderWork with a consumer thread that
//                   plays the reference's run command: it waits on the
//                   shared TimelineSemaphore for each section's waitCounter,
//                   accumulates the section into double mean RGB + weight
//                   (RunCommand.cpp:609-700) and releases.  <out> receives
//                   the means and weights (4 x H x W floats); on triggerSave
//                   the image is also saved as <prefix>_<wpp>wpp.exr.
//
// Input scene file (little-endian):
//   u32 V, T, M, L, boundaryType; f32 boundaryRadiance[3], sceneDiameter;
//   f32 pos[V*3], tbn[V*4]; u32 idx[T*3], keys[T]; u8 cull[T];
//   material[M] (52 B: u32 type, f32 albedo[3], roughness, metallic,
//   specular, cauchyFront[3], cauchyBack[3]); u32 lightTris[L];
//   f32 lightRad[L*3]; u8 twoSided[L]; f32 camera[13] (position, gaze, up,
//   fov[2], nearFar[2]).
// A scene path ending in .json is read with LoadSceneMRay (MRay scene format)
// instead; the first camera surface renders.
// Triangles are grouped into maximal runs of one material (or of consecutive
// lights) and one cull flag; each run becomes one primitive batch and one
// surface (or light surface), in triangle order.
#include "mray_tracer.hpp"

#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>

using namespace mray_amd;

namespace
{

constexpr uint32_t kLightFlag = 0x80000000u;

struct Mat
{
    uint32_t type;
    float albedo[3], roughness, metallic, specular, cauchyFront[3], cauchyBack[3];
};
static_assert(sizeof(Mat) == 52);

struct SceneFile
{
    uint32_t V = 0, T = 0, M = 0, L = 0, boundaryType = 0;
    float boundaryRad[3] = {0, 0, 0}, diameter = 0;
    std::vector<float> pos, tbn, lightRad;
    std::vector<uint32_t> idx, keys, lightTris;
    std::vector<uint8_t> cull, twoSided;
    std::vector<Mat> mats;
    float cam[13] = {};
};

template<class T>
void ReadN(FILE* f, std::vector<T>& v, size_t n)
{
    v.resize(n);
    if(n && std::fread(v.data(), sizeof(T), n, f) != n)
        throw MRayError("scene file is truncated");
}

SceneFile Load(const char* path)
{
    FILE* f = std::fopen(path, "rb");
    if(!f) throw MRayError(std::string("cannot open ") + path);
    SceneFile s;
    uint32_t hdr[5];
    if(std::fread(hdr, 4, 5, f) != 5) throw MRayError("scene file is truncated");
    s.V = hdr[0]; s.T = hdr[1]; s.M = hdr[2]; s.L = hdr[3]; s.boundaryType = hdr[4];
    if(std::fread(s.boundaryRad, 4, 3, f) != 3 || std::fread(&s.diameter, 4, 1, f) != 1)
        throw MRayError("scene file is truncated");
    ReadN(f, s.pos, size_t(s.V) * 3);
    ReadN(f, s.tbn, size_t(s.V) * 4);
    ReadN(f, s.idx, size_t(s.T) * 3);
    ReadN(f, s.keys, s.T);
    ReadN(f, s.cull, s.T);
    ReadN(f, s.mats, s.M);
    ReadN(f, s.lightTris, s.L);
    ReadN(f, s.lightRad, size_t(s.L) * 3);
    ReadN(f, s.twoSided, s.L);
    if(std::fread(s.cam, 4, 13, f) != 13) throw MRayError("scene file is truncated");
    std::fclose(f);
    return s;
}

template<class T>
TransientData Data(const std::vector<T>& v)
{
    TransientData d(std::in_place_type_t<T>{}, v.size());
    d.Push(std::span<const T>(v));
    return d;
}
TransientData Str(std::string_view s)
{
    TransientData d(std::in_place_type_t<std::string>{}, s.size());
    d.PushString(s);
    return d;
}
CommonIdRange One(uint32_t id) { return {id, id}; }

struct Run { uint32_t first, count; bool light; uint32_t key; uint8_t cull; };

// Build the scene through TracerI.  Returns the camera surface.
CamSurfaceId Feed(TracerI& t, const SceneFile& s)
{
    // ---- runs -------------------------------------------------------------
    std::vector<Run> runs;
    for(uint32_t i = 0; i < s.T; i++)
    {
        bool light = (s.keys[i] & kLightFlag) != 0;
        uint32_t li = s.keys[i] & ~kLightFlag;
        if(!runs.empty())
        {
            Run& r = runs.back();
            bool same = false;
            if(r.light && light && r.cull == s.cull[i])
            {
                uint32_t l0 = r.key;
                same = (li == l0 + r.count) &&
                       std::equal(&s.lightRad[size_t(li) * 3], &s.lightRad[size_t(li) * 3 + 3],
                                  &s.lightRad[size_t(l0) * 3]) &&
                       s.twoSided[li] == s.twoSided[l0];
            }
            else if(!r.light && !light) same = (r.key == s.keys[i] && r.cull == s.cull[i]);
            if(same) { r.count++; continue; }
        }
        runs.push_back(Run{i, 1, light, light ? li : s.keys[i], s.cull[i]});
    }
    // ---- primitives -------------------------------------------------------
    PrimGroupId pg = t.CreatePrimitiveGroup("(P)Triangle");
    std::vector<PrimCount> counts;
    std::vector<std::vector<uint32_t>> runVerts;
    for(const Run& r : runs)
    {
        std::vector<uint32_t> vs(&s.idx[size_t(r.first) * 3], &s.idx[size_t(r.first + r.count) * 3]);
        std::sort(vs.begin(), vs.end());
        vs.erase(std::unique(vs.begin(), vs.end()), vs.end());
        counts.push_back(PrimCount{r.count, uint32_t(vs.size())});
        runVerts.push_back(std::move(vs));
    }
    PrimBatchIdList batches = t.ReservePrimitiveBatches(pg, counts);
    t.CommitPrimReservations(pg);
    for(size_t k = 0; k < runs.size(); k++)
    {
        const auto& vs = runVerts[k];
        std::vector<Vector3> p(vs.size());
        std::vector<Quaternion> q(vs.size());
        std::vector<Vector2> uv(vs.size());
        for(size_t j = 0; j < vs.size(); j++)
        {
            std::copy_n(&s.pos[size_t(vs[j]) * 3], 3, p[j].v.data());
            std::copy_n(&s.tbn[size_t(vs[j]) * 4], 4, q[j].v.v.data());
        }
        std::vector<Vector3ui> ix(runs[k].count);
        for(uint32_t i = 0; i < runs[k].count; i++)
            for(int c = 0; c < 3; c++)
            {
                uint32_t g = s.idx[size_t(runs[k].first + i) * 3 + c];
                ix[i].v[c] = uint32_t(std::lower_bound(vs.begin(), vs.end(), g) - vs.begin());
            }
        t.PushPrimAttribute(pg, batches[k], 0, Data(p));
        t.PushPrimAttribute(pg, batches[k], 1, Data(q));
        t.PushPrimAttribute(pg, batches[k], 2, Data(uv));
        t.PushPrimAttribute(pg, batches[k], 3, Data(ix));
    }
    // ---- materials (reserved in file order: global index == file index) --
    const char* matNames[4] = {"(Mt)Lambert", "(Mt)Reflect", "(Mt)Refract", "(Mt)Unreal"};
    std::map<uint32_t, MatGroupId> matGroups;
    std::vector<MaterialId> matIds;
    for(const Mat& m : s.mats)
    {
        if(!matGroups.count(m.type)) matGroups[m.type] = t.CreateMaterialGroup(matNames[m.type]);
        matIds.push_back(t.ReserveMaterial(matGroups[m.type], {}));
    }
    for(auto& kv : matGroups) t.CommitMatReservations(kv.second);
    auto texNone = std::vector<Optional<TextureId>>{std::nullopt};
    for(uint32_t i = 0; i < s.M; i++)
    {
        const Mat& m = s.mats[i];
        MatGroupId g = matGroups[m.type];
        CommonIdRange r = One(uint32_t(matIds[i]));
        auto v3 = [](const float* a) { Vector3 v; std::copy_n(a, 3, v.v.data()); return std::vector<Vector3>{v}; };
        auto f1 = [](float a) { return std::vector<float>{a}; };
        switch(m.type)
        {
            case 0: t.PushMatAttribute(g, r, 0, Data(v3(m.albedo)), texNone); break;
            case 1: break;
            case 2:
                t.PushMatAttribute(g, r, 0, Data(v3(m.cauchyBack)));
                t.PushMatAttribute(g, r, 1, Data(v3(m.cauchyFront)));
                break;
            default:
                t.PushMatAttribute(g, r, 0, Data(v3(m.albedo)), texNone);
                t.PushMatAttribute(g, r, 2, Data(f1(m.roughness)), texNone);
                t.PushMatAttribute(g, r, 3, Data(f1(m.specular)), texNone);
                t.PushMatAttribute(g, r, 4, Data(f1(m.metallic)), texNone);
                break;
        }
    }
    // ---- lights -------------------------------------------------------------
    std::vector<LightId> runLight(runs.size(), LightId(0));
    bool anyLight = std::any_of(runs.begin(), runs.end(), [](const Run& r) { return r.light; });
    if(anyLight)
    {
        LightGroupId lg = t.CreateLightGroup("(L)Prim(P)Triangle", pg);
        std::vector<size_t> lr;
        std::vector<AttributeCountList> lc;
        std::vector<PrimBatchId> lb;
        for(size_t k = 0; k < runs.size(); k++)
            if(runs[k].light) { lr.push_back(k); lc.push_back({}); lb.push_back(batches[k]); }
        LightIdList ids = t.ReserveLights(lg, lc, lb);
        t.CommitLightReservations(lg);
        for(size_t j = 0; j < lr.size(); j++)
        {
            const Run& r = runs[lr[j]];
            runLight[lr[j]] = ids[j];
            Vector3 rad;
            std::copy_n(&s.lightRad[size_t(r.key) * 3], 3, rad.v.data());
            t.PushLightAttribute(lg, One(uint32_t(ids[j])), 0, Data(std::vector<Vector3>{rad}), texNone);
            TransientData two(std::in_place_type_t<bool>{}, 1);
            bool b = s.twoSided[r.key] != 0;
            two.Push(std::span<const bool>(&b, 1));
            t.PushLightAttribute(lg, One(uint32_t(ids[j])), 1, std::move(two));
        }
    }
    if(s.boundaryType == 1 || s.boundaryType == 2)
    {
        LightGroupId sg = t.CreateLightGroup(s.boundaryType == 1 ? "(L)Skysphere_CoOcta"
                                                                 : "(L)Skysphere_Spherical");
        LightId sky = t.ReserveLight(sg, {});
        t.CommitLightReservations(sg);
        Vector3 rad;
        std::copy_n(s.boundaryRad, 3, rad.v.data());
        t.PushLightAttribute(sg, One(uint32_t(sky)), 0, Data(std::vector<Vector3>{rad}), texNone);
        t.SetBoundarySurface(sky, TransformId(0));
    }
    else
        t.SetBoundarySurface(LightId(0), TransformId(0));
    // ---- surfaces (triangle order) -----------------------------------------
    for(size_t k = 0; k < runs.size(); k++)
    {
        if(runs[k].light)
            t.CreateLightSurface(LightSurfaceParams{runLight[k], TransformId(0), {}});
        else
        {
            SurfaceParams sp;
            sp.primBatches = {batches[k]};
            sp.materials = {matIds[runs[k].key]};
            sp.transformId = TransformId(0);
            sp.cullFaceFlags = {runs[k].cull != 0};
            t.CreateSurface(std::move(sp));
        }
    }
    // ---- camera ---------------------------------------------------------------
    CameraGroupId cg = t.CreateCameraGroup("(C)Pinhole");
    CameraId cam = t.ReserveCamera(cg, {});
    t.CommitCamReservations(cg);
    Vector4 fp{{s.cam[9], s.cam[10], s.cam[11], s.cam[12]}};
    Vector3 pos{{s.cam[0], s.cam[1], s.cam[2]}}, gaze{{s.cam[3], s.cam[4], s.cam[5]}},
        up{{s.cam[6], s.cam[7], s.cam[8]}};
    t.PushCamAttribute(cg, One(uint32_t(cam)), 0, Data(std::vector<Vector4>{fp}));
    t.PushCamAttribute(cg, One(uint32_t(cam)), 1, Data(std::vector<Vector3>{gaze}));
    t.PushCamAttribute(cg, One(uint32_t(cam)), 2, Data(std::vector<Vector3>{pos}));
    t.PushCamAttribute(cg, One(uint32_t(cam)), 3, Data(std::vector<Vector3>{up}));
    return t.CreateCameraSurface(CameraSurfaceParams{cam, TransformId(0), {}});
}

// Error behaviour and type queries (no GPU): every misuse must throw MRayError.
int Errors(TracerI& t)
{
    int bad = 0;
    auto expectThrow = [&](const char* what, auto&& fn)
    {
        try { fn(); std::printf("FAIL (no throw): %s\n", what); bad++; }
        catch(const MRayError& e) { std::printf("ok: %s -> %s\n", what, e.what()); }
    };
    auto expect = [&](const char* what, bool c)
    {
        std::printf("%s: %s\n", c ? "ok" : "FAIL", what);
        bad += c ? 0 : 1;
    };
    auto has = [](const TypeNameList& l, std::string_view n) { return std::find(l.begin(), l.end(), n) != l.end(); };
    expect("prim types", has(t.PrimitiveGroups(), "(P)Triangle"));
    expect("material types", has(t.MaterialGroups(), "(Mt)Unreal") && t.MaterialGroups().size() == 4);
    expect("light types", has(t.LightGroups(), "(L)Prim(P)Triangle") && has(t.LightGroups(), "(L)Skysphere_CoOcta") &&
           has(t.LightGroups(), "(L)Skysphere_Spherical"));
    expect("renderer types", has(t.Renderers(), "(R)PathTracerRGB") && has(t.Renderers(), "(R)PathTracerSpectral"));
    auto ua = t.AttributeInfoMat("(Mt)Unreal");
    expect("Unreal attributes", ua.size() == 5 && ua[0].name == "albedo" && ua[2].name == "roughness" &&
                                ua[3].name == "specular" && ua[4].name == "metallic");
    expect("Pinhole attributes", t.AttributeInfoCam("(C)Pinhole").size() == 4);
    expect("renderer attributes", t.AttributeInfoRenderer("(R)PathTracerRGB").size() == 6);
    expectThrow("unknown primitive type", [&] { (void)t.CreatePrimitiveGroup("(P)Sphere"); });
    expectThrow("unknown material type", [&] { (void)t.CreateMaterialGroup("(Mt)Glass"); });
    expectThrow("unknown renderer type", [&] { (void)t.CreateRenderer("(R)Nope"); });
    RendererId r = t.CreateRenderer("(R)PathTracerRGB");
    expectThrow("unknown renderer attribute", [&] { t.PushRendererAttribute(r, 6, Data(std::vector<uint32_t>{1})); });
    expectThrow("unknown sample mode", [&] { t.PushRendererAttribute(r, 3, Str("Bogus")); });
    expectThrow("unknown light sampler", [&] { t.PushRendererAttribute(r, 5, Str("Irradiance")); });
    try
    {
        t.PushRendererAttribute(r, 5, Str("Uniform"));
        t.PushRendererAttribute(r, 5, Str("Power"));
        expect("light samplers Uniform and Power accepted", true);
    }
    catch(const MRayError&) { expect("light samplers Uniform and Power accepted", false); }
    expectThrow("wrong attribute type", [&] { t.PushRendererAttribute(r, 0, Data(std::vector<float>{1.0f})); });
    expectThrow("StartRender before CommitSurfaces", [&] {
        (void)t.StartRender(r, CamSurfaceId(0), RenderImageParams{}, std::nullopt, std::nullopt); });
    expectThrow("CommitSurfaces with no triangles", [&] { (void)t.CommitSurfaces(); });
    PrimGroupId pg = t.CreatePrimitiveGroup("(P)Triangle");
    PrimBatchId b = t.ReservePrimitiveBatch(pg, PrimCount{1, 3});
    expectThrow("push before commit", [&] {
        t.PushPrimAttribute(pg, b, 0, Data(std::vector<Vector3>(3))); });
    t.CommitPrimReservations(pg);
    expectThrow("wrong element count", [&] { t.PushPrimAttribute(pg, b, 0, Data(std::vector<Vector3>(2))); });
    expectThrow("index outside batch", [&] {
        t.PushPrimAttribute(pg, b, 3, Data(std::vector<Vector3ui>{Vector3ui{{0, 1, 3}}})); });
    expectThrow("unknown prim attribute", [&] { t.PushPrimAttribute(pg, b, 7, Data(std::vector<Vector3>(3))); });
    MatGroupId mg = t.CreateMaterialGroup("(Mt)Lambert");
    MaterialId m = t.ReserveMaterial(mg, {});
    t.CommitMatReservations(mg);
    expectThrow("constant push to a texturable attribute", [&] {
        t.PushMatAttribute(mg, One(uint32_t(m)), 0, Data(std::vector<Vector3>(1))); });
    expectThrow("texture-backed albedo", [&] {
        t.PushMatAttribute(mg, One(uint32_t(m)), 0, Data(std::vector<Vector3>(1)),
                           std::vector<Optional<TextureId>>{TextureId(1)}); });
    expectThrow("textures", [&] { (void)t.CreateTexture2D(Vector2ui{{4, 4}}, 1, MRayTextureParameters{}); });
    expectThrow("mediums", [&] { (void)t.CreateMediumGroup("(Md)Vacuum"); });
    expectThrow("surface with unknown material", [&] {
        SurfaceParams sp;
        sp.primBatches = {b};
        sp.materials = {MaterialId(0x05000000u)};
        (void)t.CreateSurface(std::move(sp)); });
    TransGroupId tg = t.CreateTransformGroup("(T)Single");
    TransformId tr = t.ReserveTransformation(tg, {});
    t.CommitTransReservations(tg);
    // a surface transform may be any non-singular affine matrix (positions
    // only); a singular one is an error
    Matrix4x4 shear;
    shear.m[1] = 0.5f;
    try
    {
        t.PushTransAttribute(tg, One(uint32_t(tr)), 0, Data(std::vector<Matrix4x4>{shear}));
        expect("sheared transform accepted", true);
    }
    catch(const MRayError&) { expect("sheared transform accepted", false); }
    Matrix4x4 flat;
    flat.m[10] = 0.0f;
    expectThrow("singular transform", [&] {
        t.PushTransAttribute(tg, One(uint32_t(tr)), 0, Data(std::vector<Matrix4x4>{flat})); });
    return bad == 0 ? 0 : 1;
}

int Run(int argc, char** argv)
{
    if(argc == 3 && std::string(argv[2]) == "errors")
    {
        void* so = dlopen(argv[1], RTLD_NOW | RTLD_LOCAL);
        if(!so) throw MRayError(std::string("dlopen: ") + dlerror());
        auto ctor = reinterpret_cast<TracerI* (*)(const TracerParameters&)>(dlsym(so, "ConstructTracer"));
        auto dtor = reinterpret_cast<void (*)(TracerI*)>(dlsym(so, "DestroyTracer"));
        if(!ctor || !dtor) throw MRayError("ConstructTracer / DestroyTracer not exported");
        TracerI* t = ctor(TracerParameters{});
        int rc = Errors(*t);
        dtor(t);
        dlclose(so);
        return rc;
    }
    if(argc < 5) throw MRayError("usage: tracer_driver <libmray_tracer.so> <scene.bin> flatten|render <out> ...");
    void* so = dlopen(argv[1], RTLD_NOW | RTLD_LOCAL);
    if(!so) throw MRayError(std::string("dlopen: ") + dlerror());
    using Ctor = TracerI* (*)(const TracerParameters&);
    using Dtor = void (*)(TracerI*);
    using Flat = int (*)(TracerI*, const char*);
    auto ctor = reinterpret_cast<Ctor>(dlsym(so, "ConstructTracer"));
    auto dtor = reinterpret_cast<Dtor>(dlsym(so, "DestroyTracer"));
    if(!ctor || !dtor) throw MRayError("ConstructTracer / DestroyTracer not exported");
    std::string mode = argv[3];
    TracerParameters tp;
    uint32_t W = 0, H = 0, spp = 0, spectral = 0, sampleMode = 2, rr0 = 8, rr1 = 8, slots = 0;
    if(mode == "render")
    {
        if(argc < 16)
            throw MRayError("render needs: W H spp spectral mode rr0 rr1 seed slots renderMode burstSize");
        W = uint32_t(std::atoi(argv[5])); H = uint32_t(std::atoi(argv[6]));
        spp = uint32_t(std::atoi(argv[7])); spectral = uint32_t(std::atoi(argv[8]));
        sampleMode = uint32_t(std::atoi(argv[9]));
        rr0 = uint32_t(std::atoi(argv[10])); rr1 = uint32_t(std::atoi(argv[11]));
        tp.seed = std::strtoull(argv[12], nullptr, 10);
        slots = uint32_t(std::atoi(argv[13]));
        if(slots) tp.parallelizationHint = slots;
        else tp.parallelizationHint = W * H;
    }
    const std::string scenePath = argv[2];
    const bool isJson = scenePath.size() > 5 && scenePath.substr(scenePath.size() - 5) == ".json";
    SceneFile s;
    if(!isJson) s = Load(argv[2]);
    TracerI* t = ctor(tp);
    int rc = 0;
    try
    {
        CamSurfaceId cs{0};
        if(isJson)
        {
            // MRay scene file through the library's loader
            SceneIds ids = LoadSceneMRay(*t, scenePath);
            if(ids.cameraSurfaces.empty()) throw MRayError("scene has no camera surface");
            cs = ids.cameraSurfaces.front();
            std::printf("loaded: %zu surfaces, %zu light surfaces, %zu camera surfaces\n",
                        ids.surfaces.size(), ids.lightSurfaces.size(), ids.cameraSurfaces.size());
        }
        else cs = Feed(*t, s);
        if(mode == "flatten")
        {
            auto flat = reinterpret_cast<Flat>(dlsym(so, "mray_tracer_flatten_to_file"));
            if(!flat || flat(t, argv[4]) != 0) throw MRayError("flatten failed");
        }
        else if(mode == "render")
        {
            SurfaceCommitResult cr = t->CommitSurfaces();
            std::printf("committed: %zu surfaces, aabb [%g %g %g] - [%g %g %g]\n", cr.instanceCount,
                        cr.aabb.min.v[0], cr.aabb.min.v[1], cr.aabb.min.v[2],
                        cr.aabb.max.v[0], cr.aabb.max.v[1], cr.aabb.max.v[2]);
            RendererId r = t->CreateRenderer(spectral ? "(R)PathTracerSpectral" : "(R)PathTracerRGB");
            const char* modes[3] = {"Pure", "WithNEE", "WithNEEAndMIS"};
            t->PushRendererAttribute(r, 0, Data(std::vector<uint32_t>{spp}));
            t->PushRendererAttribute(r, 1, Data(std::vector<uint32_t>{uint32_t(std::atoi(argv[15]))}));
            t->PushRendererAttribute(r, 2, Str(argv[14]));
            t->PushRendererAttribute(r, 3, Str(modes[sampleMode]));
            t->PushRendererAttribute(r, 4, Data(std::vector<Vector2ui>{Vector2ui{{rr0, rr1}}}));
            t->PushRendererAttribute(r, 5, Str("Uniform"));
            TimelineSemaphore sem(0);
            t->SetupRenderEnv(&sem, 4096, 0);
            RenderImageParams ip{Vector2ui{{W, H}}, Vector2ui{{0, 0}}, Vector2ui{{W, H}}};
            RenderBufferInfo info = t->StartRender(r, cs, ip, std::nullopt, std::nullopt);
            if(reinterpret_cast<uintptr_t>(info.data) % 4096 != 0)
                throw MRayError("staging buffer is not aligned to the import alignment");
            // consumer: the run command's section loop
            const size_t P = size_t(W) * H;
            std::vector<double> acc(4 * P, 0.0);
            std::deque<RenderImageSection> queue;
            std::mutex qm;
            std::condition_variable qc;
            bool done = false;
            uint32_t consumed = 0;
            std::thread consumer([&]
            {
                for(;;)
                {
                    RenderImageSection sec;
                    {
                        std::unique_lock<std::mutex> lk(qm);
                        qc.wait(lk, [&] { return done || !queue.empty(); });
                        if(queue.empty()) return;
                        sec = queue.front();
                        queue.pop_front();
                    }
                    if(!sem.Acquire(sec.waitCounter)) return;
                    const float* rgb[3];
                    for(int c = 0; c < 3; c++)
                        rgb[c] = reinterpret_cast<const float*>(info.data + sec.pixStartOffsets[c]);
                    const float* w = reinterpret_cast<const float*>(info.data + sec.weightStartOffset);
                    // the section's pixels [pixelMin, pixelMax), scanlines
                    // packed at its own width (RunCommand.cpp:609-700)
                    const uint32_t sw = sec.pixelMax[0] - sec.pixelMin[0];
                    const uint32_t sh = sec.pixelMax[1] - sec.pixelMin[1];
                    for(uint32_t y = 0; y < sh; y++)
                        for(uint32_t x = 0; x < sw; x++)
                        {
                            const size_t k = size_t(y) * sw + x;
                            const size_t i = size_t(sec.pixelMin[1] + y) * W + sec.pixelMin[0] + x;
                            double sOut = acc[3 * P + i], total = sOut + double(w[k]);
                            for(int c = 0; c < 3; c++)
                                acc[c * P + i] = (total == 0.0) ? 0.0
                                    : (acc[c * P + i] * sOut + double(rgb[c][k])) / total;
                            acc[3 * P + i] = total;
                        }
                    consumed++;
                    sem.Release();
                }
            });
            RendererOutput out;
            uint32_t calls = 0, saves = 0;
            try
            {
                do
                {
                    out = t->DoRenderWork();
                    calls++;
                    if(out.triggerSave) saves++;
                    if(out.imageOut)
                    {
                        std::lock_guard<std::mutex> lk(qm);
                        queue.push_back(*out.imageOut);
                        qc.notify_all();
                    }
                    if(calls > 100000) throw MRayError("render did not finish");
                } while(!out.triggerSave);
                // one more call after the save request must not ask again;
                // its samples go past totalSPP (a tiled renderer keeps
                // rendering, as the reference's does for an open viewer), so
                // the saved image does not take them
                RendererOutput extra = t->DoRenderWork();
                if(extra.triggerSave) throw MRayError("triggerSave reported twice");
            }
            catch(...)
            {
                sem.Invalidate();
                { std::lock_guard<std::mutex> lk(qm); done = true; qc.notify_all(); }
                consumer.join();
                throw;
            }
            { std::lock_guard<std::mutex> lk(qm); done = true; qc.notify_all(); }
            consumer.join();
            std::printf("rendered: %u DoRenderWork calls, %u sections consumed, %u save request, "
                        "%.3f spp, last %.3f %s\n", calls + 1, consumed, saves,
                        out.analytics->workPerPixel, out.analytics->throughput,
                        out.analytics->throughputSuffix.c_str());
            std::vector<float> img(4 * P);
            for(size_t i = 0; i < 4 * P; i++) img[i] = float(acc[i]);
            FILE* f = std::fopen(argv[4], "wb");
            if(!f) throw MRayError("cannot write film");
            std::fwrite(img.data(), sizeof(float), img.size(), f);
            std::fclose(f);
            if(argc > 16)
            {
                // RunCommand.cpp:873-930: sums back from mean x weight, then
                // <prefix>_<wpp>wpp.exr
                for(size_t i = 0; i < P; i++)
                    for(int c = 0; c < 3; c++) img[c * P + i] = float(acc[c * P + i] * acc[3 * P + i]);
                using Exr = int (*)(const float*, uint32_t, uint32_t, const char*);
                auto exr = reinterpret_cast<Exr>(dlsym(so, "mray_film_write_exr"));
                if(!exr) throw MRayError("mray_film_write_exr not found");
                char name[64];
                std::snprintf(name, sizeof(name), "_%08.0fwpp.exr", std::round(out.analytics->workPerPixel));
                std::string path = std::string(argv[16]) + name;
                if(exr(img.data(), W, H, path.c_str()) != 0) throw MRayError("EXR write failed");
                std::printf("saved %s\n", path.c_str());
            }
            t->StopRender();
            t->DestroyRenderer(r);
        }
        else throw MRayError("unknown mode " + mode);
    }
    catch(const MRayError& e)
    {
        std::fprintf(stderr, "MRayError: %s\n", e.what());
        rc = 3;
    }
    dtor(t);
    dlclose(so);
    return rc;
}

}  // namespace

int main(int argc, char** argv)
{
    try
    {
        return Run(argc, argv);
    }
    catch(const MRayError& e)
    {
        std::fprintf(stderr, "MRayError: %s\n", e.what());
        return 2;
    }
}
