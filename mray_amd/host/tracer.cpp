// TracerI implementation over the C-ABI (include/mray_amd.h).
//
// Follows the reference's TracerBase (Tracer/TracerBase.cpp) for the calls a
// path-tracing front-end makes: groups and reservations, attribute pushes,
// surfaces, CommitSurfaces, renderers, StartRender / DoRenderWork.  The scene
// is kept on the host until CommitSurfaces, which flattens every surface,
// light surface and the boundary light into one MRaySceneDesc (triangle soup
// with per-triangle material / light keys) and builds the accelerator on the
// GPU.  Transforms are baked into positions there; TBN quaternions turn with
// TransformPrimitives only, as in the reference, where a surface's transform
// moves positions but leaves the shading frame in mesh space.
#include "mray_tracer.hpp"
#include "mray_amd.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>

namespace mray_amd
{
namespace
{

constexpr uint32_t kIndexBits = 24;
constexpr uint32_t kIndexMask = (1u << kIndexBits) - 1u;

inline uint32_t Pack(uint32_t group, uint32_t index)
{
    if(group > 0xFFu || index > kIndexMask)
        throw MRayError("Id space exhausted (256 groups x 16M items)");
    return (group << kIndexBits) | index;
}
inline uint32_t GroupOf(uint32_t id) { return id >> kIndexBits; }
inline uint32_t IndexOf(uint32_t id) { return id & kIndexMask; }
template<class E> inline uint32_t U(E e) { return static_cast<uint32_t>(e); }

// ---- type names (Core/TypeNameGenerators.h, TracerConstants.h:42-56) ------
constexpr std::string_view kPrimTriangle = "(P)Triangle";
constexpr std::string_view kMatNames[4] = {"(Mt)Lambert", "(Mt)Reflect", "(Mt)Refract", "(Mt)Unreal"};
constexpr std::string_view kLightPrimTri = "(L)Prim(P)Triangle";
constexpr std::string_view kLightSky     = "(L)Skysphere_CoOcta";
constexpr std::string_view kLightSkySph  = "(L)Skysphere_Spherical";
constexpr std::string_view kLightNull    = "(L)Null";
constexpr std::string_view kCamPinhole   = "(C)Pinhole";
constexpr std::string_view kTransIdentity = "(T)Identity";
constexpr std::string_view kTransSingle   = "(T)Single";
constexpr std::string_view kRendRGB      = "(R)PathTracerRGB";
constexpr std::string_view kRendSpectral = "(R)PathTracerSpectral";

// ---- attribute lists (the reference's AttributeInfo() of each type) ------
using enum AttributeOptionality;
using enum AttributeTexturable;
using enum AttributeIsArray;
using enum AttributeIsColor;

// PrimitiveDefaultTriangle.cu:169-184
const PrimAttributeInfoList kTriAttribs = {
    {PrimitiveAttributeLogic::POSITION, MRayDataEnum::MR_VECTOR_3, IS_SCALAR, MR_MANDATORY},
    {PrimitiveAttributeLogic::NORMAL, MRayDataEnum::MR_QUATERNION, IS_SCALAR, MR_MANDATORY},
    {PrimitiveAttributeLogic::UV0, MRayDataEnum::MR_VECTOR_2, IS_SCALAR, MR_MANDATORY},
    {PrimitiveAttributeLogic::INDEX, MRayDataEnum::MR_VECTOR_3UI, IS_SCALAR, MR_MANDATORY}};

// MaterialsDefault.cpp:38-52, 161-165, 236-250, 382-402
MatAttributeInfoList MatAttribs(uint32_t type)
{
    switch(type)
    {
        case MRAY_MAT_LAMBERT: return {
            {"albedo", MRayDataEnum::MR_VECTOR_3, IS_SCALAR, MR_MANDATORY, MR_TEXTURE_OR_CONSTANT, IS_COLOR},
            {"normalMap", MRayDataEnum::MR_VECTOR_3, IS_SCALAR, MR_OPTIONAL, MR_TEXTURE_ONLY, IS_PURE_DATA}};
        case MRAY_MAT_REFLECT: return {};
        case MRAY_MAT_REFRACT: return {
            {"cauchyBack", MRayDataEnum::MR_VECTOR_3, IS_SCALAR, MR_MANDATORY, MR_CONSTANT_ONLY, IS_PURE_DATA},
            {"cauchyFront", MRayDataEnum::MR_VECTOR_3, IS_SCALAR, MR_MANDATORY, MR_CONSTANT_ONLY, IS_PURE_DATA}};
        default: return {
            {"albedo", MRayDataEnum::MR_VECTOR_3, IS_SCALAR, MR_MANDATORY, MR_TEXTURE_OR_CONSTANT, IS_COLOR},
            {"normalMap", MRayDataEnum::MR_VECTOR_3, IS_SCALAR, MR_OPTIONAL, MR_TEXTURE_ONLY, IS_PURE_DATA},
            {"roughness", MRayDataEnum::MR_FLOAT, IS_SCALAR, MR_MANDATORY, MR_TEXTURE_OR_CONSTANT, IS_PURE_DATA},
            {"specular", MRayDataEnum::MR_FLOAT, IS_SCALAR, MR_MANDATORY, MR_TEXTURE_OR_CONSTANT, IS_PURE_DATA},
            {"metallic", MRayDataEnum::MR_FLOAT, IS_SCALAR, MR_MANDATORY, MR_TEXTURE_OR_CONSTANT, IS_PURE_DATA}};
    }
}
// LightsDefault.hpp:547-560, 725-738; LightsDefault.cu (Null: no attributes)
const LightAttributeInfoList kPrimLightAttribs = {
    {"radiance", MRayDataEnum::MR_VECTOR_3, IS_SCALAR, MR_MANDATORY, MR_TEXTURE_OR_CONSTANT, IS_COLOR},
    {"isTwoSided", MRayDataEnum::MR_BOOL, IS_SCALAR, MR_MANDATORY, MR_CONSTANT_ONLY, IS_PURE_DATA}};
const LightAttributeInfoList kSkyAttribs = {
    {"radiance", MRayDataEnum::MR_VECTOR_3, IS_SCALAR, MR_MANDATORY, MR_TEXTURE_OR_CONSTANT, IS_COLOR}};
// CamerasDefault.cpp:40-49
const CamAttributeInfoList kPinholeAttribs = {
    {"FovAndPlanes", MRayDataEnum::MR_VECTOR_4, IS_SCALAR, MR_MANDATORY},
    {"gaze", MRayDataEnum::MR_VECTOR_3, IS_SCALAR, MR_MANDATORY},
    {"position", MRayDataEnum::MR_VECTOR_3, IS_SCALAR, MR_MANDATORY},
    {"up", MRayDataEnum::MR_VECTOR_3, IS_SCALAR, MR_MANDATORY}};
// TransformsDefault.cu:30-36
const TransAttributeInfoList kSingleAttribs = {
    {"Transform", MRayDataEnum::MR_MATRIX_4x4, IS_SCALAR, MR_MANDATORY}};
// PathTracerRenderer.cu:1386-1400
const RendererAttributeInfoList kRendererAttribs = {
    {"totalSPP", MRayDataEnum::MR_UINT32, IS_SCALAR, MR_MANDATORY},
    {"burstSize", MRayDataEnum::MR_UINT32, IS_SCALAR, MR_OPTIONAL},
    {"renderMode", MRayDataEnum::MR_STRING, IS_SCALAR, MR_MANDATORY},
    {"sampleMode", MRayDataEnum::MR_STRING, IS_SCALAR, MR_MANDATORY},
    {"rrRange", MRayDataEnum::MR_VECTOR_2UI, IS_SCALAR, MR_MANDATORY},
    {"neeSamplerType", MRayDataEnum::MR_STRING, IS_SCALAR, MR_MANDATORY}};

[[noreturn]] void Unsupported(std::string_view what)
{
    throw MRayError(std::string(what) + " is not supported by the MI355X path tracer");
}

void CheckRc(int rc, const char* what)
{
    if(rc != MRAY_OK)
        throw MRayError(std::string(what) + ": " + mray_last_error());
}

// ---- host-side math for baking transforms ---------------------------------
struct Rigid
{
    float m[12];     // row-major 3x4
    float q[4];      // rotation part as a quaternion (w, x, y, z)
    bool identity;
};

Rigid IdentityRigid()
{
    Rigid r{{1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0}, {1, 0, 0, 0}, true};
    return r;
}

// Matrix -> rigid transform with uniform scale.  The rotation is taken from
// the normalised columns (Shepperd's method).  With `affine` any
// non-singular matrix is kept as it is (a surface's transform moves only
// positions; its rotation part is not needed).
Rigid MakeRigid(const float* m12, bool affine = false)
{
    Rigid r;
    std::copy_n(m12, 12, r.m);
    r.identity = false;
    if(affine)
    {
        const double det = double(m12[0]) * (double(m12[5]) * m12[10] - double(m12[6]) * m12[9]) -
                           double(m12[1]) * (double(m12[4]) * m12[10] - double(m12[6]) * m12[8]) +
                           double(m12[2]) * (double(m12[4]) * m12[9] - double(m12[5]) * m12[8]);
        if(det == 0.0 || !std::isfinite(det)) throw MRayError("Transform: singular matrix");
        r.q[0] = 1; r.q[1] = r.q[2] = r.q[3] = 0;
        return r;
    }
    double c[3][3];
    double len[3];
    for(int j = 0; j < 3; j++)
    {
        len[j] = std::sqrt(double(m12[j]) * m12[j] + double(m12[4 + j]) * m12[4 + j] +
                           double(m12[8 + j]) * m12[8 + j]);
        if(len[j] <= 0.0) throw MRayError("Transform: singular matrix");
        for(int i = 0; i < 3; i++) c[i][j] = m12[i * 4 + j] / len[j];
    }
    double maxL = std::max({len[0], len[1], len[2]}), minL = std::min({len[0], len[1], len[2]});
    if((maxL - minL) > 1e-4 * maxL) Unsupported("A transform with non-uniform scale");
    for(int a = 0; a < 3; a++)
        for(int b = a + 1; b < 3; b++)
        {
            double d = c[0][a] * c[0][b] + c[1][a] * c[1][b] + c[2][a] * c[2][b];
            if(std::abs(d) > 1e-4) Unsupported("A sheared transform");
        }
    double det = c[0][0] * (c[1][1] * c[2][2] - c[1][2] * c[2][1]) -
                 c[0][1] * (c[1][0] * c[2][2] - c[1][2] * c[2][0]) +
                 c[0][2] * (c[1][0] * c[2][1] - c[1][1] * c[2][0]);
    if(det < 0) Unsupported("A mirroring transform");
    double tr = c[0][0] + c[1][1] + c[2][2];
    double w, x, y, z;
    if(tr > 0)
    {
        double s = std::sqrt(tr + 1.0) * 2.0;
        w = 0.25 * s; x = (c[2][1] - c[1][2]) / s; y = (c[0][2] - c[2][0]) / s; z = (c[1][0] - c[0][1]) / s;
    }
    else if(c[0][0] > c[1][1] && c[0][0] > c[2][2])
    {
        double s = std::sqrt(1.0 + c[0][0] - c[1][1] - c[2][2]) * 2.0;
        w = (c[2][1] - c[1][2]) / s; x = 0.25 * s; y = (c[0][1] + c[1][0]) / s; z = (c[0][2] + c[2][0]) / s;
    }
    else if(c[1][1] > c[2][2])
    {
        double s = std::sqrt(1.0 + c[1][1] - c[0][0] - c[2][2]) * 2.0;
        w = (c[0][2] - c[2][0]) / s; x = (c[0][1] + c[1][0]) / s; y = 0.25 * s; z = (c[1][2] + c[2][1]) / s;
    }
    else
    {
        double s = std::sqrt(1.0 + c[2][2] - c[0][0] - c[1][1]) * 2.0;
        w = (c[1][0] - c[0][1]) / s; x = (c[0][2] + c[2][0]) / s; y = (c[1][2] + c[2][1]) / s; z = 0.25 * s;
    }
    r.q[0] = float(w); r.q[1] = float(x); r.q[2] = float(y); r.q[3] = float(z);
    return r;
}

void ApplyPoint(const Rigid& r, const float* p, float* out)
{
    if(r.identity) { std::copy_n(p, 3, out); return; }
    for(int i = 0; i < 3; i++)
        out[i] = r.m[i * 4 + 0] * p[0] + r.m[i * 4 + 1] * p[1] + r.m[i * 4 + 2] * p[2] + r.m[i * 4 + 3];
}

// World-to-tangent quaternion q of a vertex; after world' = R world the new
// one is q * conj(r) (Hamilton product).
void ApplyTbn(const Rigid& r, const float* q, float* out)
{
    if(r.identity) { std::copy_n(q, 4, out); return; }
    float bw = r.q[0], bx = -r.q[1], by = -r.q[2], bz = -r.q[3];
    float aw = q[0], ax = q[1], ay = q[2], az = q[3];
    out[0] = aw * bw - ax * bx - ay * by - az * bz;
    out[1] = aw * bx + ax * bw + ay * bz - az * by;
    out[2] = aw * by - ax * bz + ay * bw + az * bx;
    out[3] = aw * bz + ax * by - ay * bx + az * bw;
}

Rigid Compose(const Rigid& outer, const Rigid& inner, bool affine = false)
{
    if(inner.identity) return outer;
    if(outer.identity) return inner;
    float m[12];
    for(int i = 0; i < 3; i++)
    {
        for(int j = 0; j < 4; j++)
        {
            float v = outer.m[i * 4 + 0] * inner.m[0 * 4 + j] + outer.m[i * 4 + 1] * inner.m[1 * 4 + j] +
                      outer.m[i * 4 + 2] * inner.m[2 * 4 + j];
            if(j == 3) v += outer.m[i * 4 + 3];
            m[i * 4 + j] = v;
        }
    }
    return MakeRigid(m, affine);
}

// ---- host scene storage ---------------------------------------------------
struct PrimBatch
{
    uint32_t primCount = 0, vertexCount = 0;
    std::vector<Vector3> pos;
    std::vector<Quaternion> tbn;
    std::vector<Vector2> uv;
    std::vector<Vector3ui> idx;
    Rigid baked = IdentityRigid();      // TransformPrimitives
};
struct PrimGroup
{
    std::vector<PrimBatch> batches;
    bool committed = false;
};
struct MatGroup
{
    uint32_t type = 0;
    std::vector<uint32_t> globalIndex;   // into Tracer::materials
    bool committed = false;
};
struct LightEntry
{
    uint32_t primBatch = 0;              // PrimBatchId (prim-backed lights)
    float radiance[3] = {0, 0, 0};
    bool twoSided = false;
};
struct LightGroup
{
    uint32_t kind = 0;                   // 0 null, 1 prim triangle, 2 skysphere
    uint32_t primGroup = 0;
    std::vector<LightEntry> lights;
    bool committed = false;
};
struct CameraEntry { float fovPlanes[4] = {0, 0, 0, 0}; Vector3 gaze, position, up; };
struct CamGroup { std::vector<CameraEntry> cams; bool committed = false; };
struct TransGroup { bool identity = true; std::vector<Rigid> xf; bool committed = false; };

struct SurfaceRec
{
    bool isLight = false;
    SurfaceParams params;               // isLight == false
    LightSurfaceParams light;           // isLight == true
};

struct RendererRec
{
    bool spectral = false;
    uint32_t totalSPP = 1;
    uint32_t burstSize = 1;
    uint32_t sampleMode = MRAY_SAMPLE_NEE_MIS;
    uint32_t rr[2] = {4, 20};
    bool powerLights = false;   // neeSamplerType "Power"
    bool latency = false;       // renderMode "Latency"
};

uint32_t ParseSampleMode(std::string_view s)
{
    // PathTracerRendererShaders.h:25-35
    if(s == "Pure") return MRAY_SAMPLE_PURE;
    if(s == "WithNEE") return MRAY_SAMPLE_NEE;
    if(s == "WithNEEAndMIS") return MRAY_SAMPLE_NEE_MIS;
    throw MRayError("Unknown sample mode \"" + std::string(s) + "\"");
}

class Tracer final : public TracerI
{
    TracerParameters params;
    std::vector<PrimGroup> primGroups;      // index 0 = EmptyPrimGroupId
    std::vector<MatGroup> matGroups;        // index 0 = passthrough
    std::vector<MRayMaterial> materials;    // flattened, reservation order
    std::vector<TransGroup> transGroups;    // index 0 = identity
    std::vector<LightGroup> lightGroups;    // index 0 = Null
    std::vector<CamGroup> camGroups;        // index 0 = unused
    std::vector<SurfaceRec> surfaces;       // surfaces + light surfaces, creation order
    std::vector<CameraSurfaceParams> camSurfaces;
    bool hasBoundary = false;
    uint32_t boundaryLight = 0;
    uint32_t boundaryTransform = 0;
    std::map<uint32_t, RendererRec> renderers;
    uint32_t nextRenderer = 0;

    // committed products
    struct Flat
    {
        std::vector<float> pos, tbn, lightRad;
        std::vector<uint32_t> idx, keys, lightTris;
        std::vector<uint8_t> cull, lightTwoSided;
        uint32_t boundaryType = MRAY_BOUNDARY_NULL;
        float boundaryRad[3] = {0, 0, 0};
        AABB3 aabb;
        float sceneDiameter = 0;
    } flat;
    MRayAccel* accel = nullptr;
    bool committed = false;

    // render state
    MRayRenderer* renderer = nullptr;
    uint32_t activeRenderer = 0;
    RenderImageParams imgParams{};
    // image hand-over (Tracer/RenderImage.cpp:150-209): a host staging buffer
    // aligned for import by the consumer, filled with the sums splatted
    // since the previous section, guarded by the consumer's semaphore
    struct Staging
    {
        float* p = nullptr;
        size_t bytes = 0;
        ~Staging() { std::free(p); }
        void Allocate(size_t need, size_t align)
        {
            std::free(p);
            bytes = (need + align - 1) / align * align;
            p = static_cast<float*>(std::aligned_alloc(align, bytes));
            if(!p) throw MRayError("Unable to allocate the image staging buffer");
            std::memset(p, 0, bytes);
        }
    } staging;
    TimelineSemaphore* sem = nullptr;
    uint32_t importAlignment = 64;
    uint64_t acquireValue = 0;
    bool saveImage = false;
    uint64_t sampleCounter = 0;
    double lastIterMs = 0;
    // tile by tile (DoLatencyRender): the renderer's tiles and, per tile,
    // the samples it has (PathTracerRendererBase::tileSPPs)
    bool tiled = false;
    uint32_t tileCount = 1;
    uint32_t curTile = 0;
    std::vector<uint32_t> tileSPPs;

    // ---- lookups -----------------------------------------------------------
    template<class V> static auto& At(V& v, uint32_t g, const char* what)
    {
        if(g >= v.size()) throw MRayError(std::string("Unable to find ") + what + "(" + std::to_string(g) + ")");
        return v[g];
    }
    PrimBatch& Batch(uint32_t batchId)
    {
        auto& g = At(primGroups, GroupOf(batchId), "PrimitiveGroup");
        if(IndexOf(batchId) >= g.batches.size())
            throw MRayError("Unknown PrimBatchId " + std::to_string(batchId));
        return g.batches[IndexOf(batchId)];
    }
    Rigid TransformOf(TransformId id) const
    {
        uint32_t g = GroupOf(U(id)), i = IndexOf(U(id));
        if(g >= transGroups.size()) throw MRayError("Unknown TransformId " + std::to_string(U(id)));
        const TransGroup& tg = transGroups[g];
        if(tg.identity) return IdentityRigid();
        if(i >= tg.xf.size()) throw MRayError("Unknown TransformId " + std::to_string(U(id)));
        return tg.xf[i];
    }
    static void CheckRange(CommonIdRange r, uint32_t group, size_t count)
    {
        if(GroupOf(r[0]) != group || GroupOf(r[1]) != group || IndexOf(r[0]) > IndexOf(r[1]) ||
           IndexOf(r[1]) >= count)
            throw MRayError("Id range does not belong to the group");
    }
    template<class T>
    static std::span<const T> Elems(const TransientData& d, size_t n, const char* what)
    {
        auto s = d.AccessAs<T>();
        if(s.size() != n)
            throw MRayError(std::string(what) + ": expected " + std::to_string(n) + " elements, got " +
                            std::to_string(s.size()));
        return s;
    }

    public:
    explicit Tracer(const TracerParameters& p) : params(p) { ClearAll(); }
    ~Tracer() override { Release(); }

    void Release()
    {
        if(renderer) { mray_renderer_destroy(renderer); renderer = nullptr; }
        if(accel) { mray_accel_destroy(accel); accel = nullptr; }
    }

    // ================================ Generic ================================
    TypeNameList PrimitiveGroups() const override { return {kPrimTriangle}; }
    TypeNameList MaterialGroups() const override { return {kMatNames[0], kMatNames[1], kMatNames[2], kMatNames[3]}; }
    TypeNameList TransformGroups() const override { return {kTransIdentity, kTransSingle}; }
    TypeNameList CameraGroups() const override { return {kCamPinhole}; }
    TypeNameList MediumGroups() const override { return {}; }
    TypeNameList LightGroups() const override { return {kLightNull, kLightPrimTri, kLightSky, kLightSkySph}; }
    TypeNameList Renderers() const override { return {kRendRGB, kRendSpectral}; }

    PrimAttributeInfoList AttributeInfo(PrimGroupId g) const override
    {
        if(U(g) == 0 || U(g) >= primGroups.size()) throw MRayError("Unknown PrimGroupId");
        return kTriAttribs;
    }
    CamAttributeInfoList AttributeInfo(CameraGroupId g) const override
    {
        if(U(g) == 0 || U(g) >= camGroups.size()) throw MRayError("Unknown CameraGroupId");
        return kPinholeAttribs;
    }
    MediumAttributeInfoList AttributeInfo(MediumGroupId) const override { Unsupported("Mediums"); }
    MatAttributeInfoList AttributeInfo(MatGroupId g) const override
    {
        if(U(g) == 0 || U(g) >= matGroups.size()) throw MRayError("Unknown MatGroupId");
        return MatAttribs(matGroups[U(g)].type);
    }
    TransAttributeInfoList AttributeInfo(TransGroupId g) const override
    {
        if(U(g) >= transGroups.size()) throw MRayError("Unknown TransGroupId");
        return transGroups[U(g)].identity ? TransAttributeInfoList{} : kSingleAttribs;
    }
    LightAttributeInfoList AttributeInfo(LightGroupId g) const override
    {
        if(U(g) >= lightGroups.size()) throw MRayError("Unknown LightGroupId");
        switch(lightGroups[U(g)].kind)
        {
            case 1: return kPrimLightAttribs;
            case 2:
            case 3: return kSkyAttribs;
            default: return {};
        }
    }
    RendererAttributeInfoList AttributeInfo(RendererId r) const override
    {
        if(!renderers.count(U(r))) throw MRayError("Unknown RendererId");
        return kRendererAttribs;
    }
    PrimAttributeInfoList AttributeInfoPrim(std::string_view n) const override
    {
        if(n != kPrimTriangle) throw MRayError("Unknown primitive type \"" + std::string(n) + "\"");
        return kTriAttribs;
    }
    CamAttributeInfoList AttributeInfoCam(std::string_view n) const override
    {
        if(n != kCamPinhole) throw MRayError("Unknown camera type \"" + std::string(n) + "\"");
        return kPinholeAttribs;
    }
    MediumAttributeInfoList AttributeInfoMedium(std::string_view) const override { Unsupported("Mediums"); }
    MatAttributeInfoList AttributeInfoMat(std::string_view n) const override { return MatAttribs(MatType(n)); }
    TransAttributeInfoList AttributeInfoTrans(std::string_view n) const override
    {
        if(n == kTransIdentity) return {};
        if(n == kTransSingle) return kSingleAttribs;
        throw MRayError("Unknown transform type \"" + std::string(n) + "\"");
    }
    LightAttributeInfoList AttributeInfoLight(std::string_view n) const override
    {
        switch(LightKind(n))
        {
            case 1: return kPrimLightAttribs;
            case 2:
            case 3: return kSkyAttribs;
            default: return {};
        }
    }
    RendererAttributeInfoList AttributeInfoRenderer(std::string_view n) const override
    {
        RendererKind(n);
        return kRendererAttribs;
    }

    std::string TypeName(PrimGroupId) const override { return std::string(kPrimTriangle); }
    std::string TypeName(CameraGroupId) const override { return std::string(kCamPinhole); }
    std::string TypeName(MediumGroupId) const override { Unsupported("Mediums"); }
    std::string TypeName(MatGroupId g) const override
    {
        if(U(g) == 0 || U(g) >= matGroups.size()) throw MRayError("Unknown MatGroupId");
        return std::string(kMatNames[matGroups[U(g)].type]);
    }
    std::string TypeName(TransGroupId g) const override
    {
        if(U(g) >= transGroups.size()) throw MRayError("Unknown TransGroupId");
        return std::string(transGroups[U(g)].identity ? kTransIdentity : kTransSingle);
    }
    std::string TypeName(LightGroupId g) const override
    {
        if(U(g) >= lightGroups.size()) throw MRayError("Unknown LightGroupId");
        uint32_t k = lightGroups[U(g)].kind;
        return std::string(k == 1 ? kLightPrimTri : k == 2 ? kLightSky : k == 3 ? kLightSkySph : kLightNull);
    }
    std::string TypeName(RendererId r) const override
    {
        auto it = renderers.find(U(r));
        if(it == renderers.end()) throw MRayError("Unknown RendererId");
        return std::string(it->second.spectral ? kRendSpectral : kRendRGB);
    }

    static uint32_t MatType(std::string_view n)
    {
        for(uint32_t i = 0; i < 4; i++)
            if(n == kMatNames[i]) return i;
        throw MRayError("Unknown material type \"" + std::string(n) + "\"");
    }
    static uint32_t LightKind(std::string_view n)
    {
        if(n == kLightNull) return 0;
        if(n == kLightPrimTri) return 1;
        if(n == kLightSky) return 2;
        if(n == kLightSkySph) return 3;
        throw MRayError("Unknown light type \"" + std::string(n) + "\"");
    }
    static bool RendererKind(std::string_view n)
    {
        if(n == kRendRGB) return false;
        if(n == kRendSpectral) return true;
        throw MRayError("Unknown renderer type \"" + std::string(n) + "\"");
    }

    // =============================== Primitive ===============================
    PrimGroupId CreatePrimitiveGroup(std::string typeName) override
    {
        if(typeName != kPrimTriangle) throw MRayError("Unknown primitive type \"" + typeName + "\"");
        primGroups.emplace_back();
        return PrimGroupId(uint32_t(primGroups.size() - 1));
    }
    PrimBatchId ReservePrimitiveBatch(PrimGroupId g, PrimCount c) override
    {
        return ReservePrimitiveBatches(g, {c}).front();
    }
    PrimBatchIdList ReservePrimitiveBatches(PrimGroupId g, std::vector<PrimCount> counts) override
    {
        auto& grp = At(primGroups, U(g), "PrimitiveGroup");
        if(U(g) == 0) throw MRayError("The empty primitive group takes no batches");
        if(grp.committed) throw MRayError("PrimitiveGroup is already committed");
        PrimBatchIdList out;
        for(const PrimCount& c : counts)
        {
            PrimBatch b;
            b.primCount = c.primCount;
            b.vertexCount = c.attributeCount;
            out.push_back(PrimBatchId(Pack(U(g), uint32_t(grp.batches.size()))));
            grp.batches.push_back(std::move(b));
        }
        return out;
    }
    void CommitPrimReservations(PrimGroupId g) override
    {
        auto& grp = At(primGroups, U(g), "PrimitiveGroup");
        for(PrimBatch& b : grp.batches)
        {
            b.pos.resize(b.vertexCount);
            b.tbn.resize(b.vertexCount);
            b.uv.resize(b.vertexCount);
            b.idx.resize(b.primCount);
        }
        grp.committed = true;
    }
    bool IsPrimCommitted(PrimGroupId g) const override
    {
        return U(g) < primGroups.size() && primGroups[U(g)].committed;
    }
    void PushPrimAttribute(PrimGroupId g, PrimBatchId b, uint32_t attr, TransientData data) override
    {
        PrimBatch& pb = CommittedBatch(g, b);
        uint32_t n = (attr == 3) ? pb.primCount : pb.vertexCount;
        PushPrimRange(pb, attr, 0, n, data);
    }
    void PushPrimAttribute(PrimGroupId g, PrimBatchId b, uint32_t attr, Vector2ui range,
                           TransientData data) override
    {
        PrimBatch& pb = CommittedBatch(g, b);
        uint32_t n = (attr == 3) ? pb.primCount : pb.vertexCount;
        if(range[0] > range[1] || range[1] > n) throw MRayError("Sub-batch range is out of bounds");
        PushPrimRange(pb, attr, range[0], range[1], data);
    }
    PrimBatch& CommittedBatch(PrimGroupId g, PrimBatchId b)
    {
        if(GroupOf(U(b)) != U(g)) throw MRayError("PrimBatchId does not belong to the group");
        auto& grp = At(primGroups, U(g), "PrimitiveGroup");
        if(!grp.committed) throw MRayError("PrimitiveGroup is not committed");
        return Batch(U(b));
    }
    static void PushPrimRange(PrimBatch& pb, uint32_t attr, uint32_t lo, uint32_t hi,
                              const TransientData& data)
    {
        size_t n = hi - lo;
        switch(attr)
        {
            case 0: { auto s = Elems<Vector3>(data, n, "position"); std::copy(s.begin(), s.end(), pb.pos.begin() + lo); break; }
            case 1: { auto s = Elems<Quaternion>(data, n, "normal"); std::copy(s.begin(), s.end(), pb.tbn.begin() + lo); break; }
            case 2: { auto s = Elems<Vector2>(data, n, "uv"); std::copy(s.begin(), s.end(), pb.uv.begin() + lo); break; }
            case 3:
            {
                auto s = Elems<Vector3ui>(data, n, "index");
                for(const Vector3ui& t : s)
                    if(t[0] >= pb.vertexCount || t[1] >= pb.vertexCount || t[2] >= pb.vertexCount)
                        throw MRayError("Triangle index is outside its batch's vertex range");
                std::copy(s.begin(), s.end(), pb.idx.begin() + lo);
                break;
            }
            default:
                throw MRayError("(P)Triangle: Unknown Attribute Index " + std::to_string(attr));
        }
    }
    void TransformPrimitives(PrimGroupId g, std::vector<PrimBatchId> batches,
                             std::vector<Matrix3x4> xf) override
    {
        if(batches.size() != xf.size()) throw MRayError("TransformPrimitives: size mismatch");
        for(size_t i = 0; i < batches.size(); i++)
        {
            PrimBatch& pb = CommittedBatch(g, batches[i]);
            pb.baked = Compose(MakeRigid(xf[i].m.data()), pb.baked);
        }
    }

    // =============================== Material ================================
    MatGroupId CreateMaterialGroup(std::string typeName) override
    {
        MatGroup g;
        g.type = MatType(typeName);
        matGroups.push_back(std::move(g));
        return MatGroupId(uint32_t(matGroups.size() - 1));
    }
    MaterialId ReserveMaterial(MatGroupId g, AttributeCountList c) override
    {
        return ReserveMaterials(g, {std::move(c)}).front();
    }
    MaterialIdList ReserveMaterials(MatGroupId g, std::vector<AttributeCountList> counts) override
    {
        auto& grp = At(matGroups, U(g), "MaterialGroup");
        if(U(g) == 0) Unsupported("The passthrough material");
        if(grp.committed) throw MRayError("MaterialGroup is already committed");
        MaterialIdList out;
        for(size_t k = 0; k < counts.size(); k++)
        {
            MRayMaterial m{};
            m.type = grp.type;
            // IoR 1 (vacuum) until a Refract material receives its coefficients
            m.cauchyFront[0] = 1.0f;
            m.cauchyBack[0] = 1.0f;
            out.push_back(MaterialId(Pack(U(g), uint32_t(grp.globalIndex.size()))));
            grp.globalIndex.push_back(uint32_t(materials.size()));
            materials.push_back(m);
        }
        return out;
    }
    void CommitMatReservations(MatGroupId g) override { At(matGroups, U(g), "MaterialGroup").committed = true; }
    bool IsMatCommitted(MatGroupId g) const override { return U(g) < matGroups.size() && matGroups[U(g)].committed; }
    void PushMatAttribute(MatGroupId g, CommonIdRange r, uint32_t attr, TransientData data) override
    {
        auto& grp = At(matGroups, U(g), "MaterialGroup");
        if(!grp.committed) throw MRayError("MaterialGroup is not committed");
        auto info = MatAttribs(grp.type);
        if(attr >= info.size())
            throw MRayError(std::string(kMatNames[grp.type]) + ": Unknown attribute index " + std::to_string(attr));
        if(info[attr].isTexturable != MR_CONSTANT_ONLY)
            throw MRayError(std::string(kMatNames[grp.type]) + ": Attribute " + std::to_string(attr) +
                            " is not \"ConstantOnly\", wrong function is called");
        SetMatConstant(grp, r, attr, data);
    }
    void PushMatAttribute(MatGroupId g, CommonIdRange r, uint32_t attr, TransientData data,
                          std::vector<Optional<TextureId>> tex) override
    {
        auto& grp = At(matGroups, U(g), "MaterialGroup");
        if(!grp.committed) throw MRayError("MaterialGroup is not committed");
        auto info = MatAttribs(grp.type);
        if(attr >= info.size())
            throw MRayError(std::string(kMatNames[grp.type]) + ": Unknown attribute index " + std::to_string(attr));
        for(const auto& t : tex)
            if(t.has_value()) Unsupported("Texture-backed material attributes");
        if(info[attr].isTexturable == MR_TEXTURE_ONLY)
        {
            if(data.ByteSize() != 0) throw MRayError("Texture-only attribute given constant data");
            return;   // optional texture left empty
        }
        SetMatConstant(grp, r, attr, data);
    }
    void PushMatAttribute(MatGroupId g, CommonIdRange, uint32_t attr, std::vector<TextureId> tex) override
    {
        auto& grp = At(matGroups, U(g), "MaterialGroup");
        auto info = MatAttribs(grp.type);
        if(attr >= info.size() || info[attr].isTexturable != MR_TEXTURE_ONLY)
            throw MRayError(std::string(kMatNames[grp.type]) + ": Attribute " + std::to_string(attr) +
                            " is not \"TextureOnly\", wrong function is called");
        if(!tex.empty()) Unsupported("Texture-backed material attributes");
    }
    void SetMatConstant(MatGroup& grp, CommonIdRange r, uint32_t attr, const TransientData& data)
    {
        uint32_t gi = uint32_t(&grp - matGroups.data());
        CheckRange(r, gi, grp.globalIndex.size());
        size_t n = IndexOf(r[1]) - IndexOf(r[0]) + 1;
        auto info = MatAttribs(grp.type);
        for(size_t k = 0; k < n; k++)
        {
            MRayMaterial& m = materials[grp.globalIndex[IndexOf(r[0]) + k]];
            const std::string& name = info[attr].name;
            if(info[attr].dataType == MRayDataEnum::MR_VECTOR_3)
            {
                const Vector3& v = Elems<Vector3>(data, n, name.c_str())[k];
                float* dst = (name == "albedo") ? m.albedo : (name == "cauchyBack") ? m.cauchyBack : m.cauchyFront;
                std::copy_n(v.v.data(), 3, dst);
            }
            else
            {
                float v = Elems<float>(data, n, name.c_str())[k];
                if(name == "roughness") m.roughness = v;
                else if(name == "specular") m.specular = v;
                else m.metallic = v;
            }
        }
    }

    // =============================== Texture =================================
    TextureId CreateTexture2D(Vector2ui, uint32_t, MRayTextureParameters) override { Unsupported("Textures"); }
    TextureId CreateTexture3D(Vector3ui, uint32_t, MRayTextureParameters) override { Unsupported("Textures"); }
    void CommitTextures() override {}
    void PushTextureData(TextureId, uint32_t, TransientData) override { Unsupported("Textures"); }

    // ============================== Transform ================================
    TransGroupId CreateTransformGroup(std::string typeName) override
    {
        TransGroup g;
        if(typeName == kTransIdentity) g.identity = true;
        else if(typeName == kTransSingle) g.identity = false;
        else throw MRayError("Unknown transform type \"" + typeName + "\"");
        transGroups.push_back(std::move(g));
        return TransGroupId(uint32_t(transGroups.size() - 1));
    }
    TransformId ReserveTransformation(TransGroupId g, AttributeCountList c) override
    {
        return ReserveTransformations(g, {std::move(c)}).front();
    }
    TransformIdList ReserveTransformations(TransGroupId g, std::vector<AttributeCountList> counts) override
    {
        auto& grp = At(transGroups, U(g), "TransformGroup");
        if(grp.committed) throw MRayError("TransformGroup is already committed");
        TransformIdList out;
        for(size_t k = 0; k < counts.size(); k++)
        {
            out.push_back(TransformId(Pack(U(g), uint32_t(grp.xf.size()))));
            grp.xf.push_back(IdentityRigid());
        }
        return out;
    }
    void CommitTransReservations(TransGroupId g) override { At(transGroups, U(g), "TransformGroup").committed = true; }
    bool IsTransCommitted(TransGroupId g) const override { return U(g) < transGroups.size() && transGroups[U(g)].committed; }
    void PushTransAttribute(TransGroupId g, CommonIdRange r, uint32_t attr, TransientData data) override
    {
        auto& grp = At(transGroups, U(g), "TransformGroup");
        if(!grp.committed) throw MRayError("TransformGroup is not committed");
        if(grp.identity || attr != 0)
            throw MRayError(std::string(grp.identity ? kTransIdentity : kTransSingle) +
                            ": Unknown attribute index " + std::to_string(attr));
        CheckRange(r, U(g), grp.xf.size());
        size_t n = IndexOf(r[1]) - IndexOf(r[0]) + 1;
        auto s = Elems<Matrix4x4>(data, n, "Transform");
        for(size_t k = 0; k < n; k++)
        {
            const float* m = s[k].m.data();
            if(m[12] != 0.0f || m[13] != 0.0f || m[14] != 0.0f || m[15] != 1.0f)
                Unsupported("A projective transform");
            grp.xf[IndexOf(r[0]) + k] = MakeRigid(m, true);
        }
    }

    // ================================ Lights =================================
    LightGroupId CreateLightGroup(std::string typeName, PrimGroupId pg) override
    {
        LightGroup g;
        g.kind = LightKind(typeName);
        if(g.kind == 0) return LightGroupId(0);   // the implicit Null group
        if(g.kind == 1)
        {
            if(U(pg) == 0 || U(pg) >= primGroups.size())
                throw MRayError("Primitive-backed light needs a primitive group");
            g.primGroup = U(pg);
        }
        lightGroups.push_back(std::move(g));
        return LightGroupId(uint32_t(lightGroups.size() - 1));
    }
    LightId ReserveLight(LightGroupId g, AttributeCountList c, PrimBatchId b) override
    {
        return ReserveLights(g, {std::move(c)}, {b}).front();
    }
    LightIdList ReserveLights(LightGroupId g, std::vector<AttributeCountList> counts,
                              std::vector<PrimBatchId> batches) override
    {
        auto& grp = At(lightGroups, U(g), "LightGroup");
        if(U(g) == 0) return LightIdList(counts.size(), LightId(0));   // Null light
        if(grp.committed) throw MRayError("LightGroup is already committed");
        if(grp.kind == 1 && batches.size() != counts.size())
            throw MRayError("Primitive-backed lights need one primitive batch each");
        LightIdList out;
        for(size_t k = 0; k < counts.size(); k++)
        {
            LightEntry e;
            if(grp.kind == 1)
            {
                if(GroupOf(U(batches[k])) != grp.primGroup)
                    throw MRayError("Light's primitive batch is not in the light's primitive group");
                e.primBatch = U(batches[k]);
            }
            out.push_back(LightId(Pack(U(g), uint32_t(grp.lights.size()))));
            grp.lights.push_back(e);
        }
        return out;
    }
    void CommitLightReservations(LightGroupId g) override { At(lightGroups, U(g), "LightGroup").committed = true; }
    bool IsLightCommitted(LightGroupId g) const override { return U(g) < lightGroups.size() && (U(g) == 0 || lightGroups[U(g)].committed); }
    void PushLightAttribute(LightGroupId g, CommonIdRange r, uint32_t attr, TransientData data) override
    {
        auto& grp = At(lightGroups, U(g), "LightGroup");
        if(grp.kind == 1 && attr == 1)
        {
            if(!grp.committed) throw MRayError("LightGroup is not committed");
            CheckRange(r, U(g), grp.lights.size());
            size_t n = IndexOf(r[1]) - IndexOf(r[0]) + 1;
            auto s = Elems<bool>(data, n, "isTwoSided");
            for(size_t k = 0; k < n; k++) grp.lights[IndexOf(r[0]) + k].twoSided = s[k];
            return;
        }
        throw MRayError(TypeName(g) + ": Attribute " + std::to_string(attr) +
                        " is not \"ConstantOnly\", wrong function is called");
    }
    void PushLightAttribute(LightGroupId g, CommonIdRange r, uint32_t attr, TransientData data,
                            std::vector<Optional<TextureId>> tex) override
    {
        auto& grp = At(lightGroups, U(g), "LightGroup");
        if(grp.kind == 0 || attr != 0)
            throw MRayError(TypeName(g) + ": Attribute " + std::to_string(attr) +
                            " is not \"ParamVarying\", wrong function is called");
        if(!grp.committed) throw MRayError("LightGroup is not committed");
        for(const auto& t : tex)
            if(t.has_value()) Unsupported("Texture-backed light radiance");
        CheckRange(r, U(g), grp.lights.size());
        size_t n = IndexOf(r[1]) - IndexOf(r[0]) + 1;
        auto s = Elems<Vector3>(data, n, "radiance");
        for(size_t k = 0; k < n; k++) std::copy_n(s[k].v.data(), 3, grp.lights[IndexOf(r[0]) + k].radiance);
    }
    void PushLightAttribute(LightGroupId, CommonIdRange, uint32_t, std::vector<TextureId>) override
    {
        Unsupported("Texture-backed light radiance");
    }

    // =============================== Cameras =================================
    CameraGroupId CreateCameraGroup(std::string typeName) override
    {
        if(typeName != kCamPinhole) throw MRayError("Unknown camera type \"" + typeName + "\"");
        camGroups.emplace_back();
        return CameraGroupId(uint32_t(camGroups.size() - 1));
    }
    CameraId ReserveCamera(CameraGroupId g, AttributeCountList c) override
    {
        return ReserveCameras(g, {std::move(c)}).front();
    }
    CameraIdList ReserveCameras(CameraGroupId g, std::vector<AttributeCountList> counts) override
    {
        auto& grp = At(camGroups, U(g), "CameraGroup");
        if(U(g) == 0) throw MRayError("Unknown CameraGroupId 0");
        CameraIdList out;
        for(size_t k = 0; k < counts.size(); k++)
        {
            out.push_back(CameraId(Pack(U(g), uint32_t(grp.cams.size()))));
            grp.cams.emplace_back();
        }
        return out;
    }
    void CommitCamReservations(CameraGroupId g) override { At(camGroups, U(g), "CameraGroup").committed = true; }
    bool IsCamCommitted(CameraGroupId g) const override { return U(g) < camGroups.size() && camGroups[U(g)].committed; }
    void PushCamAttribute(CameraGroupId g, CommonIdRange r, uint32_t attr, TransientData data) override
    {
        auto& grp = At(camGroups, U(g), "CameraGroup");
        if(!grp.committed) throw MRayError("CameraGroup is not committed");
        CheckRange(r, U(g), grp.cams.size());
        size_t n = IndexOf(r[1]) - IndexOf(r[0]) + 1;
        for(size_t k = 0; k < n; k++)
        {
            CameraEntry& c = grp.cams[IndexOf(r[0]) + k];
            switch(attr)
            {
                case 0: std::copy_n(Elems<Vector4>(data, n, "FovAndPlanes")[k].v.data(), 4, c.fovPlanes); break;
                case 1: c.gaze = Elems<Vector3>(data, n, "gaze")[k]; break;
                case 2: c.position = Elems<Vector3>(data, n, "position")[k]; break;
                case 3: c.up = Elems<Vector3>(data, n, "up")[k]; break;
                default: throw MRayError("(C)Pinhole: Unknown attribute index " + std::to_string(attr));
            }
        }
    }

    // ================================ Medium =================================
    MediumGroupId CreateMediumGroup(std::string) override { Unsupported("Mediums"); }
    MediumId ReserveMedium(MediumGroupId, AttributeCountList) override { Unsupported("Mediums"); }
    MediumIdList ReserveMediums(MediumGroupId, std::vector<AttributeCountList>) override { Unsupported("Mediums"); }
    void CommitMediumReservations(MediumGroupId) override { Unsupported("Mediums"); }
    bool IsMediumCommitted(MediumGroupId) const override { return false; }
    void PushMediumAttribute(MediumGroupId, CommonIdRange, uint32_t, TransientData) override { Unsupported("Mediums"); }
    void PushMediumAttribute(MediumGroupId, CommonIdRange, uint32_t, TransientData,
                             std::vector<Optional<TextureId>>) override { Unsupported("Mediums"); }
    void PushMediumAttribute(MediumGroupId, CommonIdRange, uint32_t, std::vector<TextureId>) override { Unsupported("Mediums"); }

    // =============================== Surfaces ================================
    SurfaceId CreateSurface(SurfaceParams p) override
    {
        if(p.primBatches.empty() || p.primBatches.size() > MaxPrimBatchPerSurface ||
           p.materials.size() != p.primBatches.size())
            throw MRayError("Surface needs 1..8 primitive batches, one material each");
        for(const auto& a : p.alphaMaps)
            if(a.has_value()) Unsupported("Alpha maps");
        for(const auto& v : p.volumes)
            if(U(v) != UINT32_MAX && U(v) != 0) Unsupported("Volumes");
        for(size_t k = 0; k < p.primBatches.size(); k++)
        {
            Batch(U(p.primBatches[k]));
            uint32_t mg = GroupOf(U(p.materials[k]));
            if(mg == 0 || mg >= matGroups.size() || IndexOf(U(p.materials[k])) >= matGroups[mg].globalIndex.size())
                throw MRayError("Surface references an unknown MaterialId");
        }
        TransformOf(p.transformId);
        SurfaceRec s;
        s.params = std::move(p);
        surfaces.push_back(std::move(s));
        return SurfaceId(uint32_t(surfaces.size() - 1));
    }
    LightSurfaceId SetBoundarySurface(LightId l, TransformId t) override
    {
        uint32_t g = GroupOf(U(l));
        if(g >= lightGroups.size() || (g != 0 && lightGroups[g].kind < 2))
            throw MRayError("Boundary light must be a (L)Null or (L)Skysphere light");
        if(g != 0 && IndexOf(U(l)) >= lightGroups[g].lights.size()) throw MRayError("Unknown LightId");
        TransformOf(t);
        hasBoundary = true;
        boundaryLight = U(l);
        boundaryTransform = U(t);
        return LightSurfaceId(0);
    }
    LightSurfaceId CreateLightSurface(LightSurfaceParams p) override
    {
        uint32_t g = GroupOf(U(p.lightId));
        if(g == 0 || g >= lightGroups.size() || lightGroups[g].kind != 1)
            throw MRayError("Light surfaces must use primitive-backed lights");
        if(IndexOf(U(p.lightId)) >= lightGroups[g].lights.size()) throw MRayError("Unknown LightId");
        TransformOf(p.transformId);
        SurfaceRec s;
        s.isLight = true;
        s.light = std::move(p);
        surfaces.push_back(std::move(s));
        return LightSurfaceId(uint32_t(surfaces.size()));   // 0 is the boundary
    }
    CamSurfaceId CreateCameraSurface(CameraSurfaceParams p) override
    {
        uint32_t g = GroupOf(U(p.cameraId));
        if(g == 0 || g >= camGroups.size() || IndexOf(U(p.cameraId)) >= camGroups[g].cams.size())
            throw MRayError("Unknown CameraId");
        TransformOf(p.transformId);
        camSurfaces.push_back(std::move(p));
        return CamSurfaceId(uint32_t(camSurfaces.size() - 1));
    }

    // Flatten every surface (creation order) into one triangle soup.
    void Flatten()
    {
        Flat f;
        std::vector<uint8_t> used;
        auto emitBatch = [&](const PrimBatch& pb, const Rigid& xf, uint32_t key, bool cull)
        {
            Rigid r = Compose(xf, pb.baked, true);
            uint32_t base = uint32_t(f.pos.size() / 3);
            for(uint32_t v = 0; v < pb.vertexCount; v++)
            {
                float p[3], q[4];
                ApplyPoint(r, pb.pos[v].v.data(), p);
                // TransformPrimitives turns the frames with the vertices
                // (KCApplyTransformsTriangle); a surface's transform does not
                // (Triangle::GenerateSurface leaves the TBN in mesh space)
                ApplyTbn(pb.baked, pb.tbn[v].v.v.data(), q);
                f.pos.insert(f.pos.end(), p, p + 3);
                f.tbn.insert(f.tbn.end(), q, q + 4);
            }
            for(uint32_t t = 0; t < pb.primCount; t++)
            {
                for(int c = 0; c < 3; c++) f.idx.push_back(base + pb.idx[t][c]);
                f.keys.push_back(key == MRAY_LIGHT_FLAG ? (MRAY_LIGHT_FLAG | uint32_t(f.lightTris.size())) : key);
                f.cull.push_back(cull ? 1 : 0);
                if(key == MRAY_LIGHT_FLAG) f.lightTris.push_back(uint32_t(f.keys.size() - 1));
            }
        };
        for(const SurfaceRec& s : surfaces)
        {
            if(!s.isLight)
            {
                Rigid xf = TransformOf(s.params.transformId);
                for(size_t k = 0; k < s.params.primBatches.size(); k++)
                {
                    uint32_t m = U(s.params.materials[k]);
                    uint32_t gm = matGroups[GroupOf(m)].globalIndex[IndexOf(m)];
                    bool cull = k < s.params.cullFaceFlags.size() ? bool(s.params.cullFaceFlags[k]) : true;
                    emitBatch(Batch(U(s.params.primBatches[k])), xf, gm, cull);
                }
            }
            else
            {
                const LightGroup& lg = lightGroups[GroupOf(U(s.light.lightId))];
                const LightEntry& le = lg.lights[IndexOf(U(s.light.lightId))];
                size_t first = f.lightTris.size();
                emitBatch(Batch(le.primBatch), TransformOf(s.light.transformId), MRAY_LIGHT_FLAG, false);
                for(size_t k = first; k < f.lightTris.size(); k++)
                {
                    f.lightRad.insert(f.lightRad.end(), le.radiance, le.radiance + 3);
                    f.lightTwoSided.push_back(le.twoSided ? 1 : 0);
                }
            }
        }
        if(f.keys.empty()) throw MRayError("CommitSurfaces: the scene has no triangles");
        for(uint32_t k : f.keys)
            if(!(k & MRAY_LIGHT_FLAG) && k >= materials.size()) throw MRayError("Bad material key");
        if(hasBoundary && GroupOf(boundaryLight) != 0)
        {
            const LightEntry& le = lightGroups[GroupOf(boundaryLight)].lights[IndexOf(boundaryLight)];
            f.boundaryType = lightGroups[GroupOf(boundaryLight)].kind == 3
                             ? MRAY_BOUNDARY_SKYSPHERE_SPHERICAL : MRAY_BOUNDARY_SKYSPHERE_COOCTA;
            std::copy_n(le.radiance, 3, f.boundaryRad);
            // a constant sky is invariant under the boundary transform
        }
        // scene AABB and the boundary light's scene diameter: length of the
        // XZ span (TracerBase.cpp:1651-1664)
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for(uint32_t i : f.idx)
            for(int c = 0; c < 3; c++)
            {
                lo[c] = std::min(lo[c], f.pos[size_t(i) * 3 + c]);
                hi[c] = std::max(hi[c], f.pos[size_t(i) * 3 + c]);
            }
        for(int c = 0; c < 3; c++) { f.aabb.min.v[c] = lo[c]; f.aabb.max.v[c] = hi[c]; }
        float sx = hi[0] - lo[0], sz = hi[2] - lo[2];
        f.sceneDiameter = std::sqrt(sx * sx + sz * sz);
        flat = std::move(f);
    }

    MRaySceneDesc Desc(const MRayCamera& cam)
    {
        MRaySceneDesc d{};
        d.vertexCount = uint32_t(flat.pos.size() / 3);
        d.triangleCount = uint32_t(flat.keys.size());
        d.hPositions = flat.pos.data();
        d.hTbnQuats = flat.tbn.data();
        d.hIndices = flat.idx.data();
        d.hSurfaceKeys = flat.keys.data();
        d.hCullBackFace = flat.cull.data();
        d.materialCount = uint32_t(materials.size());
        d.hMaterials = materials.data();
        d.lightCount = uint32_t(flat.lightTris.size());
        d.hLightTriangles = flat.lightTris.data();
        d.hLightRadiance = flat.lightRad.data();
        d.hLightTwoSided = flat.lightTwoSided.data();
        d.boundaryType = flat.boundaryType;
        std::copy_n(flat.boundaryRad, 3, d.boundaryRadiance);
        d.sceneDiameter = flat.sceneDiameter;
        d.camera = cam;
        return d;
    }

    SurfaceCommitResult CommitSurfaces() override
    {
        Release();
        Flatten();
        MRaySceneDesc d = Desc(MRayCamera{});
        CheckRc(mray_accel_create(&d, &accel), "CommitSurfaces");
        committed = true;
        return SurfaceCommitResult{flat.aabb, surfaces.size(), 1};
    }
    // Host-side products of CommitSurfaces without the GPU build (tests).
    void FlattenOnly() { Flatten(); }
    const Flat& Flattened() const { return flat; }
    const std::vector<MRayMaterial>& Materials() const { return materials; }

    MRayCamera CameraOf(uint32_t camSurf) const
    {
        if(camSurf >= camSurfaces.size()) throw MRayError("Unknown CamSurfaceId");
        const CameraSurfaceParams& cs = camSurfaces[camSurf];
        const CameraEntry& c = camGroups[GroupOf(U(cs.cameraId))].cams[IndexOf(U(cs.cameraId))];
        Rigid xf = TransformOf(cs.transformId);
        MRayCamera cam{};
        ApplyPoint(xf, c.position.v.data(), cam.position);
        ApplyPoint(xf, c.gaze.v.data(), cam.gaze);
        float up[3] = {c.up.v[0], c.up.v[1], c.up.v[2]};
        if(!xf.identity)
            for(int i = 0; i < 3; i++)
                cam.up[i] = xf.m[i * 4 + 0] * up[0] + xf.m[i * 4 + 1] * up[1] + xf.m[i * 4 + 2] * up[2];
        else std::copy_n(up, 3, cam.up);
        cam.fov[0] = c.fovPlanes[0];
        cam.fov[1] = c.fovPlanes[1];
        cam.nearFar[0] = c.fovPlanes[2];
        cam.nearFar[1] = c.fovPlanes[3];
        return cam;
    }
    CameraTransform GetCamTransform(CamSurfaceId s) const override
    {
        MRayCamera c = CameraOf(U(s));
        CameraTransform t;
        std::copy_n(c.position, 3, t.position.v.data());
        std::copy_n(c.gaze, 3, t.gazePoint.v.data());
        std::copy_n(c.up, 3, t.up.v.data());
        return t;
    }
    VolumeId RegisterVolume(VolumeParams) override { Unsupported("Volumes"); }
    VolumeIdList RegisterVolumes(std::vector<VolumeParams>) override { Unsupported("Volumes"); }
    void SetBoundaryVolume(VolumeId) override {}

    // =============================== Renderers ===============================
    RendererId CreateRenderer(std::string typeName) override
    {
        RendererRec r;
        r.spectral = RendererKind(typeName);
        uint32_t id = nextRenderer++;
        renderers[id] = r;
        return RendererId(id);
    }
    void DestroyRenderer(RendererId r) override
    {
        if(!renderers.erase(U(r))) throw MRayError("Unknown RendererId");
        if(renderer && activeRenderer == U(r)) StopRender();
    }
    void PushRendererAttribute(RendererId id, uint32_t attr, TransientData data) override
    {
        auto it = renderers.find(U(id));
        if(it == renderers.end()) throw MRayError("Unknown RendererId");
        RendererRec& r = it->second;
        const std::string name = TypeName(id);
        // PathTracerRenderer.cu:127-142
        switch(attr)
        {
            case 0: r.totalSPP = Elems<uint32_t>(data, 1, "totalSPP")[0]; break;
            case 1: r.burstSize = std::max(1u, Elems<uint32_t>(data, 1, "burstSize")[0]); break;
            case 2:
            {
                std::string_view m = std::as_const(data).AccessAsString();
                if(m != "Throughput" && m != "Latency")
                    throw MRayError("Unknown render mode \"" + std::string(m) + "\"");
                r.latency = (m == "Latency");
                break;
            }
            case 3: r.sampleMode = ParseSampleMode(std::as_const(data).AccessAsString()); break;
            case 4:
            {
                Vector2ui v = Elems<Vector2ui>(data, 1, "rrRange")[0];
                r.rr[0] = v[0];
                r.rr[1] = v[1];
                break;
            }
            case 5:
            {
                // LightSamplerNames (PathTracerRendererBase.h:50-60) hold only
                // "Uniform"; "Power" is this tracer's alias-table sampler
                std::string_view s = std::as_const(data).AccessAsString();
                if(s == "Uniform") r.powerLights = false;
                else if(s == "Power") r.powerLights = true;
                else Unsupported("Light sampler \"" + std::string(s) + "\"");
                break;
            }
            default:
                throw MRayError(name + " Unknown attribute index " + std::to_string(attr));
        }
    }

    // =============================== Rendering ===============================
    void SetupRenderEnv(TimelineSemaphore* semaphore, uint32_t alignment,
                        uint64_t initialAcquireValue) override
    {
        // TracerBase::SetupRenderEnv -> RenderImage (TracerBase.cpp:1701-1707)
        sem = semaphore;
        importAlignment = std::max<uint32_t>(64u, alignment);
        acquireValue = initialAcquireValue;
    }

    RenderBufferInfo StartRender(RendererId rid, CamSurfaceId cs, RenderImageParams ip,
                                 Optional<uint32_t>, Optional<uint32_t>) override
    {
        if(!committed) throw MRayError("StartRender before CommitSurfaces");
        auto it = renderers.find(U(rid));
        if(it == renderers.end()) throw MRayError("Unknown RendererId");
        if(ip.regionMin[0] != 0 || ip.regionMin[1] != 0 ||
           (ip.regionMax[0] != 0 && ip.regionMax[0] != ip.resolution[0]) ||
           (ip.regionMax[1] != 0 && ip.regionMax[1] != ip.resolution[1]))
            Unsupported("Partial-region rendering");
        if(params.filmFilter.type != FilterType::GAUSSIAN) Unsupported("A non-Gaussian film filter");
        if(params.samplerType.e != SamplerType::INDEPENDENT) Unsupported("A non-independent sampler");
        if(it->second.spectral && params.wavelengthSampleMode.e != WavelengthSampleMode::HYPERBOLIC_PBRT)
            Unsupported("A wavelength sampler other than HYPERBOLIC_PBRT");
        StopRender();
        const RendererRec& rr = it->second;
        MRayRenderParams p{};
        p.width = ip.resolution[0];
        p.height = ip.resolution[1];
        p.totalSPP = rr.totalSPP;
        p.rrRange[0] = rr.rr[0];
        p.rrRange[1] = rr.rr[1];
        p.sampleMode = rr.sampleMode;
        p.spectral = rr.spectral ? 1 : 0;
        p.maxRayCount = params.parallelizationHint;
        p.seed = params.seed;
        p.filterRadius = params.filmFilter.radius;
        p.flags = rr.powerLights ? MRAY_RENDER_LIGHT_POWER : 0u;
        MRaySceneDesc d = Desc(CameraOf(U(cs)));
        // PathTracerRendererBase::DoRender (PathTracerRendererBase.cu:507-533):
        // Throughput on one tile with burstSize 1 renders bounce by bounce
        // over the whole image; everything else renders a tile at a time,
        // burstSize samples (Throughput) or one sample (Latency) per call.
        MRayRenderer* tr = nullptr;
        MRayRenderParams pt = p;
        pt.flags |= MRAY_RENDER_TILED;
        CheckRc(mray_renderer_create(accel, &d, &pt, &tr), "StartRender");
        MRayTileInfo ti{};
        CheckRc(mray_renderer_tiling(tr, &ti), "StartRender");
        tileCount = ti.tileCount[0] * ti.tileCount[1];
        tiled = rr.latency || rr.burstSize > 1 || tileCount > 1;
        if(tiled) renderer = tr;
        else
        {
            mray_renderer_destroy(tr);
            CheckRc(mray_renderer_create(accel, &d, &p, &renderer), "StartRender");
        }
        curTile = 0;
        tileSPPs.assign(tileCount, 0u);
        activeRenderer = U(rid);
        imgParams = ip;
        // a tile's planes, packed (the whole image when it is one tile)
        size_t P = tiled ? size_t(ti.tileSize[0]) * ti.tileSize[1] : size_t(p.width) * p.height;
        staging.Allocate(4 * P * sizeof(float), importAlignment);
        sampleCounter = 0;
        saveImage = true;
        RenderBufferInfo info{};
        info.data = reinterpret_cast<const std::byte*>(staging.p);
        info.totalSize = staging.bytes;
        info.renderColorSpace = params.globalTextureColorSpace;
        info.resolution = ip.resolution;
        info.curRenderLogic0 = rr.sampleMode;
        info.curRenderLogic1 = UINT32_MAX;
        return info;
    }
    void SetCameraTransform(RendererId rid, CameraTransform) override
    {
        if(!renderers.count(U(rid))) throw MRayError("Unknown RendererId");
        Unsupported("Moving the camera during a render");
    }
    void StopRender() override
    {
        if(renderer) { mray_renderer_destroy(renderer); renderer = nullptr; }
    }
    RendererOutput DoRenderWork() override
    {
        if(!renderer) return RendererOutput{};
        const RendererRec& rr = renderers.at(activeRenderer);
        if(tiled) return DoTileWork(rr);
        auto t0 = std::chrono::steady_clock::now();
        CheckRc(mray_renderer_iterate(renderer, rr.burstSize, nullptr), "DoRenderWork");
        MRayRenderStats st{};
        CheckRc(mray_renderer_stats(renderer, &st), "DoRenderWork");
        double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        lastIterMs = ms;
        sampleCounter = st.totalDeadPaths;
        size_t P = size_t(imgParams.resolution[0]) * imgParams.resolution[1];
        RendererAnalyticData a{};
        a.throughput = double(st.deadPaths) / (ms * 1e-3) / 1e6;   // PathTracerRendererBase.cu:305-311
        a.throughputSuffix = "M path/s";
        a.workPerPixel = double(sampleCounter) / double(P);
        a.wppLimit = rr.totalSPP;
        a.workPerPixelSuffix = "spp";
        a.iterationTimeMS = float(ms);
        a.renderResolution = imgParams.resolution;
        a.outputColorSpace = params.globalTextureColorSpace;
        a.usedGPUMemoryBytes = UsedDeviceMemory();
        a.customLogicSize0 = 3;                                    // sample modes
        RendererOutput out;
        out.analytics = a;
        // RenderImage::TransferToHost (RenderImage.cpp:164-209): wait until
        // the consumer has released the staging buffer, fill it, release,
        // and tell the consumer which value to wait for.  A semaphore that
        // cannot be acquired means the consumer is closing: no section.
        if(sem && !sem->Acquire(acquireValue)) return out;
        CheckRc(mray_renderer_take_film(renderer, staging.p), "DoRenderWork");
        uint64_t waitCounter = 0;
        if(sem)
        {
            sem->Release();
            acquireValue += 2;
            waitCounter = acquireValue - 1;
        }
        RenderImageSection sec{};
        sec.pixelMin = Vector2ui{{0, 0}};
        sec.pixelMax = imgParams.resolution;
        sec.globalWeight = 1.0f;                                   // PathTracerRendererBase.cu:260
        sec.waitCounter = waitCounter;
        sec.pixStartOffsets = {0, P * sizeof(float), 2 * P * sizeof(float)};
        sec.weightStartOffset = 3 * P * sizeof(float);
        out.imageOut = sec;
        // one save request when the sample limit is reached
        // (PathTracerRenderer.cu:1063-1065)
        out.triggerSave = saveImage && st.finished != 0;
        if(out.triggerSave) saveImage = false;
        return out;
    }

    // PathTracerRendererT::DoLatencyRender (PathTracerRenderer.cu:1076-1180):
    // passCount more samples of the current tile (capped at totalSPP until
    // the save was requested), that tile's section handed over with global
    // weight 1, then the next tile; the save is requested once every tile
    // has totalSPP samples.
    RendererOutput DoTileWork(const RendererRec& rr)
    {
        const uint32_t passCount = rr.latency ? 1u : rr.burstSize;
        const uint32_t t = curTile;
        uint32_t sppLimit = tileSPPs[t] + passCount;
        if(saveImage) sppLimit = std::min(sppLimit, rr.totalSPP);
        sppLimit = std::max(sppLimit, tileSPPs[t]);
        auto t0 = std::chrono::steady_clock::now();
        MRayTileResult tile{};
        MRayRenderStats st{};
        CheckRc(mray_renderer_render_tile(renderer, sppLimit - tileSPPs[t], &tile, &st), "DoRenderWork");
        tileSPPs[t] = sppLimit;
        curTile = (t + 1u) % tileCount;
        const uint32_t w = tile.pixelMax[0] - tile.pixelMin[0], h = tile.pixelMax[1] - tile.pixelMin[1];
        RendererOutput out;
        if(sem && !sem->Acquire(acquireValue)) return out;
        CheckRc(mray_renderer_take_tile(renderer, staging.p), "DoRenderWork");
        uint64_t waitCounter = 0;
        if(sem)
        {
            sem->Release();
            acquireValue += 2;
            waitCounter = acquireValue - 1;
        }
        double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        lastIterMs = ms;
        // CalculateAnalyticDataLatency (PathTracerRendererBase.cu:328-356)
        uint64_t sppSum = 0;
        for(uint32_t v : tileSPPs) sppSum += v;
        RendererAnalyticData a{};
        a.throughput = double(st.deadPaths) / (ms * 1e-3) / 1e6;
        a.throughputSuffix = "M path/s";
        a.workPerPixel = double(sppSum) / double(tileCount);
        a.wppLimit = rr.totalSPP;
        a.workPerPixelSuffix = "spp";
        a.iterationTimeMS = float(ms);
        a.renderResolution = imgParams.resolution;
        a.outputColorSpace = params.globalTextureColorSpace;
        a.usedGPUMemoryBytes = UsedDeviceMemory();
        a.customLogicSize0 = 3;
        out.analytics = a;
        const size_t P = size_t(w) * h;
        RenderImageSection sec{};
        sec.pixelMin = Vector2ui{{tile.pixelMin[0], tile.pixelMin[1]}};
        sec.pixelMax = Vector2ui{{tile.pixelMax[0], tile.pixelMax[1]}};
        sec.globalWeight = 1.0f;
        sec.waitCounter = waitCounter;
        sec.pixStartOffsets = {0, P * sizeof(float), 2 * P * sizeof(float)};
        sec.weightStartOffset = 3 * P * sizeof(float);
        out.imageOut = sec;
        out.triggerSave = saveImage && sppSum == uint64_t(rr.totalSPP) * tileCount;
        if(out.triggerSave) saveImage = false;
        return out;
    }

    // ================================ Misc ===================================
    void ClearAll() override
    {
        StopRender();
        Release();
        primGroups.assign(1, PrimGroup{});
        primGroups[0].committed = true;
        matGroups.assign(1, MatGroup{});
        materials.clear();
        transGroups.assign(1, TransGroup{});
        transGroups[0].identity = true;
        transGroups[0].committed = true;
        lightGroups.assign(1, LightGroup{});
        lightGroups[0].committed = true;
        camGroups.assign(1, CamGroup{});
        surfaces.clear();
        camSurfaces.clear();
        hasBoundary = false;
        renderers.clear();
        nextRenderer = 0;
        committed = false;
        flat = Flat{};
    }
    void Flush() const override {}
    GPUThreadInitFunction GetThreadInitFunction() const override { return []() {}; }
    void SetThreadPool(ThreadPool&) override {}
    size_t TotalDeviceMemory() const override
    {
        size_t total = 0, used = 0;
        CheckRc(mray_device_memory(&total, &used), "TotalDeviceMemory");
        return total;
    }
    size_t UsedDeviceMemory() const override
    {
        size_t total = 0, used = 0;
        CheckRc(mray_device_memory(&total, &used), "UsedDeviceMemory");
        return used;
    }
    const TracerParameters& Parameters() const override { return params; }
};

}  // namespace
}  // namespace mray_amd

extern "C" mray_amd::TracerI* ConstructTracer(const mray_amd::TracerParameters& params)
{
    return new mray_amd::Tracer(params);
}

extern "C" void DestroyTracer(mray_amd::TracerI* tracer)
{
    delete tracer;
}

// Test hook: run CommitSurfaces' host half (no GPU) and write the flattened
// scene to `path` in the layout read by tests/cpp (little-endian):
// u32 V, T, M, L, boundaryType; f32 boundaryRadiance[3], sceneDiameter;
// f32 pos[V*3], tbn[V*4]; u32 idx[T*3], keys[T]; u8 cull[T];
// MRayMaterial[M]; u32 lightTris[L]; f32 lightRad[L*3]; u8 twoSided[L].
extern "C" MRAY_AMD_TRACER_API int mray_tracer_flatten_to_file(mray_amd::TracerI* tracer, const char* path)
{
    auto* t = dynamic_cast<mray_amd::Tracer*>(tracer);
    if(!t) return MRAY_ERR_INVALID;
    try
    {
        t->FlattenOnly();
    }
    catch(const MRayError& e)
    {
        std::fprintf(stderr, "flatten: %s\n", e.what());
        return MRAY_ERR_INVALID;
    }
    const auto& f = t->Flattened();
    const auto& mats = t->Materials();
    FILE* fp = std::fopen(path, "wb");
    if(!fp) return MRAY_ERR_INVALID;
    uint32_t hdr[5] = {uint32_t(f.pos.size() / 3), uint32_t(f.keys.size()), uint32_t(mats.size()),
                       uint32_t(f.lightTris.size()), f.boundaryType};
    std::fwrite(hdr, 4, 5, fp);
    std::fwrite(f.boundaryRad, 4, 3, fp);
    std::fwrite(&f.sceneDiameter, 4, 1, fp);
    // fwrite's buffer must not be NULL, even for zero items (an empty vector's data())
    auto put = [fp](const auto& v)
    {
        if(!v.empty()) std::fwrite(v.data(), sizeof(v[0]), v.size(), fp);
    };
    put(f.pos);
    put(f.tbn);
    put(f.idx);
    put(f.keys);
    put(f.cull);
    put(mats);
    put(f.lightTris);
    put(f.lightRad);
    put(f.lightTwoSided);
    std::fclose(fp);
    return MRAY_OK;
}
