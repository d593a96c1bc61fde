// mray_tracer.hpp -- C++20 host interface of the MI355X path tracer,
// mirroring the reference's TracerI plugin interface (Core/TracerI.h:150-376).
//
// The object returned by ConstructTracer() takes the same calls, in the same
// order and with the same argument meaning, as the reference's TracerI:
// type queries, group creation / reservation / commit, attribute pushes with
// TransientData, surface creation, CommitSurfaces, renderer creation,
// attribute pushes, StartRender / DoRenderWork / StopRender.  Errors are
// reported the reference's way: an MRayError exception thrown across the
// call (TracerBase.cpp:1713-1717, PathTracerRenderer.cu:139-140).
//
// The types are this library's own, laid out for the calls the path tracer
// needs; an adapter compiled inside the reference tree forwards its TracerI
// onto this one (each method forwards 1:1).  All GPU work happens below the
// C-ABI of mray_amd.h; this layer only assembles the flat scene description.
//
// Supported types (everything else throws MRayError):
//   (P)Triangle                      positions, TBN quaternions, uv, indices
//   (Mt)Lambert (Mt)Reflect (Mt)Refract (Mt)Unreal   constant parameters
//   (L)Prim(P)Triangle (L)Skysphere_CoOcta (L)Null   constant radiance
//   (C)Pinhole
//   (T)Identity (T)Single                            baked into positions
//   (R)PathTracerRGB (R)PathTracerSpectral
#pragma once

#include <array>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <optional>
#include <span>
#include <stdexcept>
#include <string>
#include <string_view>
#include <type_traits>
#include <utility>
#include <vector>

namespace mray_amd
{

// ---------------------------------------------------------------------------
// Small value types
// ---------------------------------------------------------------------------
template<class T> using Optional = std::optional<T>;

struct Vector2ui { uint32_t v[2] = {0, 0}; uint32_t operator[](int i) const { return v[i]; } };
struct Vector3ui { uint32_t v[3] = {0, 0, 0}; uint32_t operator[](int i) const { return v[i]; } };
struct Vector2   { float v[2] = {0, 0}; };
struct Vector3   { float v[3] = {0, 0, 0}; };
struct Vector4   { float v[4] = {0, 0, 0, 0}; };
struct Quaternion { float v[4] = {1, 0, 0, 0}; };   // w, x, y, z
struct Matrix4x4 { float m[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1}; };  // row-major
struct Matrix3x4 { float m[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0}; };              // row-major, last column = translation
struct AABB3     { Vector3 min, max; };

// Exception thrown across every call (the reference's MRayError).
class MRayError : public std::runtime_error
{
    public:
    using std::runtime_error::runtime_error;
};

// ---------------------------------------------------------------------------
// TransientData: a typed, owning host buffer that is moved into the tracer
// (TracerI.h:202-204).  Created for `count` elements of T, filled by Push,
// read back by AccessAs<T>.
// ---------------------------------------------------------------------------
enum class MRayDataEnum : uint32_t
{
    MR_FLOAT, MR_VECTOR_2, MR_VECTOR_3, MR_VECTOR_4, MR_QUATERNION, MR_MATRIX_4x4,
    MR_UINT32, MR_VECTOR_2UI, MR_VECTOR_3UI, MR_BOOL, MR_STRING
};

template<class T> struct DataEnumOf;
template<> struct DataEnumOf<float>            { static constexpr auto V = MRayDataEnum::MR_FLOAT; };
template<> struct DataEnumOf<Vector2>          { static constexpr auto V = MRayDataEnum::MR_VECTOR_2; };
template<> struct DataEnumOf<Vector3>          { static constexpr auto V = MRayDataEnum::MR_VECTOR_3; };
template<> struct DataEnumOf<Vector4>          { static constexpr auto V = MRayDataEnum::MR_VECTOR_4; };
template<> struct DataEnumOf<Quaternion>       { static constexpr auto V = MRayDataEnum::MR_QUATERNION; };
template<> struct DataEnumOf<Matrix4x4>        { static constexpr auto V = MRayDataEnum::MR_MATRIX_4x4; };
template<> struct DataEnumOf<uint32_t>         { static constexpr auto V = MRayDataEnum::MR_UINT32; };
template<> struct DataEnumOf<Vector2ui>        { static constexpr auto V = MRayDataEnum::MR_VECTOR_2UI; };
template<> struct DataEnumOf<Vector3ui>        { static constexpr auto V = MRayDataEnum::MR_VECTOR_3UI; };
template<> struct DataEnumOf<bool>             { static constexpr auto V = MRayDataEnum::MR_BOOL; };
template<> struct DataEnumOf<std::string_view> { static constexpr auto V = MRayDataEnum::MR_STRING; };

class TransientData
{
    MRayDataEnum        type = MRayDataEnum::MR_FLOAT;
    size_t              elemSize = 1;
    size_t              capacity = 0;
    size_t              size = 0;
    std::vector<std::byte> bytes;

    public:
    TransientData() = default;
    // Strings are stored as `count` chars.
    template<class T>
    TransientData(std::in_place_type_t<T>, size_t count)
        : type(DataEnumOf<T>::V)
        , elemSize(std::is_same_v<T, std::string_view> ? 1 : sizeof(T))
        , capacity(count)
        , size(std::is_same_v<T, std::string_view> ? count : 0)
        , bytes(count * (std::is_same_v<T, std::string_view> ? 1 : sizeof(T)))
    {}
    TransientData(TransientData&&) noexcept = default;
    TransientData& operator=(TransientData&&) noexcept = default;
    TransientData(const TransientData&) = delete;
    TransientData& operator=(const TransientData&) = delete;

    template<class T>
    void Push(std::span<const T> data)
    {
        if(DataEnumOf<T>::V != type)
            throw MRayError("TransientData: pushed type does not match the buffer type");
        if(size + data.size() > capacity)
            throw MRayError("TransientData: push exceeds the reserved element count");
        std::memcpy(bytes.data() + size * elemSize, data.data(), data.size() * sizeof(T));
        size += data.size();
    }
    template<class T>
    std::span<const T> AccessAs() const
    {
        if(DataEnumOf<T>::V != type)
            throw MRayError("TransientData: accessed as the wrong type");
        return std::span<const T>(reinterpret_cast<const T*>(bytes.data()), size);
    }
    template<class T>
    std::span<T> AccessAs()
    {
        if(DataEnumOf<T>::V != type)
            throw MRayError("TransientData: accessed as the wrong type");
        return std::span<T>(reinterpret_cast<T*>(bytes.data()), size);
    }
    std::span<char> AccessAsString()
    {
        if(type != MRayDataEnum::MR_STRING)
            throw MRayError("TransientData: not a string");
        return std::span<char>(reinterpret_cast<char*>(bytes.data()), size);
    }
    std::string_view AccessAsString() const
    {
        if(type != MRayDataEnum::MR_STRING)
            throw MRayError("TransientData: not a string");
        return std::string_view(reinterpret_cast<const char*>(bytes.data()), size);
    }
    MRayDataEnum Type() const { return type; }
    size_t       Size() const { return size; }
    bool         IsFull() const { return size == capacity; }
};

// ---------------------------------------------------------------------------
// Ids (opaque; the reference packs group / index bit fields the same way,
// TracerTypes.h:164-178).  Group 0 of each kind is reserved for the implicit
// defaults (TracerI.h:99-124): identity transform, Null light, empty
// primitive group, passthrough material.
// ---------------------------------------------------------------------------
#define MRAY_AMD_ID(NAME) enum class NAME : uint32_t {}
MRAY_AMD_ID(PrimGroupId);   MRAY_AMD_ID(PrimBatchId);
MRAY_AMD_ID(MatGroupId);    MRAY_AMD_ID(MaterialId);
MRAY_AMD_ID(TransGroupId);  MRAY_AMD_ID(TransformId);
MRAY_AMD_ID(LightGroupId);  MRAY_AMD_ID(LightId);
MRAY_AMD_ID(CameraGroupId); MRAY_AMD_ID(CameraId);
MRAY_AMD_ID(MediumGroupId); MRAY_AMD_ID(MediumId);
MRAY_AMD_ID(TextureId);     MRAY_AMD_ID(VolumeId);
MRAY_AMD_ID(SurfaceId);     MRAY_AMD_ID(LightSurfaceId);
MRAY_AMD_ID(CamSurfaceId);  MRAY_AMD_ID(RendererId);
#undef MRAY_AMD_ID

using CommonId      = uint32_t;
using CommonIdRange = std::array<CommonId, 2>;   // [first id, last id], inclusive, as CommonId

struct PrimCount { uint32_t primCount; uint32_t attributeCount; };

using TypeNameList       = std::vector<std::string_view>;
using PrimBatchIdList    = std::vector<PrimBatchId>;
using MaterialIdList     = std::vector<MaterialId>;
using TransformIdList    = std::vector<TransformId>;
using MediumIdList       = std::vector<MediumId>;
using LightIdList        = std::vector<LightId>;
using CameraIdList       = std::vector<CameraId>;
using VolumeIdList       = std::vector<VolumeId>;
using AttributeCountList = std::vector<size_t>;

enum class AttributeOptionality : uint8_t { MR_MANDATORY, MR_OPTIONAL };
enum class AttributeTexturable  : uint8_t { MR_CONSTANT_ONLY, MR_TEXTURE_OR_CONSTANT, MR_TEXTURE_ONLY };
enum class AttributeIsArray     : uint8_t { IS_SCALAR, IS_ARRAY };
enum class AttributeIsColor     : uint8_t { IS_COLOR, IS_PURE_DATA };
enum class PrimitiveAttributeLogic : uint8_t { POSITION, INDEX, NORMAL, RADIUS, TANGENT, BITANGENT,
                                               UV0, UV1, WEIGHT, WEIGHT_INDEX };

struct GenericAttributeInfo
{
    std::string          name;
    MRayDataEnum         dataType;
    AttributeIsArray     isArray;
    AttributeOptionality isOptional;
};
struct TexturedAttributeInfo
{
    std::string          name;
    MRayDataEnum         dataType;
    AttributeIsArray     isArray;
    AttributeOptionality isOptional;
    AttributeTexturable  isTexturable;
    AttributeIsColor     isColor;
};
struct PrimAttributeInfo
{
    PrimitiveAttributeLogic logic;
    MRayDataEnum            dataType;
    AttributeIsArray        isArray;
    AttributeOptionality    isOptional;
};
using PrimAttributeInfoList     = std::vector<PrimAttributeInfo>;
using GenericAttributeInfoList  = std::vector<GenericAttributeInfo>;
using TexturedAttributeInfoList = std::vector<TexturedAttributeInfo>;
using CamAttributeInfoList      = GenericAttributeInfoList;
using TransAttributeInfoList    = GenericAttributeInfoList;
using RendererAttributeInfoList = GenericAttributeInfoList;
using MatAttributeInfoList      = TexturedAttributeInfoList;
using LightAttributeInfoList    = TexturedAttributeInfoList;
using MediumAttributeInfoList   = TexturedAttributeInfoList;

enum class AcceleratorType : uint8_t { SOFTWARE_NONE, SOFTWARE_BASIC_BVH, HARDWARE };
enum class SamplerType : uint8_t { INDEPENDENT, Z_SOBOL };
enum class WavelengthSampleMode : uint8_t { UNIFORM, GAUSSIAN_MIS, HYPERBOLIC_PBRT };
enum class MRayColorSpaceEnum : uint8_t { MR_ACES2065_1, MR_ACES_CG, MR_REC_709, MR_REC_2020,
                                          MR_DCI_P3, MR_ADOBE_RGB, MR_RAW, MR_DEFAULT };
struct FilterType
{
    enum E : uint8_t { BOX, TENT, GAUSSIAN, MITCHELL_NETRAVALI } type = GAUSSIAN;
    float radius = 1.0f;
};

// TracerI.h:45-68.  Fields the path tracer reads: seed, parallelizationHint
// (slot count), filmFilter (Gaussian radius), wavelengthSampleMode.
struct TracerParameters
{
    uint64_t             seed = 0;
    AcceleratorType      accelMode = AcceleratorType::HARDWARE;
    uint32_t             parallelizationHint = 1u << 21;
    SamplerType          samplerType = SamplerType::INDEPENDENT;
    uint32_t             clampedTexRes = UINT32_MAX;
    bool                 genMips = false;
    FilterType           mipGenFilter = FilterType{FilterType::GAUSSIAN, 2.0f};
    MRayColorSpaceEnum   globalTextureColorSpace = MRayColorSpaceEnum::MR_ACES_CG;
    FilterType           filmFilter = FilterType{FilterType::GAUSSIAN, 1.0f};
    WavelengthSampleMode wavelengthSampleMode = WavelengthSampleMode::HYPERBOLIC_PBRT;
    uint32_t             volumeTrackerEntryCount = 10000;
};

struct MRayTextureParameters { MRayColorSpaceEnum colorSpace = MRayColorSpaceEnum::MR_DEFAULT; };
struct VolumeParams { MediumId mediumId; TransformId transformId; int32_t priority; };

inline constexpr size_t MaxPrimBatchPerSurface = 8;

struct SurfaceParams
{
    std::vector<PrimBatchId>             primBatches;
    std::vector<MaterialId>              materials;
    TransformId                          transformId{0};
    std::vector<Optional<TextureId>>     alphaMaps;
    std::vector<bool>                    cullFaceFlags;
    std::vector<VolumeId>                volumes;
};
struct LightSurfaceParams
{
    LightId               lightId{0};
    TransformId           transformId{0};
    std::vector<VolumeId> nestedVolumes;
};
struct CameraSurfaceParams
{
    CameraId              cameraId{0};
    TransformId           transformId{0};
    std::vector<VolumeId> nestedVolumes;
};
struct SurfaceCommitResult
{
    AABB3  aabb;
    size_t instanceCount;
    size_t acceleratorCount;
};
struct CameraTransform
{
    Vector3 position{{0, 0, 0}};
    Vector3 gazePoint{{0, 0, -1}};
    Vector3 up{{0, 1, 0}};
};
struct RenderImageParams
{
    Vector2ui resolution;
    Vector2ui regionMin;
    Vector2ui regionMax;
};

// Common/RenderImageStructs.h:5-61, Common/AnalyticStructs.h
struct RenderBufferInfo
{
    const std::byte*   data;          // tracer-owned host buffer: R, G, B, weight planes
    size_t             totalSize;
    MRayColorSpaceEnum renderColorSpace;
    Vector2ui          resolution;
    uint32_t           curRenderLogic0 = 0;
    uint32_t           curRenderLogic1 = 0;
};
struct RenderImageSection
{
    Vector2ui             pixelMin;
    Vector2ui             pixelMax;
    float                 globalWeight;
    uint64_t              waitCounter;
    std::array<size_t, 3> pixStartOffsets;
    size_t                weightStartOffset;
};
struct RendererAnalyticData
{
    double             throughput;
    std::string        throughputSuffix;
    double             workPerPixel;
    double             wppLimit;
    std::string        workPerPixelSuffix;
    float              iterationTimeMS;
    Vector2ui          renderResolution;
    MRayColorSpaceEnum outputColorSpace;
    size_t             usedGPUMemoryBytes;
    uint32_t           customLogicSize0 = 0;
    uint32_t           customLogicSize1 = 0;
};
struct RendererOutput
{
    Optional<RendererAnalyticData> analytics;
    Optional<RenderImageSection>   imageOut;
    bool                           triggerSave = false;
};

// Host timeline semaphore shared with the image consumer (Visor, the run
// command).  The consumer owns it and hands it over in SetupRenderEnv, so it
// has the reference's member layout and protocol (Core/TimelineSemaphore.h:
// 31-110): Acquire(v) passes at once for v == 0, otherwise waits until the
// value equals v or the semaphore is invalidated (then returns false);
// Release() adds one and wakes every waiter.
class TimelineSemaphore
{
    uint64_t                value;
    std::mutex              mutex;
    std::condition_variable cVar;
    bool                    invalid;

    public:
    explicit TimelineSemaphore(uint64_t initialValue) : value(initialValue), invalid(false) {}
    [[nodiscard]] bool Acquire(uint64_t waitValue)
    {
        if(waitValue == 0) return true;
        std::unique_lock<std::mutex> lock(mutex);
        cVar.wait(lock, [&] { return value == waitValue || invalid; });
        return !invalid;
    }
    void Release()
    {
        std::unique_lock<std::mutex> lock(mutex);
        value++;
        cVar.notify_all();
    }
    void Reset()
    {
        std::unique_lock<std::mutex> lock(mutex);
        value = 0;
        cVar.notify_all();
    }
    void Invalidate()
    {
        std::unique_lock<std::mutex> lock(mutex);
        invalid = true;
        cVar.notify_all();
    }
    [[nodiscard]] bool IsInvalidated()
    {
        std::unique_lock<std::mutex> lock(mutex);
        return invalid;
    }
};
class ThreadPool;          // opaque: the host work here is single-threaded
using GPUThreadInitFunction = void (*)();

// ---------------------------------------------------------------------------
// The interface, method for method in the reference's order.
// ---------------------------------------------------------------------------
class TracerI
{
    public:
    virtual ~TracerI() = default;

    // Generic
    virtual TypeNameList PrimitiveGroups() const = 0;
    virtual TypeNameList MaterialGroups() const = 0;
    virtual TypeNameList TransformGroups() const = 0;
    virtual TypeNameList CameraGroups() const = 0;
    virtual TypeNameList MediumGroups() const = 0;
    virtual TypeNameList LightGroups() const = 0;
    virtual TypeNameList Renderers() const = 0;

    virtual PrimAttributeInfoList     AttributeInfo(PrimGroupId) const = 0;
    virtual CamAttributeInfoList      AttributeInfo(CameraGroupId) const = 0;
    virtual MediumAttributeInfoList   AttributeInfo(MediumGroupId) const = 0;
    virtual MatAttributeInfoList      AttributeInfo(MatGroupId) const = 0;
    virtual TransAttributeInfoList    AttributeInfo(TransGroupId) const = 0;
    virtual LightAttributeInfoList    AttributeInfo(LightGroupId) const = 0;
    virtual RendererAttributeInfoList AttributeInfo(RendererId) const = 0;

    virtual PrimAttributeInfoList     AttributeInfoPrim(std::string_view) const = 0;
    virtual CamAttributeInfoList      AttributeInfoCam(std::string_view) const = 0;
    virtual MediumAttributeInfoList   AttributeInfoMedium(std::string_view) const = 0;
    virtual MatAttributeInfoList      AttributeInfoMat(std::string_view) const = 0;
    virtual TransAttributeInfoList    AttributeInfoTrans(std::string_view) const = 0;
    virtual LightAttributeInfoList    AttributeInfoLight(std::string_view) const = 0;
    virtual RendererAttributeInfoList AttributeInfoRenderer(std::string_view) const = 0;

    virtual std::string TypeName(PrimGroupId) const = 0;
    virtual std::string TypeName(CameraGroupId) const = 0;
    virtual std::string TypeName(MediumGroupId) const = 0;
    virtual std::string TypeName(MatGroupId) const = 0;
    virtual std::string TypeName(TransGroupId) const = 0;
    virtual std::string TypeName(LightGroupId) const = 0;
    virtual std::string TypeName(RendererId) const = 0;

    // Primitive
    virtual PrimGroupId     CreatePrimitiveGroup(std::string typeName) = 0;
    virtual PrimBatchId     ReservePrimitiveBatch(PrimGroupId, PrimCount) = 0;
    virtual PrimBatchIdList ReservePrimitiveBatches(PrimGroupId, std::vector<PrimCount>) = 0;
    virtual void            CommitPrimReservations(PrimGroupId) = 0;
    virtual bool            IsPrimCommitted(PrimGroupId) const = 0;
    virtual void            PushPrimAttribute(PrimGroupId, PrimBatchId, uint32_t attributeIndex,
                                              TransientData data) = 0;
    virtual void            PushPrimAttribute(PrimGroupId, PrimBatchId, uint32_t attributeIndex,
                                              Vector2ui subBatchRange, TransientData data) = 0;
    virtual void            TransformPrimitives(PrimGroupId, std::vector<PrimBatchId>,
                                                std::vector<Matrix3x4>) = 0;
    // Material
    virtual MatGroupId     CreateMaterialGroup(std::string typeName) = 0;
    virtual MaterialId     ReserveMaterial(MatGroupId, AttributeCountList) = 0;
    virtual MaterialIdList ReserveMaterials(MatGroupId, std::vector<AttributeCountList>) = 0;
    virtual void           CommitMatReservations(MatGroupId) = 0;
    virtual bool           IsMatCommitted(MatGroupId) const = 0;
    virtual void           PushMatAttribute(MatGroupId, CommonIdRange range, uint32_t attributeIndex,
                                            TransientData data) = 0;
    virtual void           PushMatAttribute(MatGroupId, CommonIdRange range, uint32_t attributeIndex,
                                            TransientData data,
                                            std::vector<Optional<TextureId>>) = 0;
    virtual void           PushMatAttribute(MatGroupId, CommonIdRange range, uint32_t attributeIndex,
                                            std::vector<TextureId>) = 0;
    // Texture
    virtual TextureId CreateTexture2D(Vector2ui size, uint32_t mipCount, MRayTextureParameters) = 0;
    virtual TextureId CreateTexture3D(Vector3ui size, uint32_t mipCount, MRayTextureParameters) = 0;
    virtual void      CommitTextures() = 0;
    virtual void      PushTextureData(TextureId, uint32_t mipLevel, TransientData data) = 0;
    // Transform
    virtual TransGroupId    CreateTransformGroup(std::string typeName) = 0;
    virtual TransformId     ReserveTransformation(TransGroupId, AttributeCountList) = 0;
    virtual TransformIdList ReserveTransformations(TransGroupId, std::vector<AttributeCountList>) = 0;
    virtual void            CommitTransReservations(TransGroupId) = 0;
    virtual bool            IsTransCommitted(TransGroupId) const = 0;
    virtual void            PushTransAttribute(TransGroupId, CommonIdRange range,
                                               uint32_t attributeIndex, TransientData data) = 0;
    // Lights
    virtual LightGroupId CreateLightGroup(std::string typeName, PrimGroupId = PrimGroupId{0}) = 0;
    virtual LightId      ReserveLight(LightGroupId, AttributeCountList,
                                      PrimBatchId = PrimBatchId{0}) = 0;
    virtual LightIdList  ReserveLights(LightGroupId, std::vector<AttributeCountList>,
                                       std::vector<PrimBatchId> = std::vector<PrimBatchId>{}) = 0;
    virtual void         CommitLightReservations(LightGroupId) = 0;
    virtual bool         IsLightCommitted(LightGroupId) const = 0;
    virtual void         PushLightAttribute(LightGroupId, CommonIdRange range,
                                            uint32_t attributeIndex, TransientData data) = 0;
    virtual void         PushLightAttribute(LightGroupId, CommonIdRange range,
                                            uint32_t attributeIndex, TransientData,
                                            std::vector<Optional<TextureId>>) = 0;
    virtual void         PushLightAttribute(LightGroupId, CommonIdRange range,
                                            uint32_t attributeIndex, std::vector<TextureId>) = 0;
    // Cameras
    virtual CameraGroupId CreateCameraGroup(std::string typeName) = 0;
    virtual CameraId      ReserveCamera(CameraGroupId, AttributeCountList) = 0;
    virtual CameraIdList  ReserveCameras(CameraGroupId, std::vector<AttributeCountList>) = 0;
    virtual void          CommitCamReservations(CameraGroupId) = 0;
    virtual bool          IsCamCommitted(CameraGroupId) const = 0;
    virtual void          PushCamAttribute(CameraGroupId, CommonIdRange range,
                                           uint32_t attributeIndex, TransientData data) = 0;
    // Medium
    virtual MediumGroupId CreateMediumGroup(std::string typeName) = 0;
    virtual MediumId      ReserveMedium(MediumGroupId, AttributeCountList) = 0;
    virtual MediumIdList  ReserveMediums(MediumGroupId, std::vector<AttributeCountList>) = 0;
    virtual void          CommitMediumReservations(MediumGroupId) = 0;
    virtual bool          IsMediumCommitted(MediumGroupId) const = 0;
    virtual void          PushMediumAttribute(MediumGroupId, CommonIdRange range,
                                              uint32_t attributeIndex, TransientData data) = 0;
    virtual void          PushMediumAttribute(MediumGroupId, CommonIdRange range,
                                              uint32_t attributeIndex, TransientData,
                                              std::vector<Optional<TextureId>> textures) = 0;
    virtual void          PushMediumAttribute(MediumGroupId, CommonIdRange range,
                                              uint32_t attributeIndex,
                                              std::vector<TextureId> textures) = 0;
    // Surfaces
    virtual SurfaceId           CreateSurface(SurfaceParams) = 0;
    virtual LightSurfaceId      SetBoundarySurface(LightId, TransformId) = 0;
    virtual LightSurfaceId      CreateLightSurface(LightSurfaceParams) = 0;
    virtual CamSurfaceId        CreateCameraSurface(CameraSurfaceParams) = 0;
    virtual SurfaceCommitResult CommitSurfaces() = 0;
    virtual CameraTransform     GetCamTransform(CamSurfaceId) const = 0;
    virtual VolumeId            RegisterVolume(VolumeParams) = 0;
    virtual VolumeIdList        RegisterVolumes(std::vector<VolumeParams>) = 0;
    virtual void                SetBoundaryVolume(VolumeId) = 0;
    // Renderers
    virtual RendererId CreateRenderer(std::string typeName) = 0;
    virtual void       DestroyRenderer(RendererId) = 0;
    virtual void       PushRendererAttribute(RendererId, uint32_t attributeIndex,
                                             TransientData data) = 0;
    // Rendering
    virtual void             SetupRenderEnv(TimelineSemaphore* semaphore, uint32_t importAlignment,
                                            uint64_t initialAcquireValue) = 0;
    virtual RenderBufferInfo StartRender(RendererId, CamSurfaceId, RenderImageParams,
                                         Optional<uint32_t>, Optional<uint32_t>) = 0;
    virtual void             SetCameraTransform(RendererId, CameraTransform) = 0;
    virtual void             StopRender() = 0;
    virtual RendererOutput   DoRenderWork() = 0;
    // Misc
    virtual void                    ClearAll() = 0;
    virtual void                    Flush() const = 0;
    virtual GPUThreadInitFunction   GetThreadInitFunction() const = 0;
    virtual void                    SetThreadPool(ThreadPool&) = 0;
    virtual size_t                  TotalDeviceMemory() const = 0;
    virtual size_t                  UsedDeviceMemory() const = 0;
    virtual const TracerParameters& Parameters() const = 0;
};

// Ids of what a scene file defined (the reference's SceneLoaderI returns
// the same kind of id pack, Core/SceneLoaderI.h:49-59).
struct SceneIds
{
    std::vector<SurfaceId>      surfaces;
    std::vector<LightSurfaceId> lightSurfaces;
    std::vector<CamSurfaceId>   cameraSurfaces;
    LightSurfaceId              boundary{0};
};

// Reads an MRay JSON scene file (Docs/markdown/scene/mrayScene.md) and issues
// the TracerI calls that build it; CommitSurfaces is left to the caller.
// Throws MRayError for malformed files and for types the path tracer lacks.
SceneIds LoadSceneMRay(TracerI& tracer, const std::string& path);

}  // namespace mray_amd

// Factory pair (TracerDLL/EntryPoint.h:14-18).  The caller owns the returned
// object and must release it with DestroyTracer.
extern "C" mray_amd::TracerI* ConstructTracer(const mray_amd::TracerParameters& params);
extern "C" void DestroyTracer(mray_amd::TracerI* tracer);
