// mray_tracer.hpp -- C++20 host interface of the MI355X path tracer: the
// reference's TracerI plugin interface (Core/TracerI.h:150-376), binary
// compatible with it on the Itanium C++ ABI (Linux, libstdc++).
//
// The object returned by ConstructTracer() takes the same calls, in the same
// vtable order, with the same argument meaning and the same argument LAYOUT
// as the reference's TracerI, so the reference's front-end can load this
// library through its own factory pair (TracerDLL/EntryPoint.h:14-18) with no
// adapter:
//   * every value type that crosses a call has the reference's size,
//     alignment, member offsets and triviality (which fixes how it is
//     passed: in registers, on the stack, or through a hidden pointer);
//     tests/test_tracer_abi.py checks each against the figures derived from
//     the reference's declarations;
//   * the math value templates (Vector, Quat, Matrix, AABB) sit in the
//     global namespace under the reference's names, so typeid(T) of a pushed
//     element type names the same type on both sides: TransientData carries
//     typeid(T).hash_code() and AccessAs<T> checks it;
//   * MRayError is the reference's exception layout under its name, so the
//     front-end's catch(const MRayError&) receives what this library throws;
//   * TransientData has the reference's layout (TransientPool.h:44-53).  A
//     buffer the front-end pushes is read in place and never released here:
//     under the Itanium ABI the caller destroys a by-value argument, with
//     its own allocator.  Buffers this library's own scene loader makes use
//     this header's allocator and are released by their creator as well.
// Everything GPU-side is below the C-ABI of mray_amd.h; this layer only
// assembles the flat scene description.
//
// Supported types (everything else throws MRayError):
//   (P)Triangle                      positions, TBN quaternions, uv, indices
//   (Mt)Lambert (Mt)Reflect (Mt)Refract (Mt)Unreal   constant parameters
//   (L)Prim(P)Triangle (L)Skysphere_CoOcta (L)Skysphere_Spherical (L)Null
//   (C)Pinhole
//   (T)Identity (T)Single                            baked into positions
//   (R)PathTracerRGB (R)PathTracerSpectral
#pragma once

#include <algorithm>
#include <array>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <initializer_list>
#include <mutex>
#include <new>
#include <optional>
#include <span>
#include <stdexcept>
#include <string>
#include <string_view>
#include <type_traits>
#include <typeinfo>
#include <utility>
#include <vector>

// The library is built with hidden visibility; these are its exports.
#define MRAY_AMD_TRACER_API __attribute__((visibility("default")))

namespace mray_amd::abi
{
// Alignment of a small vector of `bytes` bytes (Core/Vector.h:19-31): 2, 4
// and 8-byte vectors align to their size, 12-byte ones to 4, 16-byte ones
// and larger to 16.
consteval unsigned int VecAlign(unsigned int bytes)
{
    return bytes <= 2 ? 2u : bytes <= 4 ? 4u : bytes <= 8 ? 8u : bytes < 16 ? 4u : 16u;
}
}  // namespace mray_amd::abi

// ---------------------------------------------------------------------------
// Math value types (global namespace, the reference's names and layouts:
// Core/Vector.h:33-43, Quaternion.h:9-15, Matrix.h:10-19,90-97, AABB.h:15-28)
// ---------------------------------------------------------------------------
template<unsigned int N, class T>
class alignas(mray_amd::abi::VecAlign(N * sizeof(T))) Vector
{
    static_assert(N >= 2 && N <= 4, "2, 3 or 4 components");
    public:
    std::array<T, N> v{};
    constexpr T&       operator[](unsigned int i)       { return v[i]; }
    constexpr const T& operator[](unsigned int i) const { return v[i]; }
};

template<class T>
class Quat
{
    public:
    Vector<4, T> v{{T(1), T(0), T(0), T(0)}};       // w, x, y, z
    constexpr T&       operator[](unsigned int i)       { return v[i]; }
    constexpr const T& operator[](unsigned int i) const { return v[i]; }
};

// Row-major N x N; the default is the identity.
template<unsigned int N, class T>
class alignas(mray_amd::abi::VecAlign(N * sizeof(T))) Matrix
{
    static constexpr std::array<T, N * N> Identity()
    {
        std::array<T, N * N> a{};
        for(unsigned int i = 0; i < N; i++) a[i * N + i] = T(1);
        return a;
    }
    public:
    std::array<T, N * N> m = Identity();
};

// Row-major 3 x 4 (last column = translation); the default is the identity.
template<class T>
class alignas(mray_amd::abi::VecAlign(4 * sizeof(T))) Matrix3x4T
{
    public:
    std::array<T, 12> m{T(1), T(0), T(0), T(0), T(0), T(1), T(0), T(0), T(0), T(0), T(1), T(0)};
};

template<unsigned int N, class T>
class AABB
{
    public:
    Vector<N, T> min;
    Vector<N, T> max;
};

// The reference's exception (Core/Error.h:10-21): thrown by value across
// every call, caught by the front-end as `const MRayError&`.  Not derived
// from std::exception, as there.
struct MRayError
{
    enum Type { OK, HAS_ERROR };

    private:
    Type        type = OK;
    std::string customInfo;

    public:
    MRayError(Type t = OK) : type(t) {}
    MRayError(std::string_view info) : type(HAS_ERROR), customInfo(info) {}
    MRayError(const std::string& info) : type(HAS_ERROR), customInfo(info) {}
    MRayError(const char* info) : type(HAS_ERROR), customInfo(info) {}
    explicit operator bool() const { return type == HAS_ERROR; }
    std::string GetError() const { return customInfo; }
    const char* what() const { return customInfo.c_str(); }
    void        AppendInfo(const std::string& s) { customInfo += s; }
};

namespace mray_amd
{

template<class T> using Optional = std::optional<T>;
using Float      = float;
using Vector2    = ::Vector<2, float>;
using Vector3    = ::Vector<3, float>;
using Vector4    = ::Vector<4, float>;
using Vector2ui  = ::Vector<2, uint32_t>;
using Vector3ui  = ::Vector<3, uint32_t>;
using Quaternion = ::Quat<float>;
using Matrix4x4  = ::Matrix<4, float>;
using Matrix3x4  = ::Matrix3x4T<float>;
using AABB3      = ::AABB<3, float>;
using MRayError  = ::MRayError;

// ---------------------------------------------------------------------------
// Fixed-capacity vector (Core/DataStructures.h:140-150): N slots of
// max(sizeof T, alignof T) bytes, then the count, aligned to at least 8.  Its
// copy and move constructors are user-provided, so it is passed by
// invisible reference like the reference's.
// ---------------------------------------------------------------------------
template<class T, size_t N>
class alignas(std::max(alignof(T), size_t(8))) StaticVector
{
    static constexpr size_t kJump = std::max(sizeof(T), alignof(T));
    alignas(T) std::byte storage[N * kJump];
    size_t count = 0;

    T*       At(size_t i)       { return std::launder(reinterpret_cast<T*>(storage + i * kJump)); }
    const T* At(size_t i) const { return std::launder(reinterpret_cast<const T*>(storage + i * kJump)); }
    void Clear()
    {
        for(size_t i = 0; i < count; i++) At(i)->~T();
        count = 0;
    }

    public:
    using value_type = T;
    StaticVector() {}
    StaticVector(std::initializer_list<T> l)
    {
        for(const T& x : l) push_back(x);
    }
    StaticVector(const StaticVector& o)
    {
        for(size_t i = 0; i < o.count; i++) push_back(o[i]);
    }
    StaticVector(StaticVector&& o)
    {
        for(size_t i = 0; i < o.count; i++) push_back(std::move(o[i]));
        o.Clear();
    }
    StaticVector& operator=(const StaticVector& o)
    {
        if(this != &o)
        {
            Clear();
            for(size_t i = 0; i < o.count; i++) push_back(o[i]);
        }
        return *this;
    }
    StaticVector& operator=(StaticVector&& o)
    {
        if(this != &o)
        {
            Clear();
            for(size_t i = 0; i < o.count; i++) push_back(std::move(o[i]));
            o.Clear();
        }
        return *this;
    }
    ~StaticVector() { Clear(); }

    void push_back(T x)
    {
        if(count == N) throw MRayError("StaticVector: capacity exceeded");
        new(storage + count * kJump) T(std::move(x));
        count++;
    }
    size_t   size() const { return count; }
    bool     empty() const { return count == 0; }
    static constexpr size_t capacity() { return N; }
    T&       operator[](size_t i)       { return *At(i); }
    const T& operator[](size_t i) const { return *At(i); }
    T*       begin()       { return At(0); }
    T*       end()         { return At(0) + count; }
    const T* begin() const { return At(0); }
    const T* end() const   { return At(0) + count; }
};
static_assert(sizeof(int) == 4 && sizeof(size_t) == 8, "LP64 layouts");

// ---------------------------------------------------------------------------
// Data type tags (Core/Definitions.h:316-382, MRayDataType.h:286-300):
// the enumerators in the reference's order, as the attribute lists report
// them to the front-end.
// ---------------------------------------------------------------------------
enum class MRayDataEnum : uint16_t
{
    MR_INT8, MR_VECTOR_2C, MR_VECTOR_3C, MR_VECTOR_4C,
    MR_INT16, MR_VECTOR_2S, MR_VECTOR_3S, MR_VECTOR_4S,
    MR_INT32, MR_VECTOR_2I, MR_VECTOR_3I, MR_VECTOR_4I,
    MR_INT64, MR_VECTOR_2L, MR_VECTOR_3L, MR_VECTOR_4L,
    MR_UINT8, MR_VECTOR_2UC, MR_VECTOR_3UC, MR_VECTOR_4UC,
    MR_UINT16, MR_VECTOR_2US, MR_VECTOR_3US, MR_VECTOR_4US,
    MR_UINT32, MR_VECTOR_2UI, MR_VECTOR_3UI, MR_VECTOR_4UI,
    MR_UINT64, MR_VECTOR_2UL, MR_VECTOR_3UL, MR_VECTOR_4UL,
    MR_FLOAT, MR_VECTOR_2, MR_VECTOR_3, MR_VECTOR_4,
    MR_QUATERNION, MR_MATRIX_4x4, MR_MATRIX_3x3, MR_AABB3, MR_RAY,
    MR_UNORM_4x8, MR_UNORM_2x16, MR_SNORM_4x8, MR_SNORM_2x16,
    MR_STRING, MR_BOOL,
    MR_ENUM_END
};
struct MRayDataTypeRT
{
    MRayDataEnum name = MRayDataEnum::MR_ENUM_END;
    constexpr MRayDataTypeRT() = default;
    constexpr MRayDataTypeRT(MRayDataEnum e) : name(e) {}
    constexpr MRayDataEnum Name() const { return name; }
    constexpr bool operator==(MRayDataTypeRT o) const { return name == o.name; }
};

enum class MRayPixelEnum : uint16_t
{
    MR_R8_UNORM, MR_RG8_UNORM, MR_RGB8_UNORM, MR_RGBA8_UNORM,
    MR_R16_UNORM, MR_RG16_UNORM, MR_RGB16_UNORM, MR_RGBA16_UNORM,
    MR_R8_SNORM, MR_RG8_SNORM, MR_RGB8_SNORM, MR_RGBA8_SNORM,
    MR_R16_SNORM, MR_RG16_SNORM, MR_RGB16_SNORM, MR_RGBA16_SNORM,
    MR_R_HALF, MR_RG_HALF, MR_RGB_HALF, MR_RGBA_HALF,
    MR_R_FLOAT, MR_RG_FLOAT, MR_RGB_FLOAT, MR_RGBA_FLOAT,
    MR_BC1_UNORM, MR_BC2_UNORM, MR_BC3_UNORM, MR_BC4_UNORM, MR_BC4_SNORM,
    MR_BC5_UNORM, MR_BC5_SNORM, MR_BC6H_UFLOAT, MR_BC6H_SFLOAT, MR_BC7_UNORM,
    MR_ENUM_END
};
struct MRayPixelTypeRT
{
    MRayPixelEnum name = MRayPixelEnum::MR_ENUM_END;
    constexpr MRayPixelTypeRT() = default;
    constexpr MRayPixelTypeRT(MRayPixelEnum e) : name(e) {}
    constexpr MRayPixelEnum Name() const { return name; }
};

// Element type of each tag this layer reads (a subset of MRayDataType.h:31-92).
template<class T> struct DataEnumOf;
template<> struct DataEnumOf<float>      { static constexpr auto V = MRayDataEnum::MR_FLOAT; };
template<> struct DataEnumOf<Vector2>    { static constexpr auto V = MRayDataEnum::MR_VECTOR_2; };
template<> struct DataEnumOf<Vector3>    { static constexpr auto V = MRayDataEnum::MR_VECTOR_3; };
template<> struct DataEnumOf<Vector4>    { static constexpr auto V = MRayDataEnum::MR_VECTOR_4; };
template<> struct DataEnumOf<Quaternion> { static constexpr auto V = MRayDataEnum::MR_QUATERNION; };
template<> struct DataEnumOf<Matrix4x4>  { static constexpr auto V = MRayDataEnum::MR_MATRIX_4x4; };
template<> struct DataEnumOf<uint32_t>   { static constexpr auto V = MRayDataEnum::MR_UINT32; };
template<> struct DataEnumOf<Vector2ui>  { static constexpr auto V = MRayDataEnum::MR_VECTOR_2UI; };
template<> struct DataEnumOf<Vector3ui>  { static constexpr auto V = MRayDataEnum::MR_VECTOR_3UI; };
template<> struct DataEnumOf<bool>       { static constexpr auto V = MRayDataEnum::MR_BOOL; };

// ---------------------------------------------------------------------------
// TransientData (TransientPool/TransientPool.h:44-53, TransientPool.hpp):
// { typeid(T).hash_code(), owned bytes as {pointer, u32 byte count} aligned
// to 16, bytes pushed, alignment }.  Strings are tagged with
// typeid(std::string) and hold their characters.
// ---------------------------------------------------------------------------
class TransientData
{
    struct alignas(16) Bytes
    {
        std::byte* ptr = nullptr;
        uint32_t   count = 0;
    };
    size_t typeHash = 0;
    Bytes  ownedMem;
    size_t usedBytes = 0;
    size_t alignment = 1;

    void Allocate(size_t bytes, size_t align)
    {
        if(bytes > UINT32_MAX) throw MRayError("TransientData: more than 4 GiB in one buffer");
        alignment = align;
        if(bytes == 0) return;
        ownedMem.ptr = static_cast<std::byte*>(::operator new(bytes, std::align_val_t(align)));
        ownedMem.count = uint32_t(bytes);
    }
    void Release()
    {
        if(ownedMem.ptr) ::operator delete(ownedMem.ptr, std::align_val_t(alignment));
        ownedMem = Bytes{};
    }
    template<class T>
    static size_t HashOf() { return typeid(std::remove_const_t<T>).hash_code(); }

    public:
    template<class T>
    TransientData(std::in_place_type_t<T>, size_t count)
    {
        if constexpr(std::is_same_v<T, std::string> || std::is_same_v<T, std::string_view>)
        {
            typeHash = HashOf<std::string>();
            Allocate(count, 1);
        }
        else
        {
            static_assert(std::is_trivially_copyable_v<T>);
            typeHash = HashOf<T>();
            Allocate(count * sizeof(T), alignof(T));
        }
    }
    TransientData(TransientData&& o) noexcept
        : typeHash(o.typeHash), ownedMem(o.ownedMem), usedBytes(o.usedBytes), alignment(o.alignment)
    {
        o.ownedMem = Bytes{};
    }
    TransientData& operator=(TransientData&& o) noexcept
    {
        if(this != &o)
        {
            Release();
            typeHash = o.typeHash;
            ownedMem = o.ownedMem;
            usedBytes = o.usedBytes;
            alignment = o.alignment;
            o.ownedMem = Bytes{};
        }
        return *this;
    }
    TransientData(const TransientData&) = delete;
    TransientData& operator=(const TransientData&) = delete;
    ~TransientData() { Release(); }

    template<class T>
    void Push(std::span<const T> data)
    {
        if(HashOf<T>() != typeHash) throw MRayError("TransientData: pushed type does not match the buffer type");
        size_t n = data.size_bytes();
        if(usedBytes + n > ownedMem.count)
            throw MRayError("TransientData: push exceeds the reserved element count");
        if(n) std::memcpy(ownedMem.ptr + usedBytes, data.data(), n);
        usedBytes += n;
    }
    void PushString(std::string_view s)
    {
        if(typeHash != HashOf<std::string>()) throw MRayError("TransientData: not a string buffer");
        if(usedBytes + s.size() > ownedMem.count)
            throw MRayError("TransientData: push exceeds the reserved element count");
        if(!s.empty()) std::memcpy(ownedMem.ptr + usedBytes, s.data(), s.size());
        usedBytes += s.size();
    }
    // The pushed elements, read in place (the reference checks the type the
    // same way, TransientPool.hpp AccessAs).
    template<class T>
    std::span<const T> AccessAs() const
    {
        if(HashOf<T>() != typeHash) throw MRayError("TransientData: accessed as the wrong type");
        return std::span<const T>(reinterpret_cast<const T*>(ownedMem.ptr), usedBytes / sizeof(T));
    }
    std::string_view AccessAsString() const
    {
        if(typeHash != HashOf<std::string>() && typeHash != HashOf<std::string_view>())
            throw MRayError("TransientData: not a string");
        return std::string_view(reinterpret_cast<const char*>(ownedMem.ptr), usedBytes);
    }
    size_t ByteSize() const { return usedBytes; }
    bool   IsFull() const { return usedBytes == ownedMem.count; }
};

// ---------------------------------------------------------------------------
// Ids (Core/TracerAttribInfo.h:44-98; u32 enums, MRay::CommonKey).  Group 0
// of each kind is reserved for the implicit defaults (TracerI.h:99-124):
// identity transform, Null light, empty primitive group, passthrough
// material.
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
using CommonIdRange = ::Vector<2, CommonId>;   // [first id, last id], inclusive

struct PrimCount { uint32_t primCount; uint32_t attributeCount; };

// TracerConstants.h:9-37
inline constexpr size_t MaxPrimBatchPerSurface    = 8;
inline constexpr size_t MaxAttributePerGroup      = 16;
inline constexpr size_t MaxRendererAttributeCount = 32;
inline constexpr size_t MaxNestedVolumes          = 8;

using TypeNameList       = std::vector<std::string_view>;
using PrimBatchIdList    = std::vector<PrimBatchId>;
using MaterialIdList     = std::vector<MaterialId>;
using TransformIdList    = std::vector<TransformId>;
using MediumIdList       = std::vector<MediumId>;
using LightIdList        = std::vector<LightId>;
using CameraIdList       = std::vector<CameraId>;
using VolumeIdList       = std::vector<VolumeId>;
using AttributeCountList = StaticVector<size_t, MaxAttributePerGroup>;

// Core/TracerEnums.h:195-218
enum class AttributeOptionality : uint8_t { MR_MANDATORY, MR_OPTIONAL };
enum class AttributeTexturable  : uint8_t { MR_CONSTANT_ONLY, MR_TEXTURE_OR_CONSTANT, MR_TEXTURE_ONLY };
enum class AttributeIsColor     : uint8_t { IS_COLOR, IS_PURE_DATA };
enum class AttributeIsArray     : uint8_t { IS_SCALAR, IS_ARRAY };

// Core/TracerEnums.h: small classes around an unscoped enum (4 bytes)
struct PrimitiveAttributeLogic
{
    enum E : unsigned int { POSITION, INDEX, NORMAL, RADIUS, TANGENT, BITANGENT, UV0, UV1,
                            WEIGHT, WEIGHT_INDEX, MR_ENUM_END };
    E e;
    constexpr PrimitiveAttributeLogic(E x) : e(x) {}
};
struct AcceleratorType
{
    enum E : unsigned int { SOFTWARE_NONE, SOFTWARE_BASIC_BVH, HARDWARE, MR_ENUM_END };
    E e;
    constexpr AcceleratorType(E x) : e(x) {}
};
struct SamplerType
{
    enum E : unsigned int { INDEPENDENT, Z_SOBOL, SOBOL, MR_ENUM_END };
    E e;
    constexpr SamplerType(E x) : e(x) {}
};
struct WavelengthSampleMode
{
    enum E : unsigned int { UNIFORM, GAUSSIAN_MIS, HYPERBOLIC_PBRT, MR_ENUM_END };
    E e;
    constexpr WavelengthSampleMode(E x) : e(x) {}
};
struct FilterType
{
    enum E : unsigned int { BOX, TENT, GAUSSIAN, MITCHELL_NETRAVALI, MR_ENUM_END };
    E     type;
    Float radius;
};

// Core/Definitions.h:438-506
enum class MRayColorSpaceEnum : uint8_t { MR_ACES2065_1, MR_ACES_CG, MR_REC_709, MR_REC_2020,
                                          MR_DCI_P3, MR_ADOBE_RGB, MR_DEFAULT, MR_ENUM_END };
enum class MRayTextureInterpEnum : uint8_t { MR_NEAREST, MR_LINEAR, MR_ENUM_END };
enum class MRayTextureEdgeResolveEnum : uint8_t { MR_WRAP, MR_CLAMP, MR_MIRROR, MR_ENUM_END };
enum class MRayTextureReadMode : uint8_t { MR_PASSTHROUGH, MR_DROP_1, MR_DROP_2, MR_DROP_3,
                                           MR_AS_3C_TS_NORMAL_BASIC, MR_AS_3C_TS_NORMAL_COOCTA,
                                           MR_ENUM_END };
enum class MRayTextureIsIlluminant : uint8_t { IS_ALBEDO, IS_ILLUMINANT };

// Core/TracerAttribInfo.h:16-57
struct GenericAttributeInfo
{
    std::string          name;
    MRayDataTypeRT       dataType;
    AttributeIsArray     isArray;
    AttributeOptionality isOptional;
};
struct TexturedAttributeInfo
{
    std::string          name;
    MRayDataTypeRT       dataType;
    AttributeIsArray     isArray;
    AttributeOptionality isOptional;
    AttributeTexturable  isTexturable;
    AttributeIsColor     isColor;
};
struct PrimAttributeInfo
{
    PrimitiveAttributeLogic logic;
    MRayDataTypeRT          dataType;
    AttributeIsArray        isArray;
    AttributeOptionality    isOptional;
};
using PrimAttributeInfoList     = StaticVector<PrimAttributeInfo, MaxAttributePerGroup>;
using GenericAttributeInfoList  = StaticVector<GenericAttributeInfo, MaxAttributePerGroup>;
using TexturedAttributeInfoList = StaticVector<TexturedAttributeInfo, MaxAttributePerGroup>;
using CamAttributeInfoList      = GenericAttributeInfoList;
using TransAttributeInfoList    = GenericAttributeInfoList;
using RendererAttributeInfoList = StaticVector<GenericAttributeInfo, MaxRendererAttributeCount>;
using MatAttributeInfoList      = TexturedAttributeInfoList;
using LightAttributeInfoList    = TexturedAttributeInfoList;
using MediumAttributeInfoList   = TexturedAttributeInfoList;

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

// TracerI.h:62-73
struct MRayTextureParameters
{
    MRayPixelTypeRT            pixelType;
    MRayColorSpaceEnum         colorSpace     = MRayColorSpaceEnum::MR_DEFAULT;
    Float                      gamma          = 1.0f;
    bool                       ignoreResClamp = false;
    AttributeIsColor           isColor        = AttributeIsColor::IS_COLOR;
    MRayTextureEdgeResolveEnum edgeResolve    = MRayTextureEdgeResolveEnum::MR_WRAP;
    MRayTextureInterpEnum      interpolation  = MRayTextureInterpEnum::MR_LINEAR;
    MRayTextureReadMode        readMode       = MRayTextureReadMode::MR_PASSTHROUGH;
    MRayTextureIsIlluminant    isIlluminant   = MRayTextureIsIlluminant::IS_ALBEDO;
};

// Core/TracerAttribInfo.h:85-92
struct VolumeParams
{
    MediumId    mediumId;
    TransformId transformId;
    int32_t     priority;
};

using SurfaceMatList       = StaticVector<MaterialId, MaxPrimBatchPerSurface>;
using SurfacePrimList      = StaticVector<PrimBatchId, MaxPrimBatchPerSurface>;
using OptionalAlphaMapList = StaticVector<Optional<TextureId>, MaxPrimBatchPerSurface>;
using CullBackfaceFlagList = StaticVector<bool, MaxPrimBatchPerSurface>;
using SurfaceVolumeList    = StaticVector<VolumeId, MaxPrimBatchPerSurface>;
using BoundaryVolumeList   = StaticVector<VolumeId, MaxNestedVolumes>;

// TracerI.h:123-146
struct SurfaceParams
{
    SurfacePrimList      primBatches;
    SurfaceMatList       materials;
    TransformId          transformId{0};
    OptionalAlphaMapList alphaMaps;
    CullBackfaceFlagList cullFaceFlags;
    SurfaceVolumeList    volumes;
};
struct LightSurfaceParams
{
    LightId            lightId{0};
    TransformId        transformId{0};
    BoundaryVolumeList nestedVolumes;
};
struct CameraSurfaceParams
{
    CameraId           cameraId{0};
    TransformId        transformId{0};
    BoundaryVolumeList nestedVolumes;
};
struct SurfaceCommitResult
{
    AABB3  aabb;
    size_t instanceCount;
    size_t acceleratorCount;
};
// Common/AnalyticStructs.h:11-16
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

// Common/RenderImageStructs.h:5-56, Common/AnalyticStructs.h:33-52
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
    Float                 globalWeight;
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
MRAY_AMD_TRACER_API SceneIds LoadSceneMRay(TracerI& tracer, const std::string& path);

}  // namespace mray_amd

// Factory pair (TracerDLL/EntryPoint.h:14-18).  The caller owns the returned
// object and must release it with DestroyTracer.
extern "C" MRAY_AMD_TRACER_API mray_amd::TracerI* ConstructTracer(const mray_amd::TracerParameters& params);
extern "C" MRAY_AMD_TRACER_API void DestroyTracer(mray_amd::TracerI* tracer);
