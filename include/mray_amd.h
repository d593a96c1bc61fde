/*
 * mray_amd.h -- C-ABI of the MI355X (gfx950) wavefront path-tracing hot path.
 *
 * Every entry point takes plain pointers and sizes.  "d" prefixed pointers are
 * HIP device pointers, "h" prefixed ones are host pointers, "stream" is a
 * hipStream_t passed as void* (NULL = the null stream).  Every call returns
 * MRAY_OK (0) or a non-zero MRayStatus; mray_last_error() gives the message.
 *
 * Reference interfaces replaced (yalcinerbora/mray, paths relative to Source/):
 *   scene / accelerator construction  TracerI::CommitSurfaces            Core/TracerI.h:272-280
 *                                     AcceleratorGroupLBVH::Construct     Tracer/AcceleratorLBVH.hpp:433-582
 *   closest / any hit casting         BaseAcceleratorI::CastRays / CastVisibilityRays
 *                                                                         Tracer/AcceleratorC.h:375-430
 *   renderer start / one bounce       TracerI::StartRender / DoRenderWork Core/TracerI.h:345-363
 *                                     PathTracerRendererT::StartRender    TracerDLL/PathTracerRenderer.cu:1182-1355
 *                                     PathTracerRendererT::DoThroughputSingleTileRender
 *                                                                         TracerDLL/PathTracerRenderer.cu:1000-1074
 *   film read-back                    RenderImage::TransferToHost         Tracer/RenderImage.cpp:163
 */
#ifndef MRAY_AMD_H
#define MRAY_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum MRayStatus
{
    MRAY_OK              = 0,
    MRAY_ERR_INVALID     = 1,   /* bad argument (the reference throws MRayError) */
    MRAY_ERR_HIP         = 2,   /* a HIP runtime call failed                     */
    MRAY_ERR_OOM         = 3,   /* device allocation failed                      */
    MRAY_ERR_UNSUPPORTED = 4    /* valid in the reference, not built here        */
} MRayStatus;

/* ---------------------------------------------------------------------- */
/* Scene description (host memory).  Mirrors the reference's default types: */
/* Triangle primitive group, Lambert / Reflect / Refract / Unreal materials, */
/* PrimLight<Triangle> area lights, Null or Skysphere boundary light and a   */
/* Pinhole camera (Tracer/TracerDLL/TracerTypeGenInput.txt).                  */
/* ---------------------------------------------------------------------- */
enum
{
    MRAY_MAT_LAMBERT = 0,   /* MaterialsDefault.hpp:6-131   */
    MRAY_MAT_REFLECT = 1,   /* MaterialsDefault.hpp:133-230 */
    MRAY_MAT_REFRACT = 2,   /* MaterialsDefault.hpp:232-473 */
    MRAY_MAT_UNREAL  = 3    /* MaterialsDefault.hpp:475-784 */
};

enum
{
    MRAY_BOUNDARY_NULL             = 0,   /* LightNull                       */
    MRAY_BOUNDARY_SKYSPHERE_COOCTA = 1,   /* LightSkysphere<CoOcta>          */
    MRAY_BOUNDARY_SKYSPHERE_SPHERICAL = 2 /* LightSkysphere<Spherical>: texture
                                             uv latitude-longitude             */
};

/* Surface key of a triangle: a material index, or a light index with this flag. */
#define MRAY_LIGHT_FLAG 0x80000000u

typedef struct MRayMaterial
{
    uint32_t type;
    float    albedo[3];
    float    roughness;
    float    metallic;
    float    specular;
    float    cauchyFront[3];
    float    cauchyBack[3];
} MRayMaterial;

typedef struct MRayCamera
{
    float position[3];
    float gaze[3];       /* look-at point                         */
    float up[3];
    float fov[2];        /* (fovX, fovY), radians                 */
    float nearFar[2];
} MRayCamera;

typedef struct MRaySceneDesc
{
    uint32_t            vertexCount;
    uint32_t            triangleCount;
    const float*        hPositions;     /* [vertexCount][3]                        */
    const float*        hTbnQuats;      /* [vertexCount][4] (w,x,y,z) world->tangent */
    const uint32_t*     hIndices;       /* [triangleCount][3]                      */
    const uint32_t*     hSurfaceKeys;   /* [triangleCount] material or light key    */
    const uint8_t*      hCullBackFace;  /* [triangleCount]                          */
    uint32_t            materialCount;
    const MRayMaterial* hMaterials;
    uint32_t            lightCount;     /* emissive triangles (boundary excluded)   */
    const uint32_t*     hLightTriangles;/* [lightCount] triangle of light i         */
    const float*        hLightRadiance; /* [lightCount][3] linear RGB               */
    const uint8_t*      hLightTwoSided; /* [lightCount]                             */
    uint32_t            boundaryType;
    float               boundaryRadiance[3];
    float               sceneDiameter;
    MRayCamera          camera;
} MRaySceneDesc;

enum
{
    MRAY_SAMPLE_PURE    = 0,  /* "Pure"                    */
    MRAY_SAMPLE_NEE     = 1,  /* "WithNextEventEstimation" */
    MRAY_SAMPLE_NEE_MIS = 2   /* "WithNEEAndMIS"           */
};

typedef struct MRayRenderParams
{
    uint32_t width;
    uint32_t height;
    uint32_t totalSPP;
    uint32_t rrRange[2];      /* russianRouletteRange (start, max depth)  */
    uint32_t sampleMode;      /* MRAY_SAMPLE_*                            */
    uint32_t spectral;        /* 0: PathTracerRGB, 1: PathTracerSpectral   */
    uint32_t maxRayCount;     /* path slots; 0 = width*height             */
    uint64_t seed;
    float    filterRadius;    /* Gaussian film filter radius (default 1)  */
    uint32_t flags;           /* MRAY_RENDER_* schedule options           */
} MRayRenderParams;

/* Material-keyed queues: classify hits (boundary/light + 4 material types)
 * and shade each queue with a specialised kernel, instead of one shading
 * kernel over all hits.  Same image either way; off by default because the
 * per-queue atomics and launches cost more than the divergence they remove
 * on the measured scenes.                                                  */
#define MRAY_RENDER_MATERIAL_QUEUES 1u
/* Launch every bounce directly instead of replaying a captured hipGraph of
 * 8 bounces (mray_renderer_iterate on the renderer's own stream).          */
#define MRAY_RENDER_NO_GRAPH 2u
/* NEE picks triangle lights in proportion to emitted power (luminance x
 * area, x2 if two-sided) through an alias table, instead of uniformly
 * (LightSampler.hpp:30-119); the boundary light keeps the uniform
 * sampler's probability 1 / (lights + 1).  This build's own sampler.     */
#define MRAY_RENDER_LIGHT_POWER 4u
/* Closest-hit traversal with lane refill on 4-wide trees of one scene or a
 * batch: a lane whose walk has ended takes the next ray of its wave's chunk
 * once 16 of the wave's 64 lanes are idle (Aila and Laine 2009), instead of
 * idling until the wave's longest walk ends.  Same hits either way.        */
#define MRAY_RENDER_LANE_REFILL 8u
/* Tiled rendering (the reference's ImageTiler, Tracer/RenderImage.cpp:20-136,
 * and DoLatencyRender, TracerDLL/PathTracerRenderer.cu:1076-1180): the image
 * is split into tiles of about maxRayCount pixels (0: 2^21) as
 * ImageTiler::FindOptimumTileSize does, the renderer holds one path slot per
 * pixel of a tile, and mray_renderer_render_tile renders a tile at a time.
 * Every pixel keeps its own random state for the whole render (on the
 * device; the reference caches them on the host, Random.cu:728-774).  One
 * scene only (mray_renderer_create); no radiance cache.                    */
#define MRAY_RENDER_TILED 16u

/* Per-iteration analytics (Common/AnalyticStructs.h:33-52). */
typedef struct MRayRenderStats
{
    uint64_t deadPaths;       /* paths completed in this call            */
    uint64_t totalDeadPaths;  /* paths completed since start            */
    uint64_t extensionRays;   /* closest-hit rays cast in this call      */
    uint64_t shadowRays;      /* visibility rays cast in this call       */
    uint32_t iterations;      /* bounces run in this call                */
    uint32_t finished;        /* all totalSPP samples are in the film    */
    uint64_t nodeVisits;      /* BVH nodes visited (counting mode)       */
    uint64_t triTests;        /* triangles tested (counting mode)        */
    uint32_t hostWaits;       /* render_all / render_frame: host waits on
                                 a burst's scheduler snapshot while later
                                 bursts stay queued (the GPU keeps working) */
    uint32_t deviceDrains;    /* waits that leave the GPU idle: 1, at the end */
    uint64_t shadowNodeVisits;  /* BVH nodes visited by any-hit walks (counting mode) */
    uint64_t shadowTriTests;    /* triangles tested by any-hit walks (counting mode) */
} MRayRenderStats;

/* ---------------------------------------------------------------------- */
/* Library                                                                 */
/* ---------------------------------------------------------------------- */
const char* mray_last_error(void);
const char* mray_version(void);
int         mray_device_count(int* hCount);
/* Device memory of the current GPU (TracerI::TotalDeviceMemory /
 * UsedDeviceMemory, Core/TracerI.h:370-371).                              */
int         mray_device_memory(size_t* hTotal, size_t* hUsed);

/* ---------------------------------------------------------------------- */
/* LBVH build stages (device pointers).  K1/K2/K4/K5/K6/K7 of the reference */
/* ---------------------------------------------------------------------- */
/* KCGenPrimCenters + KCGeneratePrimAABBs  (AcceleratorWork.kt.h:5-123).
 * dCenters [n][3], dAabbs [n][6] (min xyz, max xyz).                      */
int mray_prim_centers_aabbs(const float* dPositions, const uint32_t* dIndices,
                            uint32_t triangleCount, float* dCenters, float* dAabbs,
                            void* stream);
/* Segmented transform-reduce of the AABBs into one scene box
 * (AcceleratorLBVH.hpp:764-769).  dOut [6].                               */
int mray_reduce_aabb(const float* dAabbs, uint32_t count, float* dOut, void* stream);
/* KCGenMortonCode (AcceleratorLBVH.cu:96-168): 63-bit codes, bit-exact.   */
int mray_morton63(const float* dCenters, uint32_t count, const float* dSceneAabb,
                  uint64_t* dCodes, void* stream);
/* 30-bit codes, Compose3D<uint32_t> (GraphicsFunctions.h:585-603).         */
int mray_morton30(const float* dCenters, uint32_t count, const float* dSceneAabb,
                  uint32_t* dCodes, void* stream);
/* Stable ascending sort of (code, index) pairs over all 64 key bits
 * (AlgRadixSortCUDA.h, used at AcceleratorLBVH.hpp:828-838).  dTemp needs
 * mray_radix_sort_temp_bytes(count) bytes.                                 */
size_t mray_radix_sort_temp_bytes(uint32_t count);
int mray_radix_sort_u64_u32(uint64_t* dKeys, uint32_t* dValues, uint32_t count,
                            void* dTemp, size_t tempBytes, void* stream);
/* KCConstructLBVHInternalNodes + KCUnionLBVHBoundingBoxes
 * (AcceleratorLBVH.cu:170-435).  Sorted codes and leaf order in; dNodes
 * [count-1][3] {left, right, parent} with bit 31 = leaf, dNodeAabbs
 * [count-1][6].  A single leaf produces one node {leaf 0, invalid, invalid}. */
int mray_lbvh_build(const uint64_t* dSortedCodes, const uint32_t* dLeafOrder,
                    const float* dLeafAabbs, uint32_t count, uint32_t* dNodes,
                    float* dNodeAabbs, void* dTemp, size_t tempBytes, void* stream);
size_t mray_lbvh_temp_bytes(uint32_t count);
/* The radix sort above with no workspace argument and no allocation.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Its scratch (12.3 MiB, sized for
 * MRAY_NOALLOC_SORT_MAX pairs) is a static device array, one copy per GPU,
 * held until the process ends; larger counts return MRAY_ERR_INVALID.
 * Every call on one GPU shares that GPU's scratch: calls that overlap on one
 * GPU (two streams, two threads) race and must be serialised by the caller. */
#define MRAY_NOALLOC_SORT_MAX (1u << 20)
int mray_radix_sort_u64_u32_noalloc(uint64_t* dKeys, uint32_t* dValues, uint32_t count,
                                    void* stream);

/* ---------------------------------------------------------------------- */
/* Accelerator: scene upload, LBVH build, wide-BVH collapse, casting.       */
/* ---------------------------------------------------------------------- */
typedef struct MRayAccel MRayAccel;

int mray_accel_create(const MRaySceneDesc* hScene, MRayAccel** out);

/* Traversal-tree builder.  Both build the reference's LBVH (K1-K7) on the
 * GPU -- its products are what mray_accel_read_lbvh returns -- and differ in
 * the binary tree collapsed into the 4-wide traversal BVH: SAH (default)
 * uses a binned surface-area-heuristic tree over the same triangle boxes,
 * LBVH collapses the Karras tree itself.  Hits are identical either way.   */
enum { MRAY_BUILDER_SAH = 0, MRAY_BUILDER_LBVH = 1 };
/* The LBVH builder with 4-wide nodes collapses the Karras tree and gathers
 * the triangle soup on the GPU (the whole build stays in HBM); the SAH tree
 * and 8-wide nodes are collapsed on the host.  This flag collapses the LBVH
 * on the host as well (the same tree, its nodes numbered depth first rather
 * than level by level).                                                    */
#define MRAY_ACCEL_HOST_COLLAPSE 1u
typedef struct MRayAccelOptions
{
    uint32_t builder;
    uint32_t width;            /* 4: fp32 4-wide nodes; 8: compressed 8-wide
                                  nodes (8-bit quantised boxes); 0 = 4     */
    uint32_t flags;            /* MRAY_ACCEL_*                              */
    uint32_t reserved;
} MRayAccelOptions;
int mray_accel_create_ex(const MRaySceneDesc* hScene, const MRayAccelOptions* hOptions,
                         MRayAccel** out);
int mray_accel_destroy(MRayAccel* accel);
/* Sizes of the built structure: binary node count, wide node count, leaf
 * triangle count.                                                         */
int mray_accel_info(const MRayAccel* accel, uint32_t* hBinaryNodes,
                    uint32_t* hWideNodes, uint32_t* hLeafTriangles);
/* Copy out the reference-layout products of the build (host buffers):
 * Morton codes in triangle order [n], leaf order [n], binary nodes
 * [n-1][3] and node boxes [n-1][6].  Any pointer may be NULL.             */
int mray_accel_read_lbvh(const MRayAccel* accel, uint64_t* hCodes, uint32_t* hLeafOrder,
                         uint32_t* hNodes, float* hNodeAabbs);
/* Copy out the 4-wide traversal tree (4-wide accelerators): hNodes
 * [wideNodes][32] words per node = child boxes lox[4] hix[4] loy[4] hiy[4]
 * loz[4] hiz[4] (floats), child[4] (node index, or 0x80000000 |
 * (count-1) << 26 | first soup slot; 0xFFFFFFFF empty), 4 pad words; and
 * hSoupPrims [triangleCount], the triangle of each soup slot.  Either
 * pointer may be NULL.                                                      */
int mray_accel_read_wide(const MRayAccel* accel, uint32_t* hNodes, uint32_t* hSoupPrims);
/* Closest hit (BaseAcceleratorI::CastRays).  dRays [n][8] = RayGMem
 * {pos xyz, tMin, dir xyz, tMax}.  dPrim [n] triangle index or 0xFFFFFFFF;
 * dT [n]; dBary [n][2] = MetaHit (a, b).  Ties in t resolve to the lower
 * triangle index.                                                          */
int mray_accel_cast_closest(const MRayAccel* accel, const float* dRays, uint32_t count,
                            uint32_t* dPrim, float* dT, float* dBary, void* stream);
/* Same cast, also adding {nodes visited, triangles tested} to dCounters[2]
 * (u64); the counts give the kernel's algorithmic bytes per ray.  No
 * reference counterpart: it instruments CastRays for the roofline.         */
int mray_accel_cast_closest_counted(const MRayAccel* accel, const float* dRays, uint32_t count,
                                    uint32_t* dPrim, float* dT, float* dBary,
                                    unsigned long long* dCounters, void* stream);
/* Any hit (CastVisibilityRays): dVisible [n] = 1 when nothing is hit.       */
int mray_accel_cast_visibility(const MRayAccel* accel, const float* dRays, uint32_t count,
                               uint32_t* dVisible, void* stream);

/* Textured Skysphere (LightGroupSkysphere with a radiance texture,
 * LightsDefault.hpp:300-421, 860-907): hRgb [height][width][3] linear RGB
 * texels over concentric-octahedral uv (texel (x, y) centred at
 * ((x + 1/2) / width, (y + 1/2) / height)), read bilinearly with wrapped
 * edges; NEE samples directions by texel luminance (PWC-2D,
 * Distributions.h:117-243).  The scene's boundary must be
 * a Skysphere, MRAY_BOUNDARY_SKYSPHERE_COOCTA (texels over concentric-
 * octahedral uv) or _SPHERICAL (latitude-longitude uv); boundaryRadiance is
 * then unused;
 * renderers and merged scenes created afterwards use the texture.
 * Sizes 1..8192 per side; texels must be finite, luminance not all zero.
 * A second call replaces the texture for renderers created afterwards; the
 * earlier tables stay allocated until the accelerator is destroyed.       */
int mray_accel_set_environment(MRayAccel* accel, uint32_t width, uint32_t height,
                               const float* hRgb);
/* The luminance distribution: hCdfX [height][width] normalized row CDFs,
 * hCdfY [height] the row-marginal CDF.                                     */
int mray_accel_environment_cdfs(const MRayAccel* accel, float* hCdfX, float* hCdfY);

/* ---------------------------------------------------------------------- */
/* Path tracer renderer ("PathTracerRGB" / "PathTracerSpectral").           */
/* ---------------------------------------------------------------------- */
typedef struct MRayRenderer MRayRenderer;

int mray_renderer_create(const MRayAccel* accel, const MRaySceneDesc* hScene,
                         const MRayRenderParams* hParams, MRayRenderer** out);
int mray_renderer_destroy(MRayRenderer* r);
/* One DoRenderWork(): `iterations` wavefront bounces (refill, cast, shade,
 * shadow cast, film).  No host synchronisation inside.                     */
int mray_renderer_iterate(MRayRenderer* r, uint32_t iterations, void* stream);
/* Run bounces until every totalSPP sample is in the film.                  */
int mray_renderer_render_all(MRayRenderer* r, MRayRenderStats* hStats);
/* Synchronises, then fills the stats of everything run since the last call. */
int mray_renderer_stats(MRayRenderer* r, MRayRenderStats* hStats);
/* Film (RenderImage SoA): hOut [4][height][width] = R, G, B, weight sums;
 * row 0 is the bottom row of the camera plane.                             */
int mray_renderer_read_film(MRayRenderer* r, float* hOut);
/* Film taken for one image section: hOut as mray_renderer_read_film, then
 * the device film is cleared, so consecutive takes hold disjoint sums of the
 * paths splatted in between -- what RenderImage::TransferToHost hands the
 * runner each iteration (Tracer/RenderImage.cpp:164-209), which then
 * accumulates sum / weight (MRay/RunCommand.cpp:609-700).                    */
int mray_renderer_take_film(MRayRenderer* r, float* hOut);

/* Tiled renderers (MRAY_RENDER_TILED).  Tiles are numbered row-major,
 * x fastest; edge tiles may be smaller than tileSize.                      */
typedef struct MRayTileInfo
{
    uint32_t tileSize[2];     /* covering tile size, pixels (= path slots)  */
    uint32_t tileCount[2];
} MRayTileInfo;
typedef struct MRayTileResult
{
    uint32_t tileIndex;       /* the tile the call rendered                 */
    uint32_t pixelMin[2];     /* its pixels [pixelMin, pixelMax)            */
    uint32_t pixelMax[2];
    uint32_t tileSPP;         /* samples of that tile after the call        */
} MRayTileResult;
int mray_renderer_tiling(const MRayRenderer* r, MRayTileInfo* hOut);
/* One DoLatencyRender: `passCount` more samples of every pixel of the
 * current tile, run until all of them are in the film, then the next tile
 * becomes current.  passCount 0 renders nothing and still moves on.  The
 * tile's random states are loaded when the tile changes and stored back
 * when it is left (RNGGroupIndependent::SetupRange).  hStats as
 * mray_renderer_stats.                                                      */
int mray_renderer_render_tile(MRayRenderer* r, uint32_t passCount, MRayTileResult* hTile,
                              MRayRenderStats* hStats);
/* The last rendered tile's sums, packed: hOut [4][h][w] of that tile (R, G,
 * B, weight; row 0 = the tile's bottom row); those film pixels are then
 * cleared, as RenderImage::TransferToHost hands over a tile buffer.        */
int mray_renderer_take_tile(MRayRenderer* r, float* hOut);
/* Film [4][height][width] (R, G, B sums, weight sums; row 0 = bottom) as an
 * OpenEXR file of per-pixel means, RGB 32-bit float, top row first: the
 * image the reference's run command writes on triggerSave
 * (MRay/RunCommand.cpp:873-930, ImageLoader.cpp:90-140).  No GPU work.     */
int mray_film_write_exr(const float* hFilm, uint32_t width, uint32_t height, const char* path);
/* Reset the film and all path slots (StartRender again, same seed).        */
int mray_renderer_reset(MRayRenderer* r);
/* Measurement (no reference counterpart): profiling = HIP events around the
 * refill / trace / shade / shadow phases on the launch stream; counting =
 * closest-hit traversal also counts nodes visited and triangles tested
 * (reported in MRayRenderStats).  Resets the accumulated phase times.       */
int mray_renderer_set_measurement(MRayRenderer* r, uint32_t profiling, uint32_t counting);
/* Accumulated phase times hMs[4] (ms) and launch counts hLaunches[4] for
 * refill, trace, shade, shadow.                                              */
int mray_renderer_phase_times(MRayRenderer* r, double* hMs, uint64_t* hLaunches);
/* Device pointer of the film [4][height][width] (stays owned by r).         */
int mray_renderer_film_device(MRayRenderer* r, float** dFilm);

/* ---------------------------------------------------------------------- */
/* Batch: many scenes and viewports of uneven size in one renderer.  No    */
/* reference counterpart: a TracerI renders one camera of one committed    */
/* scene at a time (Core/TracerI.h:330-363, CommitSurfaces / StartRender / */
/* DoRenderWork).  Every bounce of the batch is one launch per kernel over */
/* one shared pool of path slots; paths take global indices view by view  */
/* (view k owns [sum of earlier width*height*spp, + its own)), so each     */
/* view's image is the one a single renderer with the same slot pool would */
/* produce for that path range.  The returned MRayRenderer works with      */
/* mray_renderer_iterate / _render_all / _stats / _reset / _destroy /      */
/* _set_measurement / _phase_times; mray_renderer_read_film returns every  */
/* view's film, concatenated in view order.                                */
/* ---------------------------------------------------------------------- */
typedef struct MRayBatchScene MRayBatchScene;

/* Merge the scenes of `count` accelerators (same BVH width) into one set of
 * device arrays: scene k keeps its triangles, tree, materials, lights,
 * boundary light and camera.  The accelerators may be destroyed afterwards. */
int mray_batch_scene_create(const MRayAccel* const* accels, uint32_t count, MRayBatchScene** out);
int mray_batch_scene_destroy(MRayBatchScene* scene);
int mray_batch_scene_info(const MRayBatchScene* scene, uint32_t* hScenes, uint32_t* hTriangles,
                          uint32_t* hNodes, uint32_t* hWidth);

typedef struct MRayBatchView
{
    uint32_t   scene;          /* index into the merged scenes               */
    uint32_t   width, height;
    uint32_t   spp;
    uint32_t   useCamera;      /* 1: `camera` below; 0: the scene's camera   */
    uint32_t   reserved;
    MRayCamera camera;
} MRayBatchView;

/* hParams: sampleMode, rrRange, spectral, seed, filterRadius, flags as for
 * one scene; maxRayCount = path slots of the pool (0: min(total pixels,
 * 2^21)); width, height and totalSPP are ignored (per view).              */
int mray_batch_create(const MRayBatchScene* scene, const MRayBatchView* hViews, uint32_t viewCount,
                      const MRayRenderParams* hParams, MRayRenderer** out);
/* Film of one view: hOut [4][height][width], as mray_renderer_read_film.   */
int mray_batch_read_film(MRayRenderer* r, uint32_t view, float* hOut);
/* Per-view latency: hMs[v] = device wall-clock time from the frame's first
 * refill to the refill that splats view v's last paths (v and every view
 * before it have then no path running or unissued), or -1 while v is
 * unfinished.                                                              */
int mray_batch_view_times(MRayRenderer* r, double* hMs);

/* Autotuner for a batch (no reference counterpart).  Renders the batch at
 * 1 spp per view for each candidate setting and keeps the fastest: BVH
 * width (one merged scene per width offered, NULL for a width not built),
 * material-keyed shading queues on/off, path-slot pool size and the
 * traversal workgroup size.  hTrials [maxTrials] (may be NULL) receives
 * every measured candidate; *hBest the fastest.                            */
typedef struct MRayBatchTune
{
    uint32_t width;            /* 4 or 8                                     */
    uint32_t queues;           /* 1: MRAY_RENDER_MATERIAL_QUEUES             */
    uint32_t slots;            /* maxRayCount                                */
    uint32_t traceTPB;         /* traversal workgroup size (128 or 256)      */
    double   mpathsPerS;       /* measured 1-spp throughput                  */
} MRayBatchTune;
int mray_batch_autotune(const MRayBatchScene* scene4, const MRayBatchScene* scene8,
                        const MRayBatchView* hViews, uint32_t viewCount,
                        const MRayRenderParams* hParams, MRayBatchTune* hBest,
                        MRayBatchTune* hTrials, uint32_t maxTrials, uint32_t* hTrialCount);
/* Use traversal workgroup size tpb (128 or 256) for a batch renderer's
 * closest-hit and shadow launches (the autotuner's knob; default 256).     */
int mray_batch_set_trace_tpb(MRayRenderer* r, uint32_t tpb);

/* The light-selection tables of a renderer created with
 * MRAY_RENDER_LIGHT_POWER, one entry per emissive triangle of its scene(s):
 * hProb [n] the bucket's own share, hAlias [n] its alias (an index within
 * its scene's lights), hPmf [n] each light's selection probability.  No
 * reference counterpart (its light sampler is uniform).                     */
int mray_renderer_light_table(const MRayRenderer* r, float* hProb, uint32_t* hAlias, float* hPmf);

/* ---------------------------------------------------------------------- */
/* Instanced scenes: meshes placed many times under a tree over instances  */
/* (the reference's two-level accelerator: a base accelerator over surface */
/* instances, AcceleratorLBVH.cu:437-495, above one tree per mesh,          */
/* AcceleratorWork.kt.h:177-244).  Transforms are rotation, per-axis scale  */
/* and translation (TRS).  A mesh's emissive triangles become one           */
/* light per instance; its materials are shared by its instances.          */
/* ---------------------------------------------------------------------- */
typedef struct MRayInstScene MRayInstScene;

typedef struct MRayInstance
{
    uint32_t mesh;             /* index into the meshes                    */
    float    rotation[4];      /* unit quaternion (w, x, y, z), object->world */
    float    scale[3];         /* per axis, > 0 (object space, applied first) */
    float    translation[3];
} MRayInstance;

typedef struct MRayInstEnv
{
    uint32_t   boundaryType;   /* MRAY_BOUNDARY_*                          */
    float      boundaryRadiance[3];
    float      sceneDiameter;  /* 0: the world bounds' XZ span             */
    MRayCamera camera;
    uint32_t   flags;          /* MRAY_INST_*                              */
} MRayInstEnv;

/* Shading frames of instanced triangles.  By default they stay in the
 * mesh's space, as the reference generates surfaces of transformed
 * triangles: positions and the geometric normal are transformed, the
 * interpolated TBN quaternion is not (PrimitiveDefaultTriangle.hpp:443-630,
 * "we can't apply a transform to tbn").  With this flag the frames rotate
 * with the instance, which is what the geometry implies.                  */
#define MRAY_INST_ROTATE_FRAMES 1u

/* meshes: accelerators of the meshes in their own space (4-wide trees);
 * they may be destroyed afterwards.                                        */
int mray_instanced_scene_create(const MRayAccel* const* meshes, uint32_t meshCount,
                                const MRayInstance* hInstances, uint32_t instanceCount,
                                const MRayInstEnv* hEnv, MRayInstScene** out);
int mray_instanced_scene_destroy(MRayInstScene* scene);
/* instances, triangles summed over instances, lights, instance-tree nodes */
int mray_instanced_scene_info(const MRayInstScene* scene, uint32_t* hInstances,
                              uint64_t* hEffectiveTriangles, uint32_t* hLights,
                              uint32_t* hTlasNodes);
/* A renderer over an instanced scene: the camera of hEnv; no material queues. */
int mray_renderer_create_instanced(const MRayInstScene* scene, const MRayRenderParams* hParams,
                                   MRayRenderer** out);
/* Several views of one instanced scene in one slot pool (the batch
 * renderer's schedule, views' scene field 0; useCamera 0 takes hEnv's
 * camera).  The batch read-film and view-times calls apply; no material
 * queues, traversal workgroups of 256.                                     */
int mray_instanced_batch_create(const MRayInstScene* scene, const MRayBatchView* hViews,
                                uint32_t viewCount, const MRayRenderParams* hParams,
                                MRayRenderer** out);
/* A textured Skysphere for an instanced scene (as mray_accel_set_environment;
 * the MRayInstEnv boundary must be a Skysphere).                           */
int mray_instanced_scene_set_environment(MRayInstScene* scene, uint32_t width, uint32_t height,
                                         const float* hRgb);

/* ---------------------------------------------------------------------- */
/* Spectral tables (SpectraLUTGen capability, SpectrumContext.cu:298-472):  */
/* RGB->spectrum coefficients lut [3][N][N][N][4] (c0, c1, c2, 0 in nm,     */
/* x fastest), observer cmf [471][3] and illuminant [471] at 360..830 nm,   */
/* XYZ->RGB [9].  Built on first use and cached next to the library.        */
/* ---------------------------------------------------------------------- */
int mray_spectral_tables(int* hN, float* hLut, float* hCmf, float* hIllum, float* hXyzToRgb);
/* The RGB->spectrum coefficients as a reference .mrspectra file
 * (SpectraLUTGen/main.cpp:590-609): "MR_SPECTRA", u32 N = 64, u32 mode = 1,
 * then nine N^3 float volumes [largest channel][c0, c1, c2], x fastest.
 * The library keeps its own tables in that format at
 * <library dir>/SpectraLUT/REC_709.mrspectra, where the reference's loader
 * looks (SpectrumContext.cu:298-316), and reads a file placed there.        */
int mray_spectral_write_mrspectra(const char* path);
/* Read and validate a .mrspectra file with the reference loader's checks
 * (SpectrumContext.cu:317-351: character code, endianness, N, mode) into
 * hLut [3][N][N][N][4] (c0, c1, c2, 0); either output may be NULL.         */
int mray_spectral_read_mrspectra(const char* path, int* hN, float* hLut);
/* Use a .mrspectra file's coefficients for every spectral renderer created
 * afterwards; an error once a spectral renderer has uploaded the tables.   */
int mray_spectral_use_mrspectra(const char* path);


/* ---------------------------------------------------------------------- */
/* Neural radiance cache (no reference counterpart: the reference keeps    */
/* only an irradiance hash grid, Tracer/HashGrid.h:17-476).  Multiresolution*/
/* hash encoding + 64-wide MLP (6 layers, bf16 MFMA, fp32 accumulation),    */
/* trained with Adam on relative L2 loss.                                    */
/* Query record: float[11] = position in [0,1]^3, direction (theta, phi)    */
/* and normal (theta, phi) in [0,1], roughness in [0,1], albedo RGB.         */
/* ---------------------------------------------------------------------- */
typedef struct MRayNrc MRayNrc;

typedef struct MRayNrcConfig
{
    uint32_t levels;           /* hash levels (<= 16), 2 features each        */
    uint32_t log2TableSize;    /* entries per level = 2^log2TableSize         */
    uint32_t minResolution;    /* N_min                                       */
    uint32_t maxResolution;    /* N_max                                       */
    uint32_t maxBatch;         /* largest training batch (multiple of 32)     */
    float    learningRate;
    float    beta1, beta2, epsilon;
    uint32_t flags;            /* MRAY_NRC_LAZY_ADAM                           */
    uint64_t seed;             /* parameter initialisation                    */
} MRayNrcConfig;

/* Lazy Adam for the hash grid: an entry whose gradient is zero in a step
 * keeps its parameter and both moments (the MLP is always updated densely).
 * A batch touches at most 8 corners per level per sample, so most of the
 * grid is skipped instead of streamed through every step.                   */
#define MRAY_NRC_LAZY_ADAM 1u

enum
{
    MRAY_NRC_WEIGHTS = 0,      /* fp32 masters [6][64 in][64 out]             */
    MRAY_NRC_WEIGHT_GRADS = 1,
    MRAY_NRC_GRID = 2,         /* fp32 masters [levels][2^log2T][2]           */
    MRAY_NRC_GRID_GRADS = 3
};

int mray_nrc_create(const MRayNrcConfig* hCfg, MRayNrc** out);
int mray_nrc_destroy(MRayNrc* nrc);
/* Radiance of `count` records (or *dCount when dCount is not NULL, read on
 * the device): dOut [n][3].                                                  */
int mray_nrc_infer(MRayNrc* nrc, const float* dRecords, uint32_t count, const uint32_t* dCount,
                   float* dOut, void* stream);
/* Gradients of the relative L2 loss of one batch (count a multiple of 32),
 * accumulated into the gradient buffers; mray_nrc_adam applies and clears them. */
int mray_nrc_backward(MRayNrc* nrc, const float* dRecords, const float* dTargets, uint32_t count,
                      void* stream);
int mray_nrc_adam(MRayNrc* nrc, void* stream);
/* backward + adam */
int mray_nrc_train(MRayNrc* nrc, const float* dRecords, const float* dTargets, uint32_t count,
                   void* stream);
int mray_nrc_last_loss(MRayNrc* nrc, float* hLoss);
int mray_nrc_param_counts(const MRayNrc* nrc, uint64_t* hMlp, uint64_t* hGrid);
int mray_nrc_read(MRayNrc* nrc, uint32_t which, float* hOut);
int mray_nrc_write(MRayNrc* nrc, uint32_t which, const float* hIn);
int mray_nrc_zero_grads(MRayNrc* nrc, void* stream);
/* Per-kernel timing of the cache (HIP events around each launch on its
 * stream; no reference counterpart, it instruments the learned part for the
 * MFMA roofline): while on, every launch is timed; mray_nrc_kernel_times
 * synchronises and returns accumulated ms and launch counts [6] for
 * inference forward, training forward, backward, weight gradients, grid
 * scatter and Adam.                                                        */
int mray_nrc_set_profiling(MRayNrc* nrc, uint32_t on);
int mray_nrc_kernel_times(MRayNrc* nrc, double* hMs, uint64_t* hLaunches);


/* Renderer + radiance cache (PathTracerRGB or PathTracerSpectral: the cache
 * answers RGB, which spectral paths take at their wavelengths as an emitter's
 * RGB radiance; spectral training targets are converted to RGB as the film
 * converts a path).  One path slot in 2^trainShift
 * is a training path (full length, its vertices become training pairs); the
 * others terminate into the cache at a surface vertex of depth >= minDepth
 * once the path spread exceeds spreadThreshold x the primary spread (Mueller
 * et al. 2021, sec. 4).  With MRAY_NRC_NEE_AT_END the terminal vertex still
 * takes its next-event estimate and the cache supplies only the radiance the
 * path would gather after it (training targets exclude each vertex's visible
 * NEE); without it the cache supplies the vertex's whole reflected radiance.
 * After every frame the cache takes trainSteps Adam steps of trainBatch pairs
 * (trainBatch * trainSteps a power of two: the pair ring).                   */
#define MRAY_NRC_NEE_AT_END 1u
typedef struct MRayNrcRenderParams
{
    uint32_t trainShift;
    float    spreadThreshold;
    uint32_t trainBatch;
    uint32_t trainSteps;
    uint32_t minDepth;         /* >= 1                                        */
    uint32_t flags;            /* MRAY_NRC_NEE_AT_END                          */
} MRayNrcRenderParams;

/* nrc == NULL detaches.                                                      */
int mray_renderer_attach_nrc(MRayRenderer* r, MRayNrc* nrc, const MRayNrcRenderParams* hParams);
/* One cache per view: count = 1 for a one-scene renderer, the view count
 * (at most 32) for a batch renderer, whose view k then terminates into, is
 * answered by and trains caches[k] (queries, training rings and the scene
 * box are per view).  count = 0 detaches.  No reference counterpart: the
 * per-viewport caches of a batched renderer.                                */
int mray_renderer_attach_nrcs(MRayRenderer* r, MRayNrc* const* caches, uint32_t count,
                              const MRayNrcRenderParams* hParams);
/* One frame: spp more samples per pixel into the film (bounces until they are
 * all finished), then the caches' training steps when caches are attached.
 * A batch renderer renders frames in frame order: each frame gives every view
 * spp samples (the views' own spp are not used); its first frame restarts it
 * in that order, and mray_renderer_reset returns it to view order (the order
 * mray_renderer_render_all needs).                                           */
int mray_renderer_render_frame(MRayRenderer* r, uint32_t spp, MRayRenderStats* hStats);
/* Cache queries answered and training pairs produced since the last reset.  */
int mray_renderer_cache_stats(MRayRenderer* r, uint64_t* hQueries, uint64_t* hTrainingPairs);

#ifdef __cplusplus
}
#endif

#endif /* MRAY_AMD_H */
