"""ctypes mirror of include/mray_amd.h (struct layouts and constants).

The C structs are the drop-in boundary's plain-data types; this module only
describes their layout so Python callers (tests, bench, the oracle wrapper)
can build them.  It loads nothing.
"""
import ctypes as C

MRAY_OK = 0

MAT_LAMBERT = 0
MAT_REFLECT = 1
MAT_REFRACT = 2
MAT_UNREAL = 3

BOUNDARY_NULL = 0
BOUNDARY_SKYSPHERE_COOCTA = 1
BOUNDARY_SKYSPHERE_SPHERICAL = 2

LIGHT_FLAG = 0x80000000

SAMPLE_PURE = 0
SAMPLE_NEE = 1
SAMPLE_NEE_MIS = 2

INVALID_PRIM = 0xFFFFFFFF

BUILDER_SAH = 0
BUILDER_LBVH = 1


class MRayAccelOptions(C.Structure):
    _fields_ = [("builder", C.c_uint32), ("width", C.c_uint32), ("flags", C.c_uint32),
                ("reserved", C.c_uint32)]


class MRayMaterial(C.Structure):
    _fields_ = [
        ("type", C.c_uint32),
        ("albedo", C.c_float * 3),
        ("roughness", C.c_float),
        ("metallic", C.c_float),
        ("specular", C.c_float),
        ("cauchyFront", C.c_float * 3),
        ("cauchyBack", C.c_float * 3),
    ]


class MRayCamera(C.Structure):
    _fields_ = [
        ("position", C.c_float * 3),
        ("gaze", C.c_float * 3),
        ("up", C.c_float * 3),
        ("fov", C.c_float * 2),
        ("nearFar", C.c_float * 2),
    ]


class MRaySceneDesc(C.Structure):
    _fields_ = [
        ("vertexCount", C.c_uint32),
        ("triangleCount", C.c_uint32),
        ("hPositions", C.POINTER(C.c_float)),
        ("hTbnQuats", C.POINTER(C.c_float)),
        ("hIndices", C.POINTER(C.c_uint32)),
        ("hSurfaceKeys", C.POINTER(C.c_uint32)),
        ("hCullBackFace", C.POINTER(C.c_uint8)),
        ("materialCount", C.c_uint32),
        ("hMaterials", C.POINTER(MRayMaterial)),
        ("lightCount", C.c_uint32),
        ("hLightTriangles", C.POINTER(C.c_uint32)),
        ("hLightRadiance", C.POINTER(C.c_float)),
        ("hLightTwoSided", C.POINTER(C.c_uint8)),
        ("boundaryType", C.c_uint32),
        ("boundaryRadiance", C.c_float * 3),
        ("sceneDiameter", C.c_float),
        ("camera", MRayCamera),
    ]


class MRayRenderParams(C.Structure):
    _fields_ = [
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("totalSPP", C.c_uint32),
        ("rrRange", C.c_uint32 * 2),
        ("sampleMode", C.c_uint32),
        ("spectral", C.c_uint32),
        ("maxRayCount", C.c_uint32),
        ("seed", C.c_uint64),
        ("filterRadius", C.c_float),
        ("flags", C.c_uint32),
    ]


class MRayBatchView(C.Structure):
    _fields_ = [
        ("scene", C.c_uint32),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("spp", C.c_uint32),
        ("useCamera", C.c_uint32),
        ("reserved", C.c_uint32),
        ("camera", MRayCamera),
    ]


class MRayBatchTune(C.Structure):
    _fields_ = [
        ("width", C.c_uint32),
        ("queues", C.c_uint32),
        ("slots", C.c_uint32),
        ("traceTPB", C.c_uint32),
        ("mpathsPerS", C.c_double),
    ]


def batch_views(specs):
    """MRayBatchView array from dicts {scene, width, height, spp[, camera]};
    a camera dict (position, gaze, up, fov, nearFar) overrides the scene's."""
    arr = (MRayBatchView * len(specs))()
    for k, d in enumerate(specs):
        v = arr[k]
        v.scene, v.width, v.height, v.spp = d["scene"], d["width"], d["height"], d["spp"]
        cam = d.get("camera")
        v.useCamera = 1 if cam is not None else 0
        if cam is not None:
            for name in ("position", "gaze", "up", "fov", "nearFar"):
                vals = cam[name]
                for i, x in enumerate(vals):
                    getattr(v.camera, name)[i] = x
    return arr


class MRayInstance(C.Structure):
    _fields_ = [
        ("mesh", C.c_uint32),
        ("rotation", C.c_float * 4),
        ("scale", C.c_float * 3),
        ("translation", C.c_float * 3),
    ]


class MRayInstEnv(C.Structure):
    _fields_ = [
        ("boundaryType", C.c_uint32),
        ("boundaryRadiance", C.c_float * 3),
        ("sceneDiameter", C.c_float),
        ("camera", MRayCamera),
        ("flags", C.c_uint32),
    ]


INST_ROTATE_FRAMES = 1


def camera_struct(cam, out=None):
    c = out if out is not None else MRayCamera()
    for k in range(3):
        c.position[k] = cam["position"][k]
        c.gaze[k] = cam["gaze"][k]
        c.up[k] = cam["up"][k]
    for k in range(2):
        c.fov[k] = cam["fov"][k]
        c.nearFar[k] = cam["nearFar"][k]
    return c


def instances_array(instances):
    """MRayInstance array from dicts {mesh, rotation (w,x,y,z), scale (one
    number or three), translation}."""
    arr = (MRayInstance * len(instances))()
    for k, d in enumerate(instances):
        arr[k].mesh = d["mesh"]
        for i in range(4):
            arr[k].rotation[i] = d["rotation"][i]
        sc = d["scale"]
        sc = (sc, sc, sc) if isinstance(sc, (int, float)) else tuple(sc)
        for i in range(3):
            arr[k].scale[i] = sc[i]
        for i in range(3):
            arr[k].translation[i] = d["translation"][i]
    return arr


def inst_env(env):
    e = MRayInstEnv()
    e.boundaryType = env["boundary_type"]
    for k in range(3):
        e.boundaryRadiance[k] = env["boundary_radiance"][k]
    e.sceneDiameter = env.get("scene_diameter") or 0.0
    camera_struct(env["camera"], e.camera)
    e.flags = INST_ROTATE_FRAMES if env.get("rotate_frames") else 0
    return e


class MRayRenderStats(C.Structure):
    _fields_ = [
        ("deadPaths", C.c_uint64),
        ("totalDeadPaths", C.c_uint64),
        ("extensionRays", C.c_uint64),
        ("shadowRays", C.c_uint64),
        ("iterations", C.c_uint32),
        ("finished", C.c_uint32),
        ("nodeVisits", C.c_uint64),
        ("triTests", C.c_uint64),
        ("hostWaits", C.c_uint32),
        ("deviceDrains", C.c_uint32),
        ("shadowNodeVisits", C.c_uint64),
        ("shadowTriTests", C.c_uint64),
    ]


NRC_WEIGHTS = 0
NRC_WEIGHT_GRADS = 1
NRC_GRID = 2
NRC_GRID_GRADS = 3


class MRayNrcConfig(C.Structure):
    _fields_ = [
        ("levels", C.c_uint32),
        ("log2TableSize", C.c_uint32),
        ("minResolution", C.c_uint32),
        ("maxResolution", C.c_uint32),
        ("maxBatch", C.c_uint32),
        ("learningRate", C.c_float),
        ("beta1", C.c_float),
        ("beta2", C.c_float),
        ("epsilon", C.c_float),
        ("flags", C.c_uint32),
        ("seed", C.c_uint64),
    ]


NRC_NEE_AT_END = 1
NRC_LAZY_ADAM = 1


class MRayNrcRenderParams(C.Structure):
    _fields_ = [
        ("trainShift", C.c_uint32),
        ("spreadThreshold", C.c_float),
        ("trainBatch", C.c_uint32),
        ("trainSteps", C.c_uint32),
        ("minDepth", C.c_uint32),
        ("flags", C.c_uint32),
    ]


RENDER_MATERIAL_QUEUES = 1
RENDER_NO_GRAPH = 2
RENDER_LIGHT_POWER = 4
RENDER_LANE_REFILL = 8
RENDER_TILED = 16
ACCEL_HOST_COLLAPSE = 1


class MRayTileInfo(C.Structure):
    _fields_ = [("tileSize", C.c_uint32 * 2), ("tileCount", C.c_uint32 * 2)]


class MRayTileResult(C.Structure):
    _fields_ = [("tileIndex", C.c_uint32), ("pixelMin", C.c_uint32 * 2),
                ("pixelMax", C.c_uint32 * 2), ("tileSPP", C.c_uint32)]


def render_params(width, height, spp, rr=(8, 8), mode=SAMPLE_NEE_MIS, spectral=False,
                  seed=0, filter_radius=1.0, max_ray_count=0, flags=0):
    """MRayRenderParams with the reference's renderer-attribute meaning
    (PathTracerRenderer.cu:127-142)."""
    p = MRayRenderParams()
    p.width = width
    p.height = height
    p.totalSPP = spp
    p.rrRange[0] = rr[0]
    p.rrRange[1] = rr[1]
    p.sampleMode = mode
    p.spectral = 1 if spectral else 0
    p.maxRayCount = max_ray_count
    p.seed = seed
    p.filterRadius = filter_radius
    p.flags = flags
    return p


def fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def u32ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def u64ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


def u8ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))
