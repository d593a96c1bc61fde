"""Seeded procedural scenes for the path-tracing workloads.

Scenes are plain numpy arrays in the layout of MRaySceneDesc
(include/mray_amd.h): indexed triangles, per-vertex world->tangent
quaternions (the reference stores vertex normals as TBN quaternions,
PrimitiveDefaultTriangle.cu:179), a surface key per triangle (material index
or MRAY_LIGHT_FLAG | light index), back-face culling flags (the scene loader's
default is `cullBackFace: true`, SceneLoaderMRay/JsonNode.hpp:32; light
surfaces never cull, Tracer/AcceleratorCommon.cu:274), a material table,
emissive triangles and a boundary light.  World is Y-up, right-handed.

Random draws come from a bit-exact std::mt19937 (MT19937 below); a uniform
real in [0, 1) is u32 / 2^32.
"""
import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import abi


# ---------------------------------------------------------------------------
# std::mt19937 (ISO C++ [rand.eng.mers]); checked against the standard's
# 10000th-output value in tests.
# ---------------------------------------------------------------------------
class MT19937:
    def __init__(self, seed):
        self.mt = [0] * 624
        self.mt[0] = seed & 0xFFFFFFFF
        for i in range(1, 624):
            prev = self.mt[i - 1]
            self.mt[i] = (1812433253 * (prev ^ (prev >> 30)) + i) & 0xFFFFFFFF
        self.index = 624

    def _twist(self):
        mt = self.mt
        for i in range(624):
            y = (mt[i] & 0x80000000) | (mt[(i + 1) % 624] & 0x7FFFFFFF)
            v = mt[(i + 397) % 624] ^ (y >> 1)
            if y & 1:
                v ^= 0x9908B0DF
            mt[i] = v
        self.index = 0

    def __call__(self):
        if self.index >= 624:
            self._twist()
        y = self.mt[self.index]
        self.index += 1
        y ^= (y >> 11)
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= (y >> 18)
        return y & 0xFFFFFFFF

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * (self() / 4294967296.0)


# ---------------------------------------------------------------------------
# Geometry helpers
# ---------------------------------------------------------------------------
def quat_world_to_tangent(n):
    """Unit quaternions (w,x,y,z) q with R(q) n = +Z, for unit normals n [k,3].

    q = normalize(1 + n.z, n x z); a normal facing -Z uses a half turn about X.
    """
    n = np.asarray(n, dtype=np.float64)
    c = n[:, 2]
    axis = np.stack([n[:, 1], -n[:, 0], np.zeros_like(c)], axis=1)   # n x (0,0,1)
    q = np.concatenate([(1.0 + c)[:, None], axis], axis=1)
    flip = c < -1.0 + 1e-7
    q[flip] = np.array([0.0, 1.0, 0.0, 0.0])
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return q.astype(np.float32)


def face_normal(p0, p1, p2):
    n = np.cross(p1 - p0, p2 - p0)
    return n / np.linalg.norm(n, axis=-1, keepdims=True)


@dataclass
class Scene:
    positions: np.ndarray                  # [V,3] f32
    tbn: np.ndarray                        # [V,4] f32
    indices: np.ndarray                    # [T,3] u32
    surface_keys: np.ndarray               # [T] u32
    cull: np.ndarray                       # [T] u8
    materials: list                        # list of dicts
    light_tris: np.ndarray                 # [L] u32
    light_radiance: np.ndarray             # [L,3] f32
    light_two_sided: np.ndarray            # [L] u8
    boundary_type: int = abi.BOUNDARY_NULL
    boundary_radiance: tuple = (0.0, 0.0, 0.0)
    scene_diameter: float = None           # None: from the AABB, as the tracer does
    camera: dict = field(default_factory=dict)
    name: str = ""
    env_map: np.ndarray = None             # [H,W,3] f32 Skysphere texels (CoOcta uv), or None
    _keep: list = field(default_factory=list, repr=False)

    @property
    def triangle_count(self):
        return int(self.indices.shape[0])

    @property
    def diameter(self):
        """Boundary-light scene diameter: length of the XZ span of the scene
        AABB (TracerBase.cpp:1651-1664)."""
        if self.scene_diameter is not None:
            return float(self.scene_diameter)
        p = np.asarray(self.positions, np.float32)[np.asarray(self.indices).reshape(-1)]
        span = p.max(axis=0) - p.min(axis=0)
        return float(np.float32(np.sqrt(np.float32(span[0] * span[0] + span[2] * span[2]))))

    def desc(self):
        """MRaySceneDesc pointing at this scene's arrays (kept alive by self)."""
        arrs = {
            "pos": np.ascontiguousarray(self.positions, dtype=np.float32),
            "tbn": np.ascontiguousarray(self.tbn, dtype=np.float32),
            "idx": np.ascontiguousarray(self.indices, dtype=np.uint32),
            "key": np.ascontiguousarray(self.surface_keys, dtype=np.uint32),
            "cull": np.ascontiguousarray(self.cull, dtype=np.uint8),
            "ltri": np.ascontiguousarray(self.light_tris, dtype=np.uint32),
            "lrad": np.ascontiguousarray(self.light_radiance, dtype=np.float32),
            "l2s": np.ascontiguousarray(self.light_two_sided, dtype=np.uint8),
        }
        mats = (abi.MRayMaterial * max(1, len(self.materials)))()
        for i, m in enumerate(self.materials):
            mats[i].type = m["type"]
            for k in range(3):
                mats[i].albedo[k] = m.get("albedo", (0.0, 0.0, 0.0))[k]
                mats[i].cauchyFront[k] = m.get("cauchyFront", (1.0, 0.0, 0.0))[k]
                mats[i].cauchyBack[k] = m.get("cauchyBack", (1.0, 0.0, 0.0))[k]
            mats[i].roughness = m.get("roughness", 0.0)
            mats[i].metallic = m.get("metallic", 0.0)
            mats[i].specular = m.get("specular", 0.0)
        d = abi.MRaySceneDesc()
        d.vertexCount = arrs["pos"].shape[0]
        d.triangleCount = arrs["idx"].shape[0]
        d.hPositions = abi.fptr(arrs["pos"])
        d.hTbnQuats = abi.fptr(arrs["tbn"])
        d.hIndices = abi.u32ptr(arrs["idx"])
        d.hSurfaceKeys = abi.u32ptr(arrs["key"])
        d.hCullBackFace = abi.u8ptr(arrs["cull"])
        d.materialCount = len(self.materials)
        d.hMaterials = C.cast(mats, C.POINTER(abi.MRayMaterial))
        d.lightCount = arrs["ltri"].shape[0]
        d.hLightTriangles = abi.u32ptr(arrs["ltri"])
        d.hLightRadiance = abi.fptr(arrs["lrad"])
        d.hLightTwoSided = abi.u8ptr(arrs["l2s"])
        d.boundaryType = self.boundary_type
        for k in range(3):
            d.boundaryRadiance[k] = self.boundary_radiance[k]
        d.sceneDiameter = self.diameter
        cam = self.camera
        for k in range(3):
            d.camera.position[k] = cam["position"][k]
            d.camera.gaze[k] = cam["gaze"][k]
            d.camera.up[k] = cam["up"][k]
        d.camera.fov[0] = cam["fov"][0]
        d.camera.fov[1] = cam["fov"][1]
        d.camera.nearFar[0] = cam["nearFar"][0]
        d.camera.nearFar[1] = cam["nearFar"][1]
        self._keep = [arrs, mats]
        return d


class _Builder:
    """Accumulates triangles with their own vertices (flat or smooth)."""

    def __init__(self):
        self.pos = []
        self.nrm = []
        self.idx = []
        self.key = []
        self.cull = []
        self.nv = 0

    def add_mesh(self, positions, indices, key, cull, normals=None):
        positions = np.asarray(positions, dtype=np.float64)
        indices = np.asarray(indices, dtype=np.int64)
        if normals is None:
            # flat: one vertex set per triangle
            p = positions[indices.reshape(-1)].reshape(-1, 3, 3)
            n = face_normal(p[:, 0], p[:, 1], p[:, 2])
            positions = p.reshape(-1, 3)
            normals = np.repeat(n, 3, axis=0)
            indices = np.arange(positions.shape[0]).reshape(-1, 3)
        self.pos.append(positions)
        self.nrm.append(np.asarray(normals, dtype=np.float64))
        self.idx.append(indices + self.nv)
        t = indices.shape[0]
        keys = np.broadcast_to(np.asarray(key, dtype=np.uint64), (t,)).astype(np.uint32)
        self.key.append(keys)
        self.cull.append(np.broadcast_to(np.asarray(cull, dtype=np.uint8), (t,)).copy())
        self.nv += positions.shape[0]
        return sum(a.shape[0] for a in self.idx) - t   # first triangle index

    def finish(self):
        pos = np.concatenate(self.pos).astype(np.float32)
        nrm = np.concatenate(self.nrm)
        nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
        return (pos, quat_world_to_tangent(nrm),
                np.concatenate(self.idx).astype(np.uint32),
                np.concatenate(self.key).astype(np.uint32),
                np.concatenate(self.cull).astype(np.uint8))


def _quad(c, u, v):
    """Two CCW triangles of the quad c +- u +- v; normal = u x v."""
    c, u, v = (np.asarray(x, dtype=np.float64) for x in (c, u, v))
    p = np.stack([c - u - v, c + u - v, c + u + v, c - u + v])
    return p, np.array([[0, 1, 2], [0, 2, 3]])


def _quad_grid(c, u, v, n):
    """The quad c +- u +- v split into n x n cells, two CCW triangles each
    (2 n^2); n = 1 is _quad."""
    if n <= 1:
        return _quad(c, u, v)
    c, u, v = (np.asarray(x, dtype=np.float64) for x in (c, u, v))
    g = np.linspace(-1.0, 1.0, n + 1)
    p = (c[None, None] + g[:, None, None] * u[None, None] + g[None, :, None] * v[None, None])
    p = p.reshape(-1, 3)                           # point (i, j) at i * (n + 1) + j
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    a = (i * (n + 1) + j).ravel()
    b = ((i + 1) * (n + 1) + j).ravel()
    cc = ((i + 1) * (n + 1) + j + 1).ravel()
    d = (i * (n + 1) + j + 1).ravel()
    t = np.stack([np.stack([a, b, cc], 1), np.stack([a, cc, d], 1)], 1).reshape(-1, 3)
    return p, t


def _box(center, half, angle_y, n=1):
    """12 outward-facing triangles of a box rotated about +Y (each face split
    into n x n cells: 12 n^2 triangles)."""
    cx, cy, cz = center
    hx, hy, hz = half
    ca, sa = math.cos(angle_y), math.sin(angle_y)
    rot = np.array([[ca, 0.0, sa], [0.0, 1.0, 0.0], [-sa, 0.0, ca]])
    faces = [((1, 0, 0), (0, 0, -1), (0, 1, 0)),
             ((-1, 0, 0), (0, 0, 1), (0, 1, 0)),
             ((0, 1, 0), (1, 0, 0), (0, 0, -1)),
             ((0, -1, 0), (1, 0, 0), (0, 0, 1)),
             ((0, 0, 1), (1, 0, 0), (0, 1, 0)),
             ((0, 0, -1), (-1, 0, 0), (0, 1, 0))]
    pts, tris = [], []
    h = np.array([hx, hy, hz])
    for n_, u_, v_ in faces:
        nn, u, v = np.array(n_, float), np.array(u_, float), np.array(v_, float)
        p, t = _quad_grid(nn * h, u * h, v * h, n)
        tris.append(t + sum(len(q) for q in pts))
        pts.append(p)
    p = np.concatenate(pts) @ rot.T + np.array([cx, cy, cz])
    return p, np.concatenate(tris)


def fov_y_from_x(fov_x, aspect):
    """SceneLoaderMRay.cpp:595-662: fovY = 2 atan(tan(fovX / 2) / aspect)."""
    return 2.0 * math.atan(math.tan(fov_x * 0.5) / aspect)


# ---------------------------------------------------------------------------
# Cornell-like box: 5 wall quads (10 tris), 2 boxes (24 tris) rotated about Y
# by mt19937(1) angles in [0, 30 deg], one 0.25 x 0.25 ceiling light (2 tris,
# radiance (17, 12, 4)); Lambert red / green / white 0.73.
# ---------------------------------------------------------------------------
def cornell(aspect=1.0, fov_deg=40.0, tess=1, seed=1):
    """The Cornell workload; tess > 1 splits every wall and box face into
    tess x tess cells (34 tess^2 + 2 triangles; the light stays two), seed
    draws the box angles (the workload itself is seed 1)."""
    rng = MT19937(seed)
    a0 = math.radians(rng.uniform(0.0, 30.0))
    a1 = math.radians(rng.uniform(0.0, 30.0))
    mats = [
        {"type": abi.MAT_LAMBERT, "albedo": (0.73, 0.73, 0.73)},   # 0 white
        {"type": abi.MAT_LAMBERT, "albedo": (0.63, 0.065, 0.05)},  # 1 red
        {"type": abi.MAT_LAMBERT, "albedo": (0.14, 0.45, 0.09)},   # 2 green
    ]
    n = max(1, int(tess))
    b = _Builder()
    # walls, normals facing into the box
    b.add_mesh(*_quad_grid((0, -1, 0), (1, 0, 0), (0, 0, -1), n), key=0, cull=1)   # floor, +Y
    b.add_mesh(*_quad_grid((0, 1, 0), (1, 0, 0), (0, 0, 1), n), key=0, cull=1)     # ceiling, -Y
    b.add_mesh(*_quad_grid((0, 0, -1), (1, 0, 0), (0, 1, 0), n), key=0, cull=1)    # back, +Z
    b.add_mesh(*_quad_grid((-1, 0, 0), (0, 0, -1), (0, 1, 0), n), key=1, cull=1)   # left red, +X
    b.add_mesh(*_quad_grid((1, 0, 0), (0, 0, 1), (0, 1, 0), n), key=2, cull=1)     # right green, -X
    b.add_mesh(*_box((0.33, -0.7, 0.3), (0.3, 0.3, 0.3), a0, n), key=0, cull=1)
    b.add_mesh(*_box((-0.33, -0.4, -0.3), (0.3, 0.6, 0.3), a1, n), key=0, cull=1)
    # ceiling light, facing -Y, just below the ceiling
    first = b.add_mesh(*_quad((0, 0.999, 0), (0.125, 0, 0), (0, 0, 0.125)),
                       key=[abi.LIGHT_FLAG | 0, abi.LIGHT_FLAG | 1], cull=0)
    pos, tbn, idx, key, cull = b.finish()
    fov = math.radians(fov_deg)
    cam = {"position": (0.0, 0.0, 3.5), "gaze": (0.0, 0.0, 0.0), "up": (0.0, 1.0, 0.0),
           "fov": (fov, fov_y_from_x(fov, aspect)), "nearFar": (0.1, 1000.0)}
    return Scene(positions=pos, tbn=tbn, indices=idx, surface_keys=key, cull=cull,
                 materials=mats,
                 light_tris=np.array([first, first + 1], dtype=np.uint32),
                 light_radiance=np.array([[17, 12, 4], [17, 12, 4]], dtype=np.float32),
                 light_two_sided=np.zeros(2, dtype=np.uint8),
                 boundary_type=abi.BOUNDARY_NULL,
                 camera=cam, name="cornell")


# ---------------------------------------------------------------------------
# Heightfield + icospheres (1M-triangle class workload)
# ---------------------------------------------------------------------------
def _perlin_table(rng):
    perm = list(range(256))
    for i in range(255, 0, -1):               # Fisher-Yates with mt19937 draws
        j = rng() % (i + 1)
        perm[i], perm[j] = perm[j], perm[i]
    ang = np.array([rng.uniform(0.0, 2.0 * math.pi) for _ in range(256)])
    grads = np.stack([np.cos(ang), np.sin(ang)], axis=1)
    return np.array(perm + perm, dtype=np.int64), grads


def _perlin2(x, y, perm, grads):
    xi = np.floor(x).astype(np.int64)
    yi = np.floor(y).astype(np.int64)
    xf = x - xi
    yf = y - yi
    xi &= 255
    yi &= 255

    def g(ix, iy, dx, dy):
        h = perm[perm[ix] + iy] & 255
        return grads[h, 0] * dx + grads[h, 1] * dy

    def fade(t):
        return t * t * t * (t * (t * 6 - 15) + 10)

    u, v = fade(xf), fade(yf)
    n00 = g(xi, yi, xf, yf)
    n10 = g(xi + 1, yi, xf - 1, yf)
    n01 = g(xi, yi + 1, xf, yf - 1)
    n11 = g(xi + 1, yi + 1, xf - 1, yf - 1)
    nx0 = n00 + u * (n10 - n00)
    nx1 = n01 + u * (n11 - n01)
    return nx0 + v * (nx1 - nx0)


def _fbm(x, y, perm, grads, octaves=6):
    total = np.zeros_like(x)
    amp, freq, norm = 1.0, 2.0, 0.0
    for _ in range(octaves):
        total += amp * _perlin2(x * freq, y * freq, perm, grads)
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return total / norm


def _icosphere1():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
                  [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
         [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
         [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
         [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    verts = list(map(tuple, v))
    cache = {}

    def mid(a, b):
        k = (min(a, b), max(a, b))
        if k not in cache:
            m = (np.array(verts[a]) + np.array(verts[b])) * 0.5
            m /= np.linalg.norm(m)
            verts.append(tuple(m))
            cache[k] = len(verts) - 1
        return cache[k]

    faces = []
    for a, b, c in f:
        ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
        faces += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
    return np.array(verts), np.array(faces, dtype=np.int64)


def terrain_spheres(grid=1024, spheres=2000, width=1920, height=1080, fov_x_deg=60.0,
                    light_grid=4, seed=0):
    """1024 x 1024 vertex heightfield (0.2 x 6-octave fBm, mt19937(2)) plus
    `spheres` icospheres (80 tris, radius log-U[0.01, 0.1], mt19937(3)).
    Materials U{0..3} (mt19937(4)) per terrain patch / sphere: Lambert
    (albedo U[0.1,0.9]^3), Unreal metallic=1 roughness U[0.05,0.5], Refract
    cauchyBack [1.5,0,0], Unreal metallic=0 specular 0.5 roughness 0.3.
    Constant Skysphere (0.5) boundary and light_grid^2 downward quad lights
    (radiance 10) at height 1.5.  seed s > 0 moves the three generator
    seeds to 2 + 3s, 3 + 3s, 4 + 3s (an independent scene of the same kind)."""
    rng_terrain = MT19937(2 + 3 * seed)
    perm, grads = _perlin_table(rng_terrain)
    xs = np.linspace(-1.0, 1.0, grid)
    X, Z = np.meshgrid(xs, xs, indexing="xy")
    Y = 0.2 * _fbm(X, Z, perm, grads)
    # smooth normals by central differences
    dx = np.gradient(Y, xs, axis=1)
    dz = np.gradient(Y, xs, axis=0)
    nrm = np.stack([-dx, np.ones_like(Y), -dz], axis=-1)
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    tpos = np.stack([X, Y, Z], axis=-1).reshape(-1, 3)
    tnrm = nrm.reshape(-1, 3)
    ii, jj = np.meshgrid(np.arange(grid - 1), np.arange(grid - 1), indexing="xy")
    v00 = (jj * grid + ii).reshape(-1)
    v10 = v00 + 1
    v01 = v00 + grid
    v11 = v01 + 1
    # CCW seen from +Y: (v00, v01, v11), (v00, v11, v10)
    tris = np.concatenate([np.stack([v00, v01, v11], 1), np.stack([v00, v11, v10], 1)], 0)
    cell = np.concatenate([np.stack([ii.reshape(-1), jj.reshape(-1)], 1)] * 2, 0)

    rng_mat = MT19937(4 + 3 * seed)
    mats = []

    def new_material():
        kind = rng_mat() % 4
        if kind == 0:
            m = {"type": abi.MAT_LAMBERT,
                 "albedo": tuple(rng_mat.uniform(0.1, 0.9) for _ in range(3))}
        elif kind == 1:
            alb = tuple(rng_mat.uniform(0.1, 0.9) for _ in range(3))
            m = {"type": abi.MAT_UNREAL, "albedo": alb, "metallic": 1.0,
                 "roughness": rng_mat.uniform(0.05, 0.5), "specular": 0.5}
        elif kind == 2:
            m = {"type": abi.MAT_REFRACT, "cauchyFront": (1.0, 0.0, 0.0),
                 "cauchyBack": (1.5, 0.0, 0.0)}
        else:
            alb = tuple(rng_mat.uniform(0.1, 0.9) for _ in range(3))
            m = {"type": abi.MAT_UNREAL, "albedo": alb, "metallic": 0.0,
                 "specular": 0.5, "roughness": 0.3}
        mats.append(m)
        return len(mats) - 1

    patch = 64
    npatch = (grid - 1 + patch - 1) // patch
    patch_mat = np.array([new_material() for _ in range(npatch * npatch)], dtype=np.uint32)
    tkey = patch_mat[(cell[:, 1] // patch) * npatch + (cell[:, 0] // patch)]
    tcull = np.array([0 if mats[k]["type"] == abi.MAT_REFRACT else 1 for k in tkey],
                     dtype=np.uint8)

    b = _Builder()
    b.add_mesh(tpos, tris, key=tkey, cull=tcull, normals=tnrm)

    ico_v, ico_f = _icosphere1()
    rng_sph = MT19937(3 + 3 * seed)
    hf = Y.reshape(-1)
    for _ in range(spheres):
        px = rng_sph.uniform(-0.95, 0.95)
        pz = rng_sph.uniform(-0.95, 0.95)
        r = math.exp(rng_sph.uniform(math.log(0.01), math.log(0.1)))
        gi = int(round((px + 1.0) * 0.5 * (grid - 1)))
        gj = int(round((pz + 1.0) * 0.5 * (grid - 1)))
        py = float(hf[gj * grid + gi]) + r * 0.8
        k = new_material()
        cull = 0 if mats[k]["type"] == abi.MAT_REFRACT else 1
        b.add_mesh(ico_v * r + np.array([px, py, pz]), ico_f, key=k, cull=cull, normals=ico_v)

    light_tris = []
    lsize = 0.06
    for j in range(light_grid):
        for i in range(light_grid):
            cx = -0.75 + 1.5 * i / max(1, light_grid - 1)
            cz = -0.75 + 1.5 * j / max(1, light_grid - 1)
            li = len(light_tris)
            first = b.add_mesh(*_quad((cx, 1.5, cz), (lsize, 0, 0), (0, 0, lsize)),
                               key=[abi.LIGHT_FLAG | li, abi.LIGHT_FLAG | (li + 1)], cull=0)
            light_tris += [first, first + 1]
    pos, tbn, idx, key, cull = b.finish()
    nl = len(light_tris)
    aspect = width / height
    fov = math.radians(fov_x_deg)
    cam = {"position": (0.0, 1.1, 2.1), "gaze": (0.0, 0.0, 0.0), "up": (0.0, 1.0, 0.0),
           "fov": (fov, fov_y_from_x(fov, aspect)), "nearFar": (0.01, 1000.0)}
    return Scene(positions=pos, tbn=tbn, indices=idx, surface_keys=key, cull=cull,
                 materials=mats, light_tris=np.array(light_tris, dtype=np.uint32),
                 light_radiance=np.full((nl, 3), 10.0, dtype=np.float32),
                 light_two_sided=np.zeros(nl, dtype=np.uint8),
                 boundary_type=abi.BOUNDARY_SKYSPHERE_COOCTA,
                 boundary_radiance=(0.5, 0.5, 0.5),
                 camera=cam, name=f"terrain{grid}_spheres{spheres}" + (f"_s{seed}" if seed else ""))


def terrain_spheres_sized(triangles, width, height, seed):
    """terrain_spheres with about `triangles` triangles, terrain and spheres
    in the 1024-grid scene's proportion (2000 spheres per 2.25 M)."""
    spheres = max(1, int(round(2000 * triangles / 2253090)))
    rest = max(2 * 16 * 16, triangles - 80 * spheres - 32)
    grid = int(round(math.sqrt(rest / 2.0))) + 1
    return terrain_spheres(grid=grid, spheres=spheres, width=width, height=height, seed=seed)


def batch_workload(count=32, seed=1, min_tris=10_000, max_tris=2_000_000, spp=16):
    """Independent terrain scenes with triangle counts log-uniform in
    [min_tris, max_tris] and resolutions drawn from 256^2, 512^2, 1280x720
    and 1920x1080 (mt19937(seed)).  Returns the scenes and the view dicts
    for mray_batch_create (view k renders scene k with its own camera)."""
    rng = MT19937(seed)
    res = [(256, 256), (512, 512), (1280, 720), (1920, 1080)]
    out, views = [], []
    for k in range(count):
        t = math.exp(rng.uniform(math.log(min_tris), math.log(max_tris)))
        w, h = res[rng() % 4]
        out.append(terrain_spheres_sized(int(t), w, h, seed=10 + k))
        views.append({"scene": k, "width": w, "height": h, "spp": spp})
    return out, views


def cornell_batch_workload(count=32, seed=5, min_tris=10_000, max_tris=2_000_000, spp=16):
    """The batched workload as specified: `count` Cornell scenes (the Cornell
    generator, box angles from their own seeds) tessellated to triangle
    counts log-uniform in [min_tris, max_tris], resolutions drawn from 256^2,
    512^2, 1280x720 and 1920x1080, all from mt19937(seed).  Returns the
    scenes and the view dicts for mray_batch_create."""
    rng = MT19937(seed)
    res = [(256, 256), (512, 512), (1280, 720), (1920, 1080)]
    out, views = [], []
    for k in range(count):
        t = math.exp(rng.uniform(math.log(min_tris), math.log(max_tris)))
        w, h = res[rng() % 4]
        n = max(1, int(round(math.sqrt(max(t - 2.0, 34.0) / 34.0))))
        sc = cornell(aspect=w / h, tess=n, seed=100 + k)
        sc.name = f"cornell-t{n}"
        out.append(sc)
        views.append({"scene": k, "width": w, "height": h, "spp": spp})
    return out, views


def random_triangles(n, seed=7, scale=1.0):
    """Triangle soup for traversal tests: n random small triangles in a cube."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-scale, scale, size=(n, 1, 3))
    p = c + rng.normal(scale=0.05 * scale, size=(n, 3, 3))
    pos = p.reshape(-1, 3).astype(np.float32)
    idx = np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)
    nrm = face_normal(p[:, 0], p[:, 1], p[:, 2])
    tbn = quat_world_to_tangent(np.repeat(nrm, 3, axis=0))
    cull = (rng.uniform(size=n) < 0.5).astype(np.uint8)
    cam = {"position": (0.0, 0.0, 3.0), "gaze": (0.0, 0.0, 0.0), "up": (0.0, 1.0, 0.0),
           "fov": (math.radians(60), math.radians(60)), "nearFar": (0.01, 100.0)}
    return Scene(positions=pos, tbn=tbn, indices=idx,
                 surface_keys=np.zeros(n, dtype=np.uint32), cull=cull,
                 materials=[{"type": abi.MAT_LAMBERT, "albedo": (0.5, 0.5, 0.5)}],
                 light_tris=np.zeros(0, dtype=np.uint32),
                 light_radiance=np.zeros((0, 3), dtype=np.float32),
                 light_two_sided=np.zeros(0, dtype=np.uint8), camera=cam, name="soup")


def to_mray_json(scene, path):
    """Write `scene` as an MRay JSON scene file (the reference's scene format,
    Docs/markdown/scene/mrayScene.md): one inline nodeTriangleIndexed
    primitive per run of triangles sharing a material (or consecutive lights)
    and a cull flag, in triangle order, so the tracer's flattened scene keeps
    the triangle order.  Normals are recovered from the TBN quaternions."""
    import json
    pos = np.asarray(scene.positions, np.float32)
    idx = np.asarray(scene.indices, np.int64)
    keys = np.asarray(scene.surface_keys, np.uint32)
    cull = np.asarray(scene.cull, np.uint8)
    q = np.asarray(scene.tbn, np.float64)
    # n = R(q)^T z for the world-to-tangent rotation q
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    nrm = np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], 1)
    lrad = np.asarray(scene.light_radiance, np.float32)
    l2s = np.asarray(scene.light_two_sided, np.uint8)
    runs = []
    for t in range(len(keys)):
        light = bool(keys[t] & abi.LIGHT_FLAG)
        k = int(keys[t] & ~np.uint32(abi.LIGHT_FLAG)) if light else int(keys[t])
        if runs:
            r = runs[-1]
            if r["light"] and light and r["cull"] == cull[t] and k == r["key"] + r["count"] and \
                    np.array_equal(lrad[k], lrad[r["key"]]) and l2s[k] == l2s[r["key"]]:
                r["count"] += 1
                continue
            if not r["light"] and not light and r["key"] == k and r["cull"] == cull[t]:
                r["count"] += 1
                continue
        runs.append({"first": t, "count": 1, "light": light, "key": k, "cull": int(cull[t])})
    prims, surfaces, lights, light_surfs = [], [], [], []
    lights.append({"id": 0, "type": "Null"})
    for pid, r in enumerate(runs):
        tri = idx[r["first"]:r["first"] + r["count"]]
        vs, inv = np.unique(tri.reshape(-1), return_inverse=True)
        prims.append({"id": pid, "type": "Triangle", "tag": "nodeTriangleIndexed",
                      "position": pos[vs].astype(float).tolist(),
                      "normal": nrm[vs].astype(float).tolist(),
                      "index": inv.reshape(-1, 3).astype(int).tolist()})
        if r["light"]:
            lid = len(lights)
            lights.append({"id": lid, "type": "Primitive", "primitive": pid,
                           "radiance": lrad[r["key"]].astype(float).tolist(),
                           "isTwoSided": bool(l2s[r["key"]])})
            light_surfs.append({"light": lid})
        else:
            surfaces.append({"transform": 0, "material": r["key"], "primitive": pid,
                             "cullBackFace": bool(r["cull"])})
    names = {abi.MAT_LAMBERT: "Lambert", abi.MAT_REFLECT: "Reflect", abi.MAT_REFRACT: "Refract",
             abi.MAT_UNREAL: "Unreal"}
    mats = []
    for mid, m in enumerate(scene.materials):
        e = {"id": mid, "type": names[m["type"]]}
        if m["type"] in (abi.MAT_LAMBERT, abi.MAT_UNREAL):
            e["albedo"] = list(m.get("albedo", (0.0, 0.0, 0.0)))
        if m["type"] == abi.MAT_UNREAL:
            e["roughness"] = m.get("roughness", 0.0)
            e["specular"] = m.get("specular", 0.0)
            e["metallic"] = m.get("metallic", 0.0)
        if m["type"] == abi.MAT_REFRACT:
            e["cauchyBack"] = list(m.get("cauchyBack", (1.0, 0.0, 0.0)))
            e["cauchyFront"] = list(m.get("cauchyFront", (1.0, 0.0, 0.0)))
        mats.append(e)
    boundary_light = 0
    if scene.boundary_type in (abi.BOUNDARY_SKYSPHERE_COOCTA, abi.BOUNDARY_SKYSPHERE_SPHERICAL):
        boundary_light = len(lights)
        sky = ("Skysphere_CoOcta" if scene.boundary_type == abi.BOUNDARY_SKYSPHERE_COOCTA
               else "Skysphere_Spherical")
        lights.append({"id": boundary_light, "type": sky, "radiance": list(scene.boundary_radiance)})
    cam = scene.camera
    fx, fy = cam["fov"]
    doc = {
        "Cameras": [{"id": 0, "type": "Pinhole", "isFovX": True, "fov": math.degrees(fx),
                     "aspect": math.tan(fx * 0.5) / math.tan(fy * 0.5),
                     "planes": list(cam["nearFar"]), "gaze": list(cam["gaze"]),
                     "position": list(cam["position"]), "up": list(cam["up"])}],
        "Lights": lights,
        "Mediums": [{"id": 0, "type": "Vacuum"}],
        "Transforms": [{"id": 0, "type": "Identity"}],
        "Textures": [],
        "Materials": mats,
        "Primitives": prims,
        "Surfaces": surfaces,
        "LightSurfaces": light_surfs,
        "CameraSurfaces": [{"camera": 0}],
        "Boundary": {"light": boundary_light, "transform": 0, "medium": 0},
    }
    with open(path, "w") as f:
        f.write("// written by mray_amd.scenes.to_mray_json\n")
        json.dump(doc, f)



# ---------------------------------------------------------------------------
# Instanced scenes (mray_instanced_scene_create) and their flattening
# ---------------------------------------------------------------------------
def quat_mul(a, b):
    """Hamilton product of (w, x, y, z) quaternions (broadcasting)."""
    aw, ax, ay, az = np.moveaxis(np.asarray(a, np.float64), -1, 0)
    bw, bx, by, bz = np.moveaxis(np.asarray(b, np.float64), -1, 0)
    return np.stack([aw * bw - ax * bx - ay * by - az * bz,
                     aw * bx + ax * bw + ay * bz - az * by,
                     aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw], axis=-1)


def quat_rotate(q, v):
    """q v q* for a unit quaternion q and vectors v [..., 3]."""
    q = np.asarray(q, np.float64)
    v = np.asarray(v, np.float64)
    u = q[1:]
    t = 2.0 * np.cross(u, v)
    return v + q[0] * t + np.cross(u, t)


def _unit_quat(r):
    r = np.asarray(r, np.float64)
    return r / np.linalg.norm(r)


def flatten_instances(meshes, instances, env):
    """The world-space scene an instanced scene describes: every instance's
    triangles transformed (positions by the similarity; tangent frames kept
    as the mesh's, as the reference's transformed triangles keep them, or
    rotated with the instance when env["rotate_frames"]), materials of each
    mesh once, one light per instance of an emissive triangle.  The oracle
    renders this."""
    rotate_frames = bool(env.get("rotate_frames"))
    mat_off, mats = [], []
    for m in meshes:
        mat_off.append(len(mats))
        mats += list(m.materials)
    pos, tbn, idx, keys, cull, ltri, lrad, l2s = [], [], [], [], [], [], [], []
    vbase = tbase = lbase = 0
    for inst in instances:
        m = meshes[inst["mesh"]]
        q = _unit_quat(inst["rotation"])
        s = np.broadcast_to(np.asarray(inst["scale"], np.float64), (3,))
        t = np.asarray(inst["translation"], np.float64)
        pos.append((quat_rotate(q, np.asarray(m.positions, np.float64) * s) + t).astype(np.float32))
        if rotate_frames:
            qc = np.array([q[0], -q[1], -q[2], -q[3]])
            tb = quat_mul(np.asarray(m.tbn, np.float64), qc)
            tbn.append((tb / np.linalg.norm(tb, axis=1, keepdims=True)).astype(np.float32))
        else:
            tbn.append(np.asarray(m.tbn, np.float32))
        idx.append(np.asarray(m.indices, np.uint32) + np.uint32(vbase))
        k = np.asarray(m.surface_keys, np.uint32)
        is_light = (k & abi.LIGHT_FLAG) != 0
        keys.append(np.where(is_light, abi.LIGHT_FLAG | ((k & ~np.uint32(abi.LIGHT_FLAG)) + np.uint32(lbase)),
                             k + np.uint32(mat_off[inst["mesh"]])).astype(np.uint32))
        cull.append(np.asarray(m.cull, np.uint8))
        ltri.append(np.asarray(m.light_tris, np.uint32) + np.uint32(tbase))
        lrad.append(np.asarray(m.light_radiance, np.float32).reshape(-1, 3))
        l2s.append(np.asarray(m.light_two_sided, np.uint8))
        vbase += len(m.positions)
        tbase += m.triangle_count
        lbase += len(m.light_tris)
    return Scene(positions=np.concatenate(pos), tbn=np.concatenate(tbn), indices=np.concatenate(idx),
                 surface_keys=np.concatenate(keys), cull=np.concatenate(cull), materials=mats,
                 light_tris=np.concatenate(ltri), light_radiance=np.concatenate(lrad),
                 light_two_sided=np.concatenate(l2s), boundary_type=env["boundary_type"],
                 boundary_radiance=tuple(env["boundary_radiance"]),
                 scene_diameter=env.get("scene_diameter"), camera=env["camera"], name="instances",
                 env_map=env.get("env_map"))


def light_panel(count_quads, size, radiance, seed=0, span=1.0):
    """A mesh of small downward-facing emissive quads (2 triangles each),
    scattered over [-span, span]^2 at y = 0 (mt19937(seed))."""
    rng = MT19937(seed)
    b = _Builder()
    light_tris = []
    for k in range(count_quads):
        cx, cz = rng.uniform(-span, span), rng.uniform(-span, span)
        li = len(light_tris)
        first = b.add_mesh(*_quad((cx, 0.0, cz), (size, 0, 0), (0, 0, size)),
                           key=[abi.LIGHT_FLAG | li, abi.LIGHT_FLAG | (li + 1)], cull=0)
        light_tris += [first, first + 1]
    pos, tbn, idx, key, cull = b.finish()
    nl = len(light_tris)
    return Scene(positions=pos, tbn=tbn, indices=idx, surface_keys=key, cull=cull,
                 materials=[{"type": abi.MAT_LAMBERT, "albedo": (0.5, 0.5, 0.5)}],
                 light_tris=np.array(light_tris, dtype=np.uint32),
                 light_radiance=np.tile(np.asarray(radiance, np.float32), (nl, 1)),
                 light_two_sided=np.zeros(nl, dtype=np.uint8), name="light_panel",
                 camera={"position": (0.0, -2.0, 0.0), "gaze": (0.0, 0.0, 0.0), "up": (0.0, 0.0, 1.0),
                         "fov": (math.radians(60.0),) * 2, "nearFar": (0.01, 100.0)})


def _random_rotation(rng, full=True):
    """A seeded unit quaternion: uniform over rotations (full) or about +Y."""
    if full:
        u1, u2, u3 = rng.uniform(0, 1), rng.uniform(0, 1), rng.uniform(0, 1)
        a, b = math.sqrt(1 - u1), math.sqrt(u1)
        return (a * math.sin(2 * math.pi * u2), a * math.cos(2 * math.pi * u2),
                b * math.sin(2 * math.pi * u3), b * math.cos(2 * math.pi * u3))
    ang = rng.uniform(0, 2 * math.pi)
    return (math.cos(ang / 2), 0.0, math.sin(ang / 2), 0.0)


def instanced_test_scene(width=64, height=48, seed=3, rotate_frames=False):
    """Small instanced scene for parity tests: two terrain meshes placed 12
    times with rotations and (some per-axis) scales, and an emissive panel mesh placed twice
    (its lights exist once per instance).  With rotate_frames the shading
    frames turn with the instances and rotations are arbitrary; without, the
    frames stay the meshes' (the reference's behaviour) and every rotation
    is about +Y, so no instance shades around a normal that points into its
    own surface."""
    meshes = [terrain_spheres(grid=20, spheres=4, width=width, height=height, light_grid=0, seed=1),
              terrain_spheres(grid=16, spheres=6, width=width, height=height, light_grid=0, seed=2),
              light_panel(3, 0.15, (9.0, 7.0, 5.0), seed=4, span=0.5)]
    rng = MT19937(seed)
    instances = []
    for k in range(12):
        instances.append({"mesh": k % 2,
                          "rotation": _random_rotation(rng, full=rotate_frames and (k % 3 != 0)),
                          "scale": (rng.uniform(0.5, 1.3) if k % 4 != 1 else
                                    (rng.uniform(0.5, 1.3), rng.uniform(0.5, 1.3), rng.uniform(0.5, 1.3))),
                          "translation": (rng.uniform(-2.5, 2.5), rng.uniform(-0.3, 0.3),
                                          rng.uniform(-2.5, 2.5))})
    for k in range(2):
        instances.append({"mesh": 2, "rotation": (1.0, 0.0, 0.0, 0.0) if k == 0 else
                          _random_rotation(rng, full=False),
                          "scale": 1.0 + 0.5 * k, "translation": (-1.0 + 2.0 * k, 2.2, 0.3 * k)})
    fov = math.radians(60.0)
    cam = {"position": (0.0, 4.5, 6.0), "gaze": (0.0, 0.0, 0.0), "up": (0.0, 1.0, 0.0),
           "fov": (fov, fov_y_from_x(fov, width / height)), "nearFar": (0.01, 1000.0)}
    env = {"boundary_type": abi.BOUNDARY_SKYSPHERE_COOCTA, "boundary_radiance": (0.3, 0.3, 0.35),
           "scene_diameter": None, "camera": cam, "rotate_frames": rotate_frames}
    flat = flatten_instances(meshes, instances, env)
    env["scene_diameter"] = flat.diameter
    flat.scene_diameter = env["scene_diameter"]
    return meshes, instances, env, flat


def instanced_workload(mesh_count=8, instance_count=1000, tri_range=(50_000, 200_000),
                       emissive_tris=10_000, width=1920, height=1080, seed=5):
    """Two-level workload: mesh_count terrain meshes (triangles log-uniform in
    tri_range) placed instance_count times over a 60 x 60 field with seeded
    TRS transforms (rotations about +Y, per-axis scales in [0.6, 1.6]), plus
    one panel of emissive_tris emissive triangles above it; constant
    Skysphere."""
    rng = MT19937(seed)
    meshes = []
    for k in range(mesh_count):
        t = math.exp(rng.uniform(math.log(tri_range[0]), math.log(tri_range[1])))
        spheres = max(1, int(round(2000 * t / 2253090)))
        grid = int(round(math.sqrt(max(512, t - 80 * spheres) / 2.0))) + 1
        meshes.append(terrain_spheres(grid=grid, spheres=spheres, width=width, height=height,
                                      light_grid=0, seed=20 + k))
    meshes.append(light_panel(emissive_tris // 2, 0.08, (12.0, 11.0, 9.0), seed=seed, span=28.0))
    instances = []
    for k in range(instance_count):
        instances.append({"mesh": int(rng() % mesh_count), "rotation": _random_rotation(rng, full=False),
                          "scale": (rng.uniform(0.6, 1.6), rng.uniform(0.6, 1.6), rng.uniform(0.6, 1.6)),
                          "translation": (rng.uniform(-30, 30), rng.uniform(-0.2, 0.2),
                                          rng.uniform(-30, 30))})
    instances.append({"mesh": mesh_count, "rotation": (1.0, 0.0, 0.0, 0.0), "scale": 1.0,
                      "translation": (0.0, 4.0, 0.0)})
    fov = math.radians(60.0)
    cam = {"position": (0.0, 9.0, 36.0), "gaze": (0.0, 0.0, 0.0), "up": (0.0, 1.0, 0.0),
           "fov": (fov, fov_y_from_x(fov, width / height)), "nearFar": (0.01, 1000.0)}
    env = {"boundary_type": abi.BOUNDARY_SKYSPHERE_COOCTA, "boundary_radiance": (0.5, 0.5, 0.5),
           "scene_diameter": None, "camera": cam}
    return meshes, instances, env


def instanced_views(spp=16, count=8):
    """count viewports of uneven sizes (1080p down to 512^2) looking at the
    instanced workload's field from around it."""
    sizes = [(1920, 1080), (1280, 720), (1280, 720), (960, 540), (960, 540), (640, 360),
             (640, 360), (512, 512)]
    views = []
    for k in range(count):
        w, h = sizes[k % len(sizes)]
        ang = 2.0 * math.pi * k / count
        fov = math.radians(60.0 if k % 2 == 0 else 45.0)
        cam = {"position": (30.0 * math.sin(ang), 6.0 + 2.0 * (k % 3), 30.0 * math.cos(ang)),
               "gaze": (0.0, 0.0, 0.0), "up": (0.0, 1.0, 0.0),
               "fov": (fov, fov_y_from_x(fov, w / h)), "nearFar": (0.01, 1000.0)}
        views.append({"scene": 0, "width": w, "height": h, "spp": spp, "camera": cam})
    return views


def coocta_to_dir_yup(u, v):
    """Concentric octahedral uv in [0, 1]^2 -> Y-up unit directions (numpy,
    float64): the Skysphere's texture parameterization."""
    a, b = 2.0 * np.asarray(u, np.float64) - 1.0, 2.0 * np.asarray(v, np.float64) - 1.0
    ua, va = np.abs(a), np.abs(b)
    d = 1.0 - (ua + va)
    r = 1.0 - np.abs(d)
    with np.errstate(invalid="ignore", divide="ignore"):
        phi = np.where(r != 0.0, ((va - ua) / np.where(r != 0.0, r, 1.0) + 1.0) * (math.pi / 4), 0.0)
    zs = np.where(np.signbit(d), -1.0, 1.0) * (1.0 - r * r)
    xy = r * np.sqrt(np.maximum(0.0, 2.0 - r * r))
    x = np.where(np.signbit(a), -1.0, 1.0) * np.cos(phi) * xy
    y = np.where(np.signbit(b), -1.0, 1.0) * np.sin(phi) * xy
    return np.stack([y, zs, x], axis=-1)         # Z-up (x, y, z) -> Y-up (y, z, x)


def spherical_to_dir_yup(u, v):
    """Latitude-longitude uv in [0, 1]^2 -> Y-up unit directions (numpy):
    azimuth (u * 2 pi - pi) and inclination ((1 - v) pi) about Z-up."""
    th = np.asarray(u, np.float64) * 2.0 * math.pi - math.pi
    ph = (1.0 - np.asarray(v, np.float64)) * math.pi
    x, y, z = np.cos(th) * np.sin(ph), np.sin(th) * np.sin(ph), np.cos(ph)
    return np.stack([y, z, x], axis=-1)


def sky_env(width=128, height=128, sun=(0.4, 0.6, 0.3), sun_size=0.06, sun_radiance=40.0,
            zenith=(0.15, 0.3, 0.8), horizon=(0.8, 0.85, 0.9), ground=(0.12, 0.1, 0.08),
            mapping="coocta"):
    """A procedural sky texture [height, width, 3] over concentric-octahedral
    (or, with mapping="spherical", latitude-longitude) uv: an elevation
    gradient from ground through horizon to zenith plus a bright sun lobe
    (cosine power) around `sun`."""
    ys, xs = np.meshgrid((np.arange(height) + 0.5) / height, (np.arange(width) + 0.5) / width,
                         indexing="ij")
    d = spherical_to_dir_yup(xs, ys) if mapping == "spherical" else coocta_to_dir_yup(xs, ys)
    el = d[..., 1]
    t = np.clip(el, 0.0, 1.0)[..., None] ** 0.5
    sky = np.asarray(horizon) * (1 - t) + np.asarray(zenith) * t
    rgb = np.where((el >= 0.0)[..., None], sky, np.asarray(ground))
    s = np.asarray(sun, np.float64) / np.linalg.norm(sun)
    c = np.clip(d @ s, 0.0, 1.0)
    lobe = np.exp((c - 1.0) / (sun_size ** 2))
    rgb = rgb + sun_radiance * lobe[..., None] * np.asarray([1.0, 0.9, 0.75])
    return rgb.astype(np.float32)
