"""ctypes wrapper of oracle/librefcpu.so -- the CPU parity oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the mray_amd package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from mray_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "librefcpu.so")
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


_V4_FLAGS = ("avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl")


def host_cpu():
    """(model name, True when the host has the x86-64-v4 AVX-512 set)."""
    model, flags = "unknown x86_64", set()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name") and model == "unknown x86_64":
                model = line.split(":", 1)[1].strip()
            elif line.startswith("flags") and not flags:
                flags = set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return model, all(f in flags for f in _V4_FLAGS)


def use_native_isa():
    """Load the x86-64-v4 build when this host supports it (the CPU
    baseline's build); call before the first lib().  Returns the ISA level."""
    global _LIB
    if _lib is not None:
        raise RuntimeError("the oracle library is already loaded")
    _, v4 = host_cpu()
    v4_lib = os.path.join(_HERE, "librefcpu_v4.so")
    if v4 and os.path.exists(v4_lib):
        _LIB = v4_lib
        return "x86-64-v4"
    return "x86-64-v3"


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        P = C.c_void_p
        f32 = C.POINTER(C.c_float)
        u32 = C.POINTER(C.c_uint32)
        u64 = C.POINTER(C.c_uint64)
        L.rc_morton63_compose.restype = C.c_uint64
        L.rc_morton63_compose.argtypes = [C.c_uint32] * 3
        L.rc_morton30_compose.restype = C.c_uint32
        L.rc_morton30_compose.argtypes = [C.c_uint32] * 3
        L.rc_morton63_decompose.argtypes = [C.c_uint64, u32]
        L.rc_morton30_decompose.argtypes = [C.c_uint32, u32]
        L.rc_prim_centers_aabbs.argtypes = [f32, u32, C.c_uint32, f32, f32]
        L.rc_morton63_codes.argtypes = [f32, C.c_uint32, f32, u64]
        L.rc_morton30_codes.argtypes = [f32, C.c_uint32, f32, u32]
        L.rc_lbvh_from_sorted.argtypes = [u64, u32, f32, C.c_uint32, u32, f32]
        L.rc_lbvh_tree_valid.restype = C.c_int
        L.rc_lbvh_tree_valid.argtypes = [u64, u32, C.c_uint32, C.c_int]
        L.rc_scene_create.restype = P
        L.rc_scene_create.argtypes = [C.POINTER(abi.MRaySceneDesc)]
        L.rc_scene_destroy.argtypes = [P]
        L.rc_scene_lbvh.argtypes = [P, u64, u32, u32, f32, f32]
        L.rc_cast_closest.argtypes = [P, f32, C.c_uint32, C.c_int, u32, f32, f32, u64, C.c_uint32]
        L.rc_cast_visibility.argtypes = [P, f32, C.c_uint32, u32, C.c_uint32]
        L.rc_seed_states.argtypes = [C.c_uint64, C.c_uint32, u32]
        L.rc_pcg_next.restype = C.c_uint32
        L.rc_pcg_next.argtypes = [u32]
        L.rc_to_float01.restype = C.c_float
        L.rc_to_float01.argtypes = [C.c_uint32]
        L.rc_camera_rays.argtypes = [P, C.POINTER(abi.MRayRenderParams), C.c_uint32, f32, f32]
        L.rc_render.restype = C.c_int
        L.rc_render.argtypes = [P, C.POINTER(abi.MRayRenderParams), f32, C.c_uint32,
                                C.c_uint32, C.c_uint64, u64]
        L.rc_render_tiled.restype = C.c_int
        L.rc_render_tiled.argtypes = [P, C.POINTER(abi.MRayRenderParams), u32, C.c_uint32, f32,
                                      C.c_uint32, u32, u32, u64]
        L.rc_render_batch.restype = C.c_int
        L.rc_render_batch.argtypes = [P, C.c_uint32, P, C.c_uint32, C.POINTER(abi.MRayRenderParams),
                                      P, C.c_uint32, u64]
        L.rc_spectral_load.argtypes = [C.c_int, f32]
        L.rc_spectral_colour_tables.argtypes = [f32, f32, f32]
        L.rc_sample_wavelengths.argtypes = [f32, C.c_uint32, f32, f32]
        L.rc_rgb_to_spectrum.argtypes = [f32, f32, C.c_uint32, C.c_int, f32]
        L.rc_spectrum_to_rgb.argtypes = [f32, f32, f32, C.c_uint32, f32]
        L.rc_bsdf_probe.argtypes = [C.POINTER(abi.MRayMaterial), f32, f32, C.c_int, f32, f32,
                                    f32, C.c_int, f32]
        L.rc_light_probe.argtypes = [P, f32, f32, f32]
        L.rc_light_pdf.restype = C.c_float
        L.rc_light_pdf.argtypes = [P, C.c_uint32, C.c_float, C.c_float, f32, f32]
        L.rc_light_power_table.argtypes = [P, f32, u32, f32]
        L.rc_scene_set_environment.restype = C.c_int
        L.rc_scene_set_environment.argtypes = [P, C.c_uint32, C.c_uint32, f32]
        L.rc_env_cdfs.argtypes = [P, f32, f32]
        L.rc_env_sample.argtypes = [P, f32, C.c_uint32, f32]
        L.rc_env_pdf_uv.argtypes = [P, f32, C.c_uint32, f32]
        L.rc_env_eval.argtypes = [P, f32, C.c_uint32, f32, f32]
        L.rc_coocta_to_dir.argtypes = [f32, C.c_uint32, f32]
        L.rc_dir_to_coocta.argtypes = [f32, C.c_uint32, f32]
        L.rc_bsdf_eval_many.argtypes = [C.POINTER(abi.MRayMaterial), f32, f32, C.c_int, f32, f32,
                                        C.c_uint32, f32, f32]
        L.rc_inv_erf.argtypes = [f32, C.c_uint32, f32]
        L.rc_fmath.argtypes = [C.c_int, f32, f32, C.c_uint32, f32]
        L.rc_sample_gaussian.argtypes = [f32, C.c_uint32, C.c_float, f32, f32]
        L.rc_spherical_to_dir.argtypes = [f32, C.c_uint32, f32, f32]
        L.rc_dir_to_spherical.argtypes = [f32, C.c_uint32, f32]
        L.rc_sky_solid_angle_pdf.restype = C.c_float
        L.rc_sky_solid_angle_pdf.argtypes = [C.c_uint32, C.c_float, C.c_float]
        _lib = L
    return _lib


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def morton63_compose(x, y, z):
    return int(lib().rc_morton63_compose(x, y, z))


def morton30_compose(x, y, z):
    return int(lib().rc_morton30_compose(x, y, z))


def morton63_decompose(c):
    out = np.zeros(3, dtype=np.uint32)
    lib().rc_morton63_decompose(c, abi.u32ptr(out))
    return tuple(int(v) for v in out)


def morton30_decompose(c):
    out = np.zeros(3, dtype=np.uint32)
    lib().rc_morton30_decompose(c, abi.u32ptr(out))
    return tuple(int(v) for v in out)


def prim_centers_aabbs(positions, indices):
    pos = _f(positions)
    idx = np.ascontiguousarray(indices, dtype=np.uint32)
    n = idx.shape[0]
    centers = np.zeros((n, 3), np.float32)
    aabbs = np.zeros((n, 6), np.float32)
    lib().rc_prim_centers_aabbs(abi.fptr(pos), abi.u32ptr(idx), n, abi.fptr(centers),
                                abi.fptr(aabbs))
    return centers, aabbs


def morton63_codes(centers, aabb):
    c = _f(centers)
    a = _f(aabb)
    out = np.zeros(c.shape[0], np.uint64)
    lib().rc_morton63_codes(abi.fptr(c), c.shape[0], abi.fptr(a), abi.u64ptr(out))
    return out


def morton30_codes(centers, aabb):
    c = _f(centers)
    a = _f(aabb)
    out = np.zeros(c.shape[0], np.uint32)
    lib().rc_morton30_codes(abi.fptr(c), c.shape[0], abi.fptr(a), abi.u32ptr(out))
    return out


def lbvh_from_sorted(sorted_codes, leaf_order, leaf_aabbs):
    n = len(sorted_codes)
    sc = np.ascontiguousarray(sorted_codes, dtype=np.uint64)
    lo = np.ascontiguousarray(leaf_order, dtype=np.uint32)
    la = _f(leaf_aabbs)
    m = max(n - 1, 1)
    nodes = np.zeros((m, 3), np.uint32)
    boxes = np.zeros((m, 6), np.float32)
    lib().rc_lbvh_from_sorted(abi.u64ptr(sc), abi.u32ptr(lo), abi.fptr(la), n,
                              abi.u32ptr(nodes), abi.fptr(boxes))
    return nodes, boxes


def lbvh_tree_valid(sorted_codes, leaf_order, strict=False):
    sc = np.ascontiguousarray(sorted_codes, dtype=np.uint64)
    lo = np.ascontiguousarray(leaf_order, dtype=np.uint32)
    return bool(lib().rc_lbvh_tree_valid(abi.u64ptr(sc), abi.u32ptr(lo), len(sc), int(strict)))


class OracleScene:
    """The oracle's copy of a scene with its reference-exact LBVH."""

    def __init__(self, scene):
        self.scene = scene
        self.desc = scene.desc()
        self.h = lib().rc_scene_create(C.byref(self.desc))
        self.n = scene.triangle_count
        self.env_shape = None
        if getattr(scene, "env_map", None) is not None:
            self.set_environment(scene.env_map)

    def set_environment(self, rgb):
        env = np.ascontiguousarray(rgb, dtype=np.float32)
        if lib().rc_scene_set_environment(self.h, env.shape[1], env.shape[0], abi.fptr(env)) != 0:
            raise ValueError("environment rejected (boundary not a Skysphere, bad size or zero luminance)")
        self.env_shape = env.shape[:2]
        self._env = env

    def light_power_table(self):
        n = len(self.scene.light_tris)
        prob, pmf = np.zeros(n, np.float32), np.zeros(n, np.float32)
        alias = np.zeros(n, np.uint32)
        lib().rc_light_power_table(self.h, abi.fptr(prob), abi.u32ptr(alias), abi.fptr(pmf))
        return prob, alias, pmf

    def env_cdfs(self):
        h, w = self.env_shape
        cx, cy = np.zeros((h, w), np.float32), np.zeros(h, np.float32)
        lib().rc_env_cdfs(self.h, abi.fptr(cx), abi.fptr(cy))
        return cx, cy

    def env_sample(self, xi):
        xi = _f(xi).reshape(-1, 2)
        out = np.zeros((len(xi), 3), np.float32)
        lib().rc_env_sample(self.h, abi.fptr(xi), len(xi), abi.fptr(out))
        return out

    def env_pdf_uv(self, uv):
        uv = _f(uv).reshape(-1, 2)
        out = np.zeros(len(uv), np.float32)
        lib().rc_env_pdf_uv(self.h, abi.fptr(uv), len(uv), abi.fptr(out))
        return out

    def env_eval(self, dirs):
        """Solid-angle pdf and radiance of the Skysphere at Y-up directions."""
        d = _f(dirs).reshape(-1, 3)
        pdf = np.zeros(len(d), np.float32)
        rgb = np.zeros((len(d), 3), np.float32)
        lib().rc_env_eval(self.h, abi.fptr(d), len(d), abi.fptr(pdf), abi.fptr(rgb))
        return pdf, rgb

    def close(self):
        if self.h:
            lib().rc_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def lbvh(self):
        n = self.n
        m = max(n - 1, 1)
        codes = np.zeros(n, np.uint64)
        order = np.zeros(n, np.uint32)
        nodes = np.zeros((m, 3), np.uint32)
        boxes = np.zeros((m, 6), np.float32)
        aabb = np.zeros(6, np.float32)
        lib().rc_scene_lbvh(self.h, abi.u64ptr(codes), abi.u32ptr(order), abi.u32ptr(nodes),
                            abi.fptr(boxes), abi.fptr(aabb))
        return codes, order, nodes, boxes, aabb

    def cast_closest(self, rays, brute=False, threads=0, count_visits=False):
        rays = _f(rays).reshape(-1, 8)
        n = rays.shape[0]
        prim = np.zeros(n, np.uint32)
        t = np.zeros(n, np.float32)
        bary = np.zeros((n, 2), np.float32)
        ctr = np.zeros(2, np.uint64) if count_visits else None
        lib().rc_cast_closest(self.h, abi.fptr(rays), n, 1 if brute else 0, abi.u32ptr(prim),
                              abi.fptr(t), abi.fptr(bary),
                              abi.u64ptr(ctr) if ctr is not None else None,
                              threads or os.cpu_count() or 1)
        if count_visits:
            return prim, t, bary, (int(ctr[0]), int(ctr[1]))
        return prim, t, bary

    def cast_visibility(self, rays, threads=0):
        rays = _f(rays).reshape(-1, 8)
        n = rays.shape[0]
        vis = np.zeros(n, np.uint32)
        lib().rc_cast_visibility(self.h, abi.fptr(rays), n, abi.u32ptr(vis),
                                 threads or os.cpu_count() or 1)
        return vis

    def camera_rays(self, params, n):
        rays = np.zeros((n, 8), np.float32)
        w = np.zeros(n, np.float32)
        lib().rc_camera_rays(self.h, C.byref(params), n, abi.fptr(rays), abi.fptr(w))
        return rays, w

    def render(self, params, threads=0, max_iterations=0, path_limit=0):
        film = np.zeros((4, params.height, params.width), np.float32)
        stats = np.zeros(4, np.uint64)
        rc = lib().rc_render(self.h, C.byref(params), abi.fptr(film), threads, max_iterations,
                             path_limit, abi.u64ptr(stats))
        if rc != 0:
            raise RuntimeError("rc_render failed (spectral tables not loaded?)")
        return film, {"dead_paths": int(stats[0]), "ext_rays": int(stats[1]),
                      "shadow_rays": int(stats[2]), "iterations": int(stats[3])}

    def render_tiled(self, params, passes, threads=0):
        """Tiled render (DoLatencyRender): one call per entry of `passes`, each
        that many samples of the current tile.  Returns the film accumulated
        over every call, the calls' tiles (index, x0, y0, x1, y1, tile spp),
        (tile size, tile count) and the statistics."""
        film = np.zeros((4, params.height, params.width), np.float32)
        calls = np.ascontiguousarray(passes, np.uint32)
        tiles = np.zeros((len(calls), 6), np.uint32)
        info = np.zeros(4, np.uint32)
        stats = np.zeros(4, np.uint64)
        rc = lib().rc_render_tiled(self.h, C.byref(params), calls.ctypes.data_as(C.POINTER(C.c_uint32)),
                                   len(calls), abi.fptr(film), threads,
                                   tiles.ctypes.data_as(C.POINTER(C.c_uint32)),
                                   info.ctypes.data_as(C.POINTER(C.c_uint32)), abi.u64ptr(stats))
        if rc != 0:
            raise RuntimeError("rc_render_tiled failed (spectral tables not loaded?)")
        return film, tiles, (tuple(int(v) for v in info[:2]), tuple(int(v) for v in info[2:])), {
            "dead_paths": int(stats[0]), "ext_rays": int(stats[1]),
            "shadow_rays": int(stats[2]), "iterations": int(stats[3])}

    def light_probe(self, point, rn):
        out = np.zeros(8, np.float32)
        lib().rc_light_probe(self.h, abi.fptr(_f(point)), abi.fptr(_f(rn)), abi.fptr(out))
        return out

    def light_pdf(self, li, a, b, ray_pos, ray_dir):
        return float(lib().rc_light_pdf(self.h, li, a, b, abi.fptr(_f(ray_pos)),
                                        abi.fptr(_f(ray_dir))))


def coocta_to_dir(uv):
    uv = _f(uv).reshape(-1, 2)
    out = np.zeros((len(uv), 3), np.float32)
    lib().rc_coocta_to_dir(abi.fptr(uv), len(uv), abi.fptr(out))
    return out


def dir_to_coocta(dirs):
    d = _f(dirs).reshape(-1, 3)
    out = np.zeros((len(d), 2), np.float32)
    lib().rc_dir_to_coocta(abi.fptr(d), len(d), abi.fptr(out))
    return out


def render_batch(scenes, views, params, threads=0):
    """Batch render over OracleScenes (views as for mray_batch_create);
    returns one [4][H][W] film per view and the stats."""
    arr = abi.batch_views(views)
    films = [np.zeros((4, v["height"], v["width"]), np.float32) for v in views]
    fp = (C.c_void_p * len(films))(*[f.ctypes.data for f in films])
    sp = (C.c_void_p * len(scenes))(*[s.h for s in scenes])
    stats = np.zeros(4, np.uint64)
    rc = lib().rc_render_batch(sp, len(scenes), arr, len(views), C.byref(params), fp, threads,
                               abi.u64ptr(stats))
    if rc != 0:
        raise RuntimeError(f"rc_render_batch failed ({rc})")
    return films, {"dead_paths": int(stats[0]), "ext_rays": int(stats[1]),
                   "shadow_rays": int(stats[2]), "iterations": int(stats[3])}


def seed_states(seed, count):
    out = np.zeros(count, np.uint32)
    lib().rc_seed_states(seed, count, abi.u32ptr(out))
    return out


def pcg_sequence(state, count):
    s = np.array([state], dtype=np.uint32)
    return [int(lib().rc_pcg_next(abi.u32ptr(s))) for _ in range(count)]


def to_float01(v):
    return float(lib().rc_to_float01(v))


def _material(mat):
    m = abi.MRayMaterial()
    m.type = mat["type"]
    for k in range(3):
        m.albedo[k] = mat.get("albedo", (0, 0, 0))[k]
        m.cauchyFront[k] = mat.get("cauchyFront", (1, 0, 0))[k]
        m.cauchyBack[k] = mat.get("cauchyBack", (1, 0, 0))[k]
    m.roughness = mat.get("roughness", 0.0)
    m.metallic = mat.get("metallic", 0.0)
    m.specular = mat.get("specular", 0.0)
    return m


def bsdf_eval_many(mat, tbn, geo_n, back_side, wo, wi):
    """f*cos [n][4] and pdf [n] of one material at one wo over wi [n][3]."""
    m = _material(mat)
    wi = _f(wi).reshape(-1, 3)
    out = np.zeros((len(wi), 4), np.float32)
    pdf = np.zeros(len(wi), np.float32)
    lib().rc_bsdf_eval_many(C.byref(m), abi.fptr(_f(tbn)), abi.fptr(_f(geo_n)), int(back_side),
                            abi.fptr(_f(wo)), abi.fptr(wi), len(wi), abi.fptr(out), abi.fptr(pdf))
    return out, pdf


FMATH = {"sin": 0, "cos": 1, "exp": 2, "log": 3, "atan2": 4, "acos": 5, "asin": 6,
         "atanh": 7, "cosh": 8}


def fmath(name, x, y=None):
    """The shared elementary function `name` (fmath.h) at float32 points x
    (atan2 takes (x, y) as (y-coordinate, x-coordinate))."""
    x = _f(x).ravel()
    yy = _f(y).ravel() if y is not None else x
    out = np.zeros_like(x)
    lib().rc_fmath(FMATH[name], abi.fptr(x), abi.fptr(yy), len(x), abi.fptr(out))
    return out


def inv_erf(x):
    x = _f(x).ravel()
    out = np.zeros_like(x)
    lib().rc_inv_erf(abi.fptr(x), len(x), abi.fptr(out))
    return out


def sample_gaussian(xi, sigma):
    xi = _f(xi).ravel()
    v = np.zeros_like(xi)
    p = np.zeros_like(xi)
    lib().rc_sample_gaussian(abi.fptr(xi), len(xi), float(sigma), abi.fptr(v), abi.fptr(p))
    return v, p


def spherical_to_dir(uv):
    """Latitude-longitude uv -> (Z-up directions, sin(inclination))."""
    uv = _f(uv).reshape(-1, 2)
    d = np.zeros((len(uv), 3), np.float32)
    s = np.zeros(len(uv), np.float32)
    lib().rc_spherical_to_dir(abi.fptr(uv), len(uv), abi.fptr(d), abi.fptr(s))
    return d, s


def dir_to_spherical(dirs):
    d = _f(dirs).reshape(-1, 3)
    out = np.zeros((len(d), 2), np.float32)
    lib().rc_dir_to_spherical(abi.fptr(d), len(d), abi.fptr(out))
    return out


def sky_solid_angle_pdf(boundary_type, uv_pdf, sin_incl):
    return float(lib().rc_sky_solid_angle_pdf(boundary_type, uv_pdf, sin_incl))


def bsdf_probe(mat, tbn, geo_n, back_side, wo, wi=None, rn=None, op=0):
    m = abi.MRayMaterial()
    m.type = mat["type"]
    for k in range(3):
        m.albedo[k] = mat.get("albedo", (0, 0, 0))[k]
        m.cauchyFront[k] = mat.get("cauchyFront", (1, 0, 0))[k]
        m.cauchyBack[k] = mat.get("cauchyBack", (1, 0, 0))[k]
    m.roughness = mat.get("roughness", 0.0)
    m.metallic = mat.get("metallic", 0.0)
    m.specular = mat.get("specular", 0.0)
    out = np.zeros(8, np.float32)
    lib().rc_bsdf_probe(C.byref(m), abi.fptr(_f(tbn)), abi.fptr(_f(geo_n)), int(back_side),
                        abi.fptr(_f(wo)), abi.fptr(_f(wi if wi is not None else (0, 0, 1))),
                        abi.fptr(_f(rn if rn is not None else (0, 0, 0, 0))), op, abi.fptr(out))
    return out


def spectral_load(tables):
    """Load the RGB->spectrum coefficient LUT (tables["lut"], as returned by
    mray_amd.native.spectral_tables); the oracle builds its own observer,
    illuminant and XYZ->RGB matrix from the CIE 1931 / D65 data."""
    lut = _f(tables["lut"])
    lib().rc_spectral_load(int(tables["N"]), abi.fptr(lut))
    _keep.append(lut)


def spectral_colour_tables():
    """The oracle's observer [471, 3], illuminant [471] and XYZ->RGB [3, 3]."""
    cmf = np.zeros((471, 3), np.float32)
    illum = np.zeros(471, np.float32)
    m = np.zeros(9, np.float32)
    lib().rc_spectral_colour_tables(abi.fptr(cmf), abi.fptr(illum), abi.fptr(m))
    return cmf, illum, m.reshape(3, 3)


_keep = []


def sample_wavelengths(xi):
    xi = _f(xi)
    n = xi.shape[0]
    w = np.zeros((n, 4), np.float32)
    p = np.zeros((n, 4), np.float32)
    lib().rc_sample_wavelengths(abi.fptr(xi), n, abi.fptr(w), abi.fptr(p))
    return w, p


def rgb_to_spectrum(rgb, waves, radiance=False):
    rgb = _f(rgb).reshape(-1, 3)
    waves = _f(waves).reshape(-1, 4)
    out = np.zeros((rgb.shape[0], 4), np.float32)
    lib().rc_rgb_to_spectrum(abi.fptr(rgb), abi.fptr(waves), rgb.shape[0], int(radiance),
                             abi.fptr(out))
    return out


def spectrum_to_rgb(value, waves, pdfs):
    value = _f(value).reshape(-1, 4)
    out = np.zeros((value.shape[0], 3), np.float32)
    lib().rc_spectrum_to_rgb(abi.fptr(value), abi.fptr(_f(waves)), abi.fptr(_f(pdfs)),
                             value.shape[0], abi.fptr(out))
    return out
