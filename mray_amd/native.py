"""ctypes binding of libmray_amd.so (the gfx950 hot path, include/mray_amd.h).

Device buffers are torch tensors on `cuda:0` (PyTorch is only the allocator
and stream plumbing here); every compute call goes through the C-ABI.  There is
no CPU fallback: a missing library or GPU raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmray_amd.so")
_lib = None


class MRayError(RuntimeError):
    """Raised when a C-ABI call fails (the reference throws MRayError)."""


# Every symbol include/mray_amd.h declares (checked by tests/test_abi.py).
EXPORTS = [
    "mray_last_error", "mray_version", "mray_device_count",
    "mray_prim_centers_aabbs", "mray_reduce_aabb", "mray_morton63", "mray_morton30",
    "mray_radix_sort_temp_bytes", "mray_radix_sort_u64_u32", "mray_lbvh_build",
    "mray_lbvh_temp_bytes", "mray_radix_sort_u64_u32_noalloc",
    "mray_accel_create", "mray_accel_create_ex", "mray_accel_destroy", "mray_accel_info", "mray_accel_read_lbvh",
    "mray_accel_cast_closest", "mray_accel_cast_closest_counted", "mray_accel_cast_visibility",
    "mray_renderer_create", "mray_renderer_destroy", "mray_renderer_iterate",
    "mray_renderer_render_all", "mray_renderer_stats", "mray_renderer_read_film",
    "mray_renderer_reset", "mray_renderer_film_device", "mray_spectral_tables",
    "mray_renderer_set_measurement", "mray_renderer_phase_times",
    "mray_nrc_create", "mray_nrc_destroy", "mray_nrc_infer", "mray_nrc_backward", "mray_nrc_adam",
    "mray_nrc_train", "mray_nrc_last_loss", "mray_nrc_param_counts", "mray_nrc_read",
    "mray_nrc_write", "mray_nrc_zero_grads", "mray_nrc_set_profiling", "mray_nrc_kernel_times",
    "mray_renderer_attach_nrc", "mray_renderer_attach_nrcs", "mray_renderer_render_frame",
    "mray_renderer_cache_stats",
    "mray_device_memory",
    "mray_batch_scene_create", "mray_batch_scene_destroy", "mray_batch_scene_info",
    "mray_batch_create", "mray_batch_read_film", "mray_batch_view_times", "mray_batch_autotune",
    "mray_batch_set_trace_tpb",
    "mray_spectral_write_mrspectra", "mray_spectral_read_mrspectra", "mray_spectral_use_mrspectra",
    "mray_renderer_take_film", "mray_film_write_exr",
    "mray_instanced_scene_create", "mray_instanced_scene_destroy", "mray_instanced_scene_info",
    "mray_renderer_create_instanced", "mray_instanced_batch_create",
    "mray_accel_set_environment", "mray_accel_environment_cdfs",
    "mray_instanced_scene_set_environment", "mray_renderer_light_table",
    "mray_renderer_tiling", "mray_renderer_render_tile", "mray_renderer_take_tile",
    "mray_accel_read_wide",
]


def lib():
    """Load libmray_amd.so (built by __graft_entry__.build()); raise if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MRayError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
    L = C.CDLL(LIB_PATH)
    P = C.c_void_p
    u32 = C.c_uint32
    L.mray_last_error.restype = C.c_char_p
    L.mray_version.restype = C.c_char_p
    L.mray_device_count.argtypes = [C.POINTER(C.c_int)]
    L.mray_device_memory.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    L.mray_prim_centers_aabbs.argtypes = [P, P, u32, P, P, P]
    L.mray_reduce_aabb.argtypes = [P, u32, P, P]
    L.mray_morton63.argtypes = [P, u32, P, P, P]
    L.mray_morton30.argtypes = [P, u32, P, P, P]
    L.mray_radix_sort_temp_bytes.restype = C.c_size_t
    L.mray_radix_sort_temp_bytes.argtypes = [u32]
    L.mray_radix_sort_u64_u32.argtypes = [P, P, u32, P, C.c_size_t, P]
    L.mray_radix_sort_u64_u32_noalloc.argtypes = [P, P, u32, P]
    L.mray_lbvh_temp_bytes.restype = C.c_size_t
    L.mray_lbvh_temp_bytes.argtypes = [u32]
    L.mray_lbvh_build.argtypes = [P, P, P, u32, P, P, P, C.c_size_t, P]
    L.mray_accel_create.argtypes = [C.POINTER(abi.MRaySceneDesc), C.POINTER(P)]
    L.mray_accel_create_ex.argtypes = [C.POINTER(abi.MRaySceneDesc), C.POINTER(abi.MRayAccelOptions),
                                       C.POINTER(P)]
    L.mray_accel_destroy.argtypes = [P]
    L.mray_accel_info.argtypes = [P, C.POINTER(u32), C.POINTER(u32), C.POINTER(u32)]
    L.mray_accel_read_lbvh.argtypes = [P, P, P, P, P]
    L.mray_accel_cast_closest.argtypes = [P, P, u32, P, P, P, P]
    L.mray_accel_cast_closest_counted.argtypes = [P, P, u32, P, P, P, P, P]
    L.mray_accel_cast_visibility.argtypes = [P, P, u32, P, P]
    L.mray_renderer_create.argtypes = [P, C.POINTER(abi.MRaySceneDesc),
                                       C.POINTER(abi.MRayRenderParams), C.POINTER(P)]
    L.mray_renderer_destroy.argtypes = [P]
    L.mray_renderer_iterate.argtypes = [P, u32, P]
    L.mray_renderer_render_all.argtypes = [P, C.POINTER(abi.MRayRenderStats)]
    L.mray_renderer_stats.argtypes = [P, C.POINTER(abi.MRayRenderStats)]
    L.mray_renderer_read_film.argtypes = [P, P]
    L.mray_renderer_take_film.argtypes = [P, P]
    L.mray_renderer_tiling.argtypes = [P, C.POINTER(abi.MRayTileInfo)]
    L.mray_renderer_render_tile.argtypes = [P, u32, C.POINTER(abi.MRayTileResult),
                                            C.POINTER(abi.MRayRenderStats)]
    L.mray_renderer_take_tile.argtypes = [P, P]
    L.mray_accel_read_wide.argtypes = [P, P, P]
    L.mray_film_write_exr.argtypes = [P, u32, u32, C.c_char_p]
    L.mray_renderer_reset.argtypes = [P]
    L.mray_renderer_film_device.argtypes = [P, C.POINTER(P)]
    L.mray_spectral_tables.argtypes = [C.POINTER(C.c_int), P, P, P, P]
    L.mray_spectral_write_mrspectra.argtypes = [C.c_char_p]
    L.mray_spectral_read_mrspectra.argtypes = [C.c_char_p, C.POINTER(C.c_int), P]
    L.mray_spectral_use_mrspectra.argtypes = [C.c_char_p]
    L.mray_nrc_create.argtypes = [C.POINTER(abi.MRayNrcConfig), C.POINTER(P)]
    L.mray_nrc_destroy.argtypes = [P]
    L.mray_nrc_infer.argtypes = [P, P, u32, P, P, P]
    L.mray_nrc_backward.argtypes = [P, P, P, u32, P]
    L.mray_nrc_adam.argtypes = [P, P]
    L.mray_nrc_train.argtypes = [P, P, P, u32, P]
    L.mray_nrc_last_loss.argtypes = [P, C.POINTER(C.c_float)]
    L.mray_nrc_param_counts.argtypes = [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.mray_nrc_read.argtypes = [P, u32, P]
    L.mray_nrc_write.argtypes = [P, u32, P]
    L.mray_nrc_zero_grads.argtypes = [P, P]
    L.mray_nrc_set_profiling.argtypes = [P, u32]
    L.mray_nrc_kernel_times.argtypes = [P, P, P]
    L.mray_renderer_attach_nrc.argtypes = [P, P, C.POINTER(abi.MRayNrcRenderParams)]
    L.mray_renderer_attach_nrcs.argtypes = [P, P, u32, P]
    L.mray_renderer_render_frame.argtypes = [P, u32, C.POINTER(abi.MRayRenderStats)]
    L.mray_renderer_cache_stats.argtypes = [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.mray_renderer_set_measurement.argtypes = [P, u32, u32]
    L.mray_renderer_phase_times.argtypes = [P, P, P]
    L.mray_batch_scene_create.argtypes = [P, u32, C.POINTER(P)]
    L.mray_batch_scene_destroy.argtypes = [P]
    L.mray_batch_scene_info.argtypes = [P, C.POINTER(u32), C.POINTER(u32), C.POINTER(u32),
                                        C.POINTER(u32)]
    L.mray_batch_create.argtypes = [P, P, u32, C.POINTER(abi.MRayRenderParams), C.POINTER(P)]
    L.mray_batch_read_film.argtypes = [P, u32, P]
    L.mray_batch_view_times.argtypes = [P, P]
    L.mray_batch_set_trace_tpb.argtypes = [P, u32]
    L.mray_batch_autotune.argtypes = [P, P, P, u32, C.POINTER(abi.MRayRenderParams),
                                      C.POINTER(abi.MRayBatchTune), P, u32, C.POINTER(u32)]
    L.mray_instanced_scene_create.argtypes = [P, u32, P, u32, C.POINTER(abi.MRayInstEnv), C.POINTER(P)]
    L.mray_instanced_scene_destroy.argtypes = [P]
    L.mray_instanced_scene_info.argtypes = [P, C.POINTER(u32), C.POINTER(C.c_uint64), C.POINTER(u32),
                                            C.POINTER(u32)]
    L.mray_renderer_create_instanced.argtypes = [P, C.POINTER(abi.MRayRenderParams), C.POINTER(P)]
    L.mray_instanced_batch_create.argtypes = [P, P, u32, C.POINTER(abi.MRayRenderParams), C.POINTER(P)]
    L.mray_accel_set_environment.argtypes = [P, u32, u32, P]
    L.mray_accel_environment_cdfs.argtypes = [P, P, P]
    L.mray_instanced_scene_set_environment.argtypes = [P, u32, u32, P]
    L.mray_renderer_light_table.argtypes = [P, P, P, P]
    _lib = L
    return L


def check(rc):
    if rc != abi.MRAY_OK:
        msg = lib().mray_last_error().decode(errors="replace")
        raise MRayError(f"mray call failed ({rc}): {msg}")


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def require_gpu():
    import torch
    if not torch.cuda.is_available():
        raise MRayError("no GPU visible: the gfx950 path has no CPU fallback")
    return torch


# ---------------------------------------------------------------------------
# Stage-level entry points (device tensors in, device tensors out)
# ---------------------------------------------------------------------------
def prim_centers_aabbs(d_pos, d_idx):
    torch = require_gpu()
    n = d_idx.shape[0]
    centers = torch.empty((n, 3), dtype=torch.float32, device=d_pos.device)
    aabbs = torch.empty((n, 6), dtype=torch.float32, device=d_pos.device)
    check(lib().mray_prim_centers_aabbs(_ptr(d_pos), _ptr(d_idx), n, _ptr(centers), _ptr(aabbs),
                                        _stream()))
    return centers, aabbs


def reduce_aabb(d_aabbs):
    torch = require_gpu()
    out = torch.empty(6, dtype=torch.float32, device=d_aabbs.device)
    check(lib().mray_reduce_aabb(_ptr(d_aabbs), d_aabbs.shape[0], _ptr(out), _stream()))
    return out


def morton63(d_centers, d_box):
    torch = require_gpu()
    n = d_centers.shape[0]
    out = torch.empty(n, dtype=torch.int64, device=d_centers.device)
    check(lib().mray_morton63(_ptr(d_centers), n, _ptr(d_box), _ptr(out), _stream()))
    return out


def morton30(d_centers, d_box):
    torch = require_gpu()
    n = d_centers.shape[0]
    out = torch.empty(n, dtype=torch.int32, device=d_centers.device)
    check(lib().mray_morton30(_ptr(d_centers), n, _ptr(d_box), _ptr(out), _stream()))
    return out


def radix_sort_u64_u32(d_keys, d_vals):
    """In-place stable ascending sort; keys int64 tensor (bit pattern u64)."""
    torch = require_gpu()
    n = d_keys.shape[0]
    nb = lib().mray_radix_sort_temp_bytes(n)
    tmp = torch.empty(max(1, nb), dtype=torch.uint8, device=d_keys.device)
    check(lib().mray_radix_sort_u64_u32(_ptr(d_keys), _ptr(d_vals), n, _ptr(tmp), nb, _stream()))


NOALLOC_SORT_MAX = 1 << 20     # MRAY_NOALLOC_SORT_MAX


def radix_sort_u64_u32_noalloc(d_keys, d_vals):
    """The same sort with the library's static scratch: no allocation at all
    (at most NOALLOC_SORT_MAX pairs; calls on one GPU must not overlap)."""
    require_gpu()
    n = d_keys.shape[0]
    check(lib().mray_radix_sort_u64_u32_noalloc(_ptr(d_keys), _ptr(d_vals), n, _stream()))


def lbvh_build(d_sorted_codes, d_leaf_order, d_leaf_aabbs):
    torch = require_gpu()
    n = d_sorted_codes.shape[0]
    m = max(n - 1, 1)
    nodes = torch.empty((m, 3), dtype=torch.int32, device=d_sorted_codes.device)
    boxes = torch.empty((m, 6), dtype=torch.float32, device=d_sorted_codes.device)
    nb = lib().mray_lbvh_temp_bytes(n)
    tmp = torch.empty(max(1, nb), dtype=torch.uint8, device=d_sorted_codes.device)
    check(lib().mray_lbvh_build(_ptr(d_sorted_codes), _ptr(d_leaf_order), _ptr(d_leaf_aabbs), n,
                                _ptr(nodes), _ptr(boxes), _ptr(tmp), nb, _stream()))
    return nodes, boxes


def write_mrspectra(path):
    """The library's RGB->spectrum coefficients as a reference .mrspectra file."""
    check(lib().mray_spectral_write_mrspectra(str(path).encode()))


def read_mrspectra(path):
    """Validated .mrspectra contents as lut [3][N][N][N][4] (c0, c1, c2, 0)."""
    n = C.c_int()
    check(lib().mray_spectral_read_mrspectra(str(path).encode(), C.byref(n), None))
    lut = np.zeros((3, n.value, n.value, n.value, 4), np.float32)
    check(lib().mray_spectral_read_mrspectra(str(path).encode(), None, lut.ctypes.data))
    return lut


def write_exr(film, path):
    """Film [4][H][W] (sums + weights) as the OpenEXR image of per-pixel means."""
    film = np.ascontiguousarray(film, np.float32)
    _, h, w = film.shape
    check(lib().mray_film_write_exr(film.ctypes.data, w, h, str(path).encode()))


def spectral_tables():
    """Host copies of the RGB->spectrum tables (built on first use)."""
    n = C.c_int()
    check(lib().mray_spectral_tables(C.byref(n), None, None, None, None))
    N = n.value
    lut = np.zeros((3, N, N, N, 4), np.float32)
    cmf = np.zeros((471, 3), np.float32)
    illum = np.zeros(471, np.float32)
    m = np.zeros(9, np.float32)
    check(lib().mray_spectral_tables(C.byref(n), lut.ctypes.data, cmf.ctypes.data,
                                     illum.ctypes.data, m.ctypes.data))
    return {"N": N, "lut": lut, "cmf": cmf, "illum": illum, "xyz_to_rgb": m}


# ---------------------------------------------------------------------------
# Accelerator and renderer objects
# ---------------------------------------------------------------------------
class Accel:
    """Device scene + LBVH + 4-wide BVH (MRayAccel)."""

    def __init__(self, scene, builder="sah", width=4, host_collapse=False):
        require_gpu()
        self.scene = scene
        self.desc = scene.desc()
        h = C.c_void_p()
        opt = abi.MRayAccelOptions()
        opt.builder = {"sah": abi.BUILDER_SAH, "lbvh": abi.BUILDER_LBVH}[builder]
        opt.width = width
        opt.flags = abi.ACCEL_HOST_COLLAPSE if host_collapse else 0
        check(lib().mray_accel_create_ex(C.byref(self.desc), C.byref(opt), C.byref(h)))
        self.h = h
        self.n = scene.triangle_count
        if getattr(scene, "env_map", None) is not None:
            self.set_environment(scene.env_map)

    def set_environment(self, rgb):
        """Textured Skysphere: rgb [H, W, 3] over concentric-octahedral uv."""
        env = np.ascontiguousarray(rgb, dtype=np.float32)
        check(lib().mray_accel_set_environment(self.h, env.shape[1], env.shape[0], env.ctypes.data))

    def environment_cdfs(self, height, width):
        cx = np.zeros((height, width), np.float32)
        cy = np.zeros(height, np.float32)
        check(lib().mray_accel_environment_cdfs(self.h, cx.ctypes.data, cy.ctypes.data))
        return cx, cy

    def close(self):
        if getattr(self, "h", None):
            lib().mray_accel_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def read_wide(self):
        """The 4-wide tree: (nodes [wide][32] uint32 words, soup triangle ids)."""
        wide = self.info()["wide_nodes"]
        nodes = np.zeros((wide, 32), np.uint32)
        prims = np.zeros(self.n, np.uint32)
        check(lib().mray_accel_read_wide(self.h, nodes.ctypes.data, prims.ctypes.data))
        return nodes, prims

    def info(self):
        a, b, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
        check(lib().mray_accel_info(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return {"binary_nodes": a.value, "wide_nodes": b.value, "triangles": c.value}

    def read_lbvh(self):
        n = self.n
        m = max(n - 1, 1)
        codes = np.zeros(n, np.uint64)
        order = np.zeros(n, np.uint32)
        nodes = np.zeros((m, 3), np.uint32)
        boxes = np.zeros((m, 6), np.float32)
        check(lib().mray_accel_read_lbvh(self.h, codes.ctypes.data, order.ctypes.data,
                                         nodes.ctypes.data, boxes.ctypes.data))
        return codes, order, nodes, boxes

    def cast_closest(self, d_rays, count_visits=False):
        torch = require_gpu()
        n = d_rays.shape[0]
        dev = d_rays.device
        prim = torch.empty(n, dtype=torch.int32, device=dev)
        t = torch.empty(n, dtype=torch.float32, device=dev)
        bary = torch.empty((n, 2), dtype=torch.float32, device=dev)
        if count_visits:
            ctr = torch.zeros(2, dtype=torch.int64, device=dev)
            check(lib().mray_accel_cast_closest_counted(self.h, _ptr(d_rays), n, _ptr(prim),
                                                        _ptr(t), _ptr(bary), _ptr(ctr),
                                                        _stream()))
            return prim, t, bary, ctr
        check(lib().mray_accel_cast_closest(self.h, _ptr(d_rays), n, _ptr(prim), _ptr(t),
                                            _ptr(bary), _stream()))
        return prim, t, bary

    def cast_visibility(self, d_rays):
        torch = require_gpu()
        n = d_rays.shape[0]
        vis = torch.empty(n, dtype=torch.int32, device=d_rays.device)
        check(lib().mray_accel_cast_visibility(self.h, _ptr(d_rays), n, _ptr(vis), _stream()))
        return vis


class Renderer:
    """Wavefront path tracer (MRayRenderer)."""

    def __init__(self, accel, params, scene=None):
        self.accel = accel
        self.params = params
        self.width, self.height = params.width, params.height
        h = C.c_void_p()
        desc = accel.desc if scene is None else scene.desc()
        check(lib().mray_renderer_create(accel.h, C.byref(desc), C.byref(params), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().mray_renderer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def iterate(self, iterations=1, stream=None):
        s = C.c_void_p(stream) if stream is not None else None
        check(lib().mray_renderer_iterate(self.h, iterations, s))

    def stats(self):
        st = abi.MRayRenderStats()
        check(lib().mray_renderer_stats(self.h, C.byref(st)))
        return {"dead_paths": st.deadPaths, "total_dead_paths": st.totalDeadPaths,
                "ext_rays": st.extensionRays, "shadow_rays": st.shadowRays,
                "iterations": st.iterations, "finished": bool(st.finished),
                "node_visits": st.nodeVisits, "tri_tests": st.triTests}

    def render_all(self):
        st = abi.MRayRenderStats()
        check(lib().mray_renderer_render_all(self.h, C.byref(st)))
        return {"dead_paths": st.deadPaths, "ext_rays": st.extensionRays,
                "shadow_rays": st.shadowRays, "iterations": st.iterations,
                "finished": bool(st.finished), "node_visits": st.nodeVisits,
                "tri_tests": st.triTests, "host_waits": st.hostWaits,
                "device_drains": st.deviceDrains, "shadow_node_visits": st.shadowNodeVisits,
                "shadow_tri_tests": st.shadowTriTests}

    def set_measurement(self, profiling=False, counting=False):
        check(lib().mray_renderer_set_measurement(self.h, int(profiling), int(counting)))

    def phase_times(self):
        """Accumulated ms and launch counts for refill, trace, shade, shadow."""
        ms = np.zeros(4, np.float64)
        n = np.zeros(4, np.uint64)
        check(lib().mray_renderer_phase_times(self.h, ms.ctypes.data, n.ctypes.data))
        names = ("refill", "trace", "shade", "shadow")
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(names)}

    def reset(self):
        check(lib().mray_renderer_reset(self.h))

    def film(self):
        out = np.zeros((4, self.height, self.width), np.float32)
        check(lib().mray_renderer_read_film(self.h, out.ctypes.data))
        return out

    def take_film(self):
        """Sums splatted since the last take; the device film is cleared."""
        out = np.zeros((4, self.height, self.width), np.float32)
        check(lib().mray_renderer_take_film(self.h, out.ctypes.data))
        return out

    def tiling(self):
        """(tile size, tile count) of a tiled renderer (MRAY_RENDER_TILED)."""
        t = abi.MRayTileInfo()
        check(lib().mray_renderer_tiling(self.h, C.byref(t)))
        return tuple(t.tileSize), tuple(t.tileCount)

    def render_tile(self, pass_count=1):
        """`pass_count` more samples of the current tile, then the next tile
        is current; returns the tile rendered and the call's statistics."""
        t = abi.MRayTileResult()
        st = abi.MRayRenderStats()
        check(lib().mray_renderer_render_tile(self.h, pass_count, C.byref(t), C.byref(st)))
        tile = {"index": t.tileIndex, "min": tuple(t.pixelMin), "max": tuple(t.pixelMax),
                "spp": t.tileSPP}
        return tile, {"dead_paths": st.deadPaths, "ext_rays": st.extensionRays,
                      "shadow_rays": st.shadowRays, "iterations": st.iterations,
                      "finished": bool(st.finished), "host_waits": st.hostWaits}

    def take_tile(self, tile):
        """The last tile's sums [4][h][w]; those film pixels are cleared."""
        w = tile["max"][0] - tile["min"][0]
        h = tile["max"][1] - tile["min"][1]
        out = np.zeros((4, h, w), np.float32)
        check(lib().mray_renderer_take_tile(self.h, out.ctypes.data))
        return out

    def light_table(self, n):
        """(prob, alias, pmf) of the n lights' power-sampling tables."""
        prob, pmf = np.zeros(n, np.float32), np.zeros(n, np.float32)
        alias = np.zeros(n, np.uint32)
        check(lib().mray_renderer_light_table(self.h, prob.ctypes.data, alias.ctypes.data,
                                              pmf.ctypes.data))
        return prob, alias, pmf

    @staticmethod
    def _nrc_params(train_shift, spread, batch, steps, min_depth, nee_at_end):
        p = abi.MRayNrcRenderParams()
        p.trainShift, p.spreadThreshold, p.trainBatch, p.trainSteps = train_shift, spread, batch, steps
        p.minDepth = min_depth
        p.flags = abi.NRC_NEE_AT_END if nee_at_end else 0
        return p

    def attach_nrc(self, nrc, train_shift=4, spread=0.01, batch=1 << 14, steps=4, min_depth=1,
                   nee_at_end=False):
        """Terminate paths into `nrc` and train it after every frame;
        nrc=None detaches."""
        p = self._nrc_params(train_shift, spread, batch, steps, min_depth, nee_at_end)
        self.nrc = nrc
        check(lib().mray_renderer_attach_nrc(self.h, nrc.h if nrc is not None else None,
                                             C.byref(p)))

    def attach_nrcs(self, caches, train_shift=4, spread=0.01, batch=1 << 14, steps=4,
                    min_depth=1, nee_at_end=False):
        """One cache per view (a batch renderer's views in order); [] detaches."""
        p = self._nrc_params(train_shift, spread, batch, steps, min_depth, nee_at_end)
        self.caches = list(caches)
        arr = (C.c_void_p * max(1, len(caches)))(*[c.h.value for c in caches])
        check(lib().mray_renderer_attach_nrcs(self.h, arr if caches else None, len(caches),
                                              C.byref(p)))

    def render_frame(self, spp=1):
        st = abi.MRayRenderStats()
        check(lib().mray_renderer_render_frame(self.h, spp, C.byref(st)))
        return {"dead_paths": st.deadPaths, "ext_rays": st.extensionRays,
                "shadow_rays": st.shadowRays, "iterations": st.iterations,
                "host_waits": st.hostWaits, "device_drains": st.deviceDrains}

    def cache_stats(self):
        q, t = C.c_uint64(), C.c_uint64()
        check(lib().mray_renderer_cache_stats(self.h, C.byref(q), C.byref(t)))
        return {"queries": q.value, "training_pairs": t.value}


class Nrc:
    """Neural radiance cache (mray_nrc_*): hash grid + 64-wide MLP on the GPU.
    Records are float32 [n, 11] device tensors (see include/mray_amd.h)."""

    def __init__(self, levels=16, log2_table=19, min_res=16, max_res=2048, max_batch=1 << 16,
                 lr=1e-2, beta1=0.9, beta2=0.99, eps=1e-15, seed=1, lazy_adam=False):
        require_gpu()
        cfg = abi.MRayNrcConfig()
        cfg.flags = abi.NRC_LAZY_ADAM if lazy_adam else 0
        cfg.levels, cfg.log2TableSize = levels, log2_table
        cfg.minResolution, cfg.maxResolution = min_res, max_res
        cfg.maxBatch = max_batch
        cfg.learningRate, cfg.beta1, cfg.beta2, cfg.epsilon = lr, beta1, beta2, eps
        cfg.seed = seed
        self.h = C.c_void_p()
        check(lib().mray_nrc_create(C.byref(cfg), C.byref(self.h)))
        m, g = C.c_uint64(), C.c_uint64()
        check(lib().mray_nrc_param_counts(self.h, C.byref(m), C.byref(g)))
        self.mlp_params, self.grid_params = m.value, g.value
        self.levels, self.log2_table = levels, log2_table

    def close(self):
        if self.h:
            lib().mray_nrc_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def infer(self, d_rec):
        torch = require_gpu()
        n = d_rec.shape[0]
        out = torch.empty((n, 3), dtype=torch.float32, device=d_rec.device)
        check(lib().mray_nrc_infer(self.h, _ptr(d_rec), n, None, _ptr(out), None))
        torch.cuda.synchronize()
        return out

    def backward(self, d_rec, d_target):
        check(lib().mray_nrc_backward(self.h, _ptr(d_rec), _ptr(d_target), d_rec.shape[0], None))

    def adam(self):
        check(lib().mray_nrc_adam(self.h, None))

    def train(self, d_rec, d_target):
        check(lib().mray_nrc_train(self.h, _ptr(d_rec), _ptr(d_target), d_rec.shape[0], None))

    def loss(self):
        v = C.c_float()
        check(lib().mray_nrc_last_loss(self.h, C.byref(v)))
        return v.value

    def zero_grads(self):
        check(lib().mray_nrc_zero_grads(self.h, None))

    NRC_KERNELS = ("infer_forward", "train_forward", "backward", "wgrad", "grid_bwd", "adam")

    def set_profiling(self, on=True):
        check(lib().mray_nrc_set_profiling(self.h, int(bool(on))))

    def kernel_times(self):
        """{kernel: (accumulated ms, launches)} since profiling was turned on."""
        import numpy as np
        ms = np.zeros(6, np.float64)
        n = np.zeros(6, np.uint64)
        check(lib().mray_nrc_kernel_times(self.h, ms.ctypes.data, n.ctypes.data))
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(self.NRC_KERNELS)}

    def read(self, which):
        import numpy as np
        n = self.mlp_params if which in (abi.NRC_WEIGHTS, abi.NRC_WEIGHT_GRADS) else self.grid_params
        out = np.empty(n, np.float32)
        check(lib().mray_nrc_read(self.h, which, abi.fptr(out)))
        if which in (abi.NRC_WEIGHTS, abi.NRC_WEIGHT_GRADS):
            return out.reshape(6, 64, 64)
        return out.reshape(self.levels, 1 << self.log2_table, 2)

    def write(self, which, arr):
        import numpy as np
        a = np.ascontiguousarray(arr, np.float32).reshape(-1)
        check(lib().mray_nrc_write(self.h, which, abi.fptr(a)))


class BatchScene:
    """Scenes of several accelerators merged for one batch (MRayBatchScene)."""

    def __init__(self, accels):
        require_gpu()
        arr = (C.c_void_p * len(accels))(*[a.h.value for a in accels])
        h = C.c_void_p()
        check(lib().mray_batch_scene_create(arr, len(accels), C.byref(h)))
        self.h = h
        self.count = len(accels)

    def close(self):
        if getattr(self, "h", None):
            lib().mray_batch_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        v = [C.c_uint32() for _ in range(4)]
        check(lib().mray_batch_scene_info(self.h, *[C.byref(x) for x in v]))
        return {"scenes": v[0].value, "triangles": v[1].value, "nodes": v[2].value,
                "width": v[3].value}


def _tune_dict(t):
    return {"width": t.width, "queues": t.queues, "slots": t.slots, "trace_tpb": t.traceTPB,
            "mpaths_per_s": t.mpathsPerS}


def batch_autotune(scene4, scene8, views, params, max_trials=64):
    """Autotuner: best setting and every trial (see mray_batch_autotune)."""
    arr = abi.batch_views(views)
    best = abi.MRayBatchTune()
    trials = (abi.MRayBatchTune * max_trials)()
    n = C.c_uint32()
    check(lib().mray_batch_autotune(scene4.h if scene4 is not None else None,
                                    scene8.h if scene8 is not None else None,
                                    arr, len(views), C.byref(params), C.byref(best), trials,
                                    max_trials, C.byref(n)))
    return _tune_dict(best), [_tune_dict(trials[k]) for k in range(min(n.value, max_trials))]


class BatchRenderer(Renderer):
    """Batch of views in one renderer: over a BatchScene (mray_batch_create)
    or over one InstancedScene, every view's scene 0
    (mray_instanced_batch_create)."""

    def __init__(self, bscene, views, params):
        self.bscene = bscene
        self.params = params
        self.views = [dict(v) for v in views]
        arr = abi.batch_views(views)
        h = C.c_void_p()
        if isinstance(bscene, InstancedScene):
            check(lib().mray_instanced_batch_create(bscene.h, arr, len(views), C.byref(params),
                                                    C.byref(h)))
        else:
            check(lib().mray_batch_create(bscene.h, arr, len(views), C.byref(params), C.byref(h)))
        self.h = h

    def film(self, view=None):
        if view is None:
            return [self.film(k) for k in range(len(self.views))]
        v = self.views[view]
        out = np.zeros((4, v["height"], v["width"]), np.float32)
        check(lib().mray_batch_read_film(self.h, view, out.ctypes.data))
        return out

    def view_times(self):
        ms = np.zeros(len(self.views), np.float64)
        check(lib().mray_batch_view_times(self.h, ms.ctypes.data))
        return ms

    def set_trace_tpb(self, tpb):
        check(lib().mray_batch_set_trace_tpb(self.h, tpb))


class InstancedScene:
    """Meshes placed by instances under a tree over instances (MRayInstScene).
    instances: dicts {mesh, rotation (w,x,y,z), scale, translation};
    env: {boundary_type, boundary_radiance, scene_diameter, camera}."""

    def __init__(self, mesh_accels, instances, env):
        require_gpu()
        arr = (C.c_void_p * len(mesh_accels))(*[a.h.value for a in mesh_accels])
        inst = abi.instances_array(instances)
        e = abi.inst_env(env)
        h = C.c_void_p()
        check(lib().mray_instanced_scene_create(arr, len(mesh_accels), inst, len(instances),
                                                C.byref(e), C.byref(h)))
        self.h = h
        if env.get("env_map") is not None:
            rgb = np.ascontiguousarray(env["env_map"], dtype=np.float32)
            check(lib().mray_instanced_scene_set_environment(self.h, rgb.shape[1], rgb.shape[0],
                                                             rgb.ctypes.data))

    def close(self):
        if getattr(self, "h", None):
            lib().mray_instanced_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        n, l, t = C.c_uint32(), C.c_uint32(), C.c_uint32()
        tri = C.c_uint64()
        check(lib().mray_instanced_scene_info(self.h, C.byref(n), C.byref(tri), C.byref(l), C.byref(t)))
        return {"instances": n.value, "effective_triangles": tri.value, "lights": l.value,
                "tlas_nodes": t.value}


class InstancedRenderer(Renderer):
    """Wavefront path tracer over an InstancedScene."""

    def __init__(self, iscene, params):
        self.iscene = iscene
        self.params = params
        self.width, self.height = params.width, params.height
        h = C.c_void_p()
        check(lib().mray_renderer_create_instanced(iscene.h, C.byref(params), C.byref(h)))
        self.h = h
