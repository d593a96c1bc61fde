"""Tiled rendering on the GPU (MRAY_RENDER_TILED: ImageTiler +
DoLatencyRender, Tracer/RenderImage.cpp:20-136,
TracerDLL/PathTracerRenderer.cu:1076-1180) against the oracle's tiled
schedule on the same calls.

* Cornell box, 101 x 61 in 2 x 2 tiles of 51 x 31 (edge tiles 50 and 30
  wide), RGB, NEE+MIS: eight calls, a round of 3 then a round of 5 samples
  per tile (the second round's passes are longer than the edge tiles, so
  slots past an edge tile run with the states they kept).  Each call's tile,
  the path counts and every tile section (taken with
  mray_renderer_take_tile) match the oracle's; the sections assembled equal
  the oracle's film up to the order of float atomics.
* Spectral terrain, 96 x 64 in tiles of about 1500 pixels, render_all.
* One tile: the tiled renderer is the untiled one (same film, same counts).
"""
import numpy as np
import pytest

from mray_amd import abi, scenes
from oracle import refcpu
from tests.conftest import film_image, relmse

pytestmark = pytest.mark.gpu


def _gpu_calls(acc, params, passes):
    from mray_amd import native
    r = native.Renderer(acc, params)
    W, H = params.width, params.height
    film = np.zeros((4, H, W), np.float32)
    tiles, ext = [], 0
    for n in passes:
        t, st = r.render_tile(n)
        sec = r.take_tile(t)
        x0, y0 = t["min"]
        x1, y1 = t["max"]
        film[:, y0:y1, x0:x1] += sec
        tiles.append([t["index"], x0, y0, x1, y1, t["spp"]])
        ext += st["ext_rays"]
    info = r.tiling()
    rest = r.film()
    r.close()
    # every section was taken: nothing is left in the device film
    assert not rest.any()
    return film, tiles, info, ext


def test_tiles_match_oracle(torch_gpu):
    from mray_amd import native
    W, H = 101, 61
    scene = scenes.cornell(aspect=W / H)
    p = abi.render_params(W, H, 8, rr=(3, 8), mode=abi.SAMPLE_NEE_MIS, seed=3,
                          max_ray_count=1000, flags=abi.RENDER_TILED)
    passes = [3, 3, 3, 3, 5, 5, 5, 5]
    acc = native.Accel(scene)
    film, tiles, info, ext = _gpu_calls(acc, p, passes)
    acc.close()
    o = refcpu.OracleScene(scene)
    ref, rtiles, rinfo, rst = o.render_tiled(p, passes, threads=8)
    o.close()
    print(f"tiles {info} gpu ext {ext} oracle {rst}", flush=True)
    assert info == rinfo == ((51, 31), (2, 2))
    assert tiles == rtiles.tolist()
    assert ext == rst["ext_rays"]
    assert np.allclose(film[3], 8.0, rtol=1e-5)
    e = relmse(film_image(film), film_image(ref))
    print(f"relMSE vs oracle {e:.3e}", flush=True)
    assert e <= 1e-10


def test_tiled_spectral_render_all(torch_gpu):
    from mray_amd import native
    W, H = 96, 64
    scene = scenes.terrain_spheres(grid=64, spheres=40, width=W, height=H)
    p = abi.render_params(W, H, 2, rr=(10, 10), mode=abi.SAMPLE_NEE_MIS, spectral=True, seed=9,
                          max_ray_count=1500, flags=abi.RENDER_TILED)
    acc = native.Accel(scene)
    r = native.Renderer(acc, p)
    st = r.render_all()
    size, count = r.tiling()
    film = r.film()
    r.close()
    acc.close()
    refcpu.spectral_load(native.spectral_tables())
    o = refcpu.OracleScene(scene)
    ref, _, rinfo, rst = o.render_tiled(p, [2] * (count[0] * count[1]), threads=8)
    o.close()
    print(f"tiles {size} x {count}: gpu {st} oracle {rst}", flush=True)
    assert (size, count) == rinfo and count[0] * count[1] > 1
    assert st["dead_paths"] == rst["dead_paths"] == W * H * 2
    assert st["ext_rays"] == rst["ext_rays"]
    assert np.isfinite(film).all()
    e = relmse(film_image(film), film_image(ref))
    print(f"relMSE vs oracle {e:.3e}", flush=True)
    assert e <= 1e-10


def test_one_tile_is_untiled(torch_gpu):
    from mray_amd import native
    scene = scenes.cornell()
    acc = native.Accel(scene)
    out = []
    for flags in (0, abi.RENDER_TILED):
        r = native.Renderer(acc, abi.render_params(64, 64, 4, rr=(3, 8), seed=1, flags=flags))
        st = r.render_all()
        out.append((r.film(), st))
        r.close()
    acc.close()
    (fa, sa), (fb, sb) = out
    assert sa["dead_paths"] == sb["dead_paths"] and sa["ext_rays"] == sb["ext_rays"]
    assert relmse(film_image(fb), film_image(fa)) <= 1e-10
