"""Edge cases of the one-scene renderer against the oracle at equal samples:
a 1x1 image, a ragged image with fewer path slots than pixels, a scene of
one triangle (a tree whose root holds a single leaf) under a Skysphere, and
a path depth limit of 1 (camera hits and NEE only)."""
import math

import numpy as np
import pytest

from mray_amd import abi, scenes
from oracle import refcpu
from tests.conftest import film_image, relmse

pytestmark = pytest.mark.gpu


def _one_triangle():
    b = scenes._Builder()
    b.add_mesh(np.array([[-1.0, 0.0, -1.0], [1.0, 0.0, -1.0], [0.0, 0.0, 1.0]]),
               np.array([[0, 2, 1]]), key=0, cull=0)
    pos, tbn, idx, key, cull = b.finish()
    cam = {"position": (0.0, 2.0, 2.0), "gaze": (0.0, 0.0, 0.0), "up": (0.0, 1.0, 0.0),
           "fov": (math.radians(60.0),) * 2, "nearFar": (0.01, 100.0)}
    return scenes.Scene(positions=pos, tbn=tbn, indices=idx, surface_keys=key, cull=cull,
                        materials=[{"type": abi.MAT_LAMBERT, "albedo": (0.7, 0.5, 0.3)}],
                        light_tris=np.zeros(0, np.uint32), light_radiance=np.zeros((0, 3), np.float32),
                        light_two_sided=np.zeros(0, np.uint8),
                        boundary_type=abi.BOUNDARY_SKYSPHERE_COOCTA,
                        boundary_radiance=(0.6, 0.7, 0.9), camera=cam)


def _check(scene, params, bound=1e-3):
    from mray_amd import native
    acc = native.Accel(scene)
    r = native.Renderer(acc, params)
    st = r.render_all()
    film = r.film()
    r.close()
    acc.close()
    ofilm, ost = refcpu.OracleScene(scene).render(params)
    assert st["dead_paths"] == ost["dead_paths"] == params.width * params.height * params.totalSPP
    assert np.allclose(film[3], ofilm[3], rtol=1e-5)
    e = relmse(film_image(film), film_image(ofilm))
    print(f"{params.width}x{params.height}x{params.totalSPP}: relMSE={e:.3e}", flush=True)
    assert e <= bound
    return film


@pytest.mark.parametrize("spectral", [False, True])
def test_one_pixel(torch_gpu, spectral):
    from mray_amd import native
    if spectral:
        refcpu.spectral_load(native.spectral_tables())
    _check(scenes.cornell(), abi.render_params(1, 1, 7, rr=(3, 8), seed=2, spectral=spectral))


def test_ragged_image_few_slots(torch_gpu):
    _check(scenes.cornell(), abi.render_params(7, 3, 5, rr=(3, 8), seed=4, max_ray_count=5))


def test_one_triangle_scene(torch_gpu):
    from mray_amd import native
    sc = _one_triangle()
    acc = native.Accel(sc)
    assert acc.info()["triangles"] == 1
    acc.close()
    film = _check(sc, abi.render_params(24, 16, 4, rr=(3, 8), seed=6))
    assert np.all(film[3] > 0)


def test_depth_limit_one(torch_gpu):
    _check(scenes.cornell(), abi.render_params(16, 16, 4, rr=(1, 1), seed=9))
