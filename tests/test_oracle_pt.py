"""Oracle path tracer against closed-form scenes.

* sky only: every camera ray misses, the pixel is the Skysphere radiance
  exactly (LightWorkFunctionWithNEE, camera rays need no MIS);
* an emitter filling the view: the pixel is its radiance exactly;
* a large Lambert plane (albedo a) under a constant sky L: one bounce to the
  sky, E[pixel] = a L -- for Pure, NEE and NEE+MIS alike;
* Cornell: Pure / NEE / NEE+MIS agree in expectation.
"""
import math

import numpy as np
import pytest

from mray_amd import abi, scenes
from oracle import refcpu
from tests.conftest import film_image

SKY = (0.5, 0.25, 0.125)


def _scene(positions, indices, keys, mats, cull, lights=(), light_rad=(), boundary=SKY,
           cam=None):
    pos = np.asarray(positions, np.float32)
    idx = np.asarray(indices, np.uint32)
    p = pos[idx]
    n = scenes.face_normal(p[:, 0].astype(np.float64), p[:, 1].astype(np.float64),
                           p[:, 2].astype(np.float64))
    vn = np.zeros((len(pos), 3))
    for t, tri in enumerate(idx):
        vn[tri] = n[t]
    cam = cam or {"position": (0.0, 3.0, 0.0), "gaze": (0.0, 0.0, 0.0), "up": (0.0, 0.0, -1.0),
                  "fov": (math.radians(30),) * 2, "nearFar": (0.01, 1000.0)}
    return scenes.Scene(positions=pos, tbn=scenes.quat_world_to_tangent(vn), indices=idx,
                        surface_keys=np.asarray(keys, np.uint32), cull=np.asarray(cull, np.uint8),
                        materials=mats, light_tris=np.asarray(lights, np.uint32),
                        light_radiance=np.asarray(light_rad, np.float32).reshape(-1, 3),
                        light_two_sided=np.zeros(len(lights), np.uint8),
                        boundary_type=abi.BOUNDARY_SKYSPHERE_COOCTA, boundary_radiance=boundary,
                        scene_diameter=100.0, camera=cam)


LAMB = [{"type": abi.MAT_LAMBERT, "albedo": (0.6, 0.6, 0.6)}]


@pytest.mark.parametrize("mode", [abi.SAMPLE_PURE, abi.SAMPLE_NEE, abi.SAMPLE_NEE_MIS])
def test_sky_only_exact(mode):
    # one small triangle far behind the camera
    s = _scene([[0, 50, 0], [1, 50, 0], [0, 50, 1]], [[0, 1, 2]], [0], LAMB, [1])
    film, st = refcpu.OracleScene(s).render(abi.render_params(16, 16, 4, mode=mode))
    img = film_image(film)
    assert st["dead_paths"] == 16 * 16 * 4
    for c in range(3):
        assert np.all(img[c] == np.float32(SKY[c]))
    assert np.all(film[3] == 4.0)


def test_emitter_fills_view_exact():
    L = (3.0, 2.0, 1.0)
    quad = [[-10, 0, -10], [10, 0, -10], [10, 0, 10], [-10, 0, 10]]
    # facing +Y (towards the camera): CCW seen from above
    s = _scene(quad, [[0, 3, 2], [0, 2, 1]], [abi.LIGHT_FLAG | 0, abi.LIGHT_FLAG | 1], LAMB,
               [0, 0], lights=[0, 1], light_rad=[L, L])
    film, _ = refcpu.OracleScene(s).render(abi.render_params(16, 16, 2))
    img = film_image(film)
    for c in range(3):
        assert np.allclose(img[c], L[c], rtol=1e-6)


def _plane_scene(albedo):
    quad = [[-200, 0, -200], [200, 0, -200], [200, 0, 200], [-200, 0, 200]]
    mats = [{"type": abi.MAT_LAMBERT, "albedo": albedo}]
    return _scene(quad, [[0, 3, 2], [0, 2, 1]], [0, 0], mats, [1, 1])


@pytest.mark.parametrize("mode", [abi.SAMPLE_PURE, abi.SAMPLE_NEE, abi.SAMPLE_NEE_MIS])
def test_lambert_plane_under_sky(mode):
    a = (0.8, 0.5, 0.3)
    s = _plane_scene(a)
    film, _ = refcpu.OracleScene(s).render(abi.render_params(48, 48, 16, rr=(8, 8), mode=mode,
                                                             seed=3))
    img = film_image(film)
    for c in range(3):
        assert img[c].mean() == pytest.approx(a[c] * SKY[c], rel=0.02)


def test_sample_modes_agree_on_cornell():
    s = scenes.cornell()
    o = refcpu.OracleScene(s)
    means = []
    for mode in (abi.SAMPLE_PURE, abi.SAMPLE_NEE, abi.SAMPLE_NEE_MIS):
        film, _ = o.render(abi.render_params(48, 48, 64, rr=(8, 8), mode=mode, seed=11))
        means.append(film_image(film).mean(axis=(1, 2)))
    means = np.array(means)
    # Pure PT is noisy with a small light: compare NEE and MIS tightly, Pure loosely
    assert np.allclose(means[1], means[2], rtol=0.03)
    assert np.allclose(means[0], means[2], rtol=0.15)


def test_render_is_deterministic_and_seed_dependent():
    s = scenes.cornell()
    o = refcpu.OracleScene(s)
    a, _ = o.render(abi.render_params(32, 32, 4, seed=1), threads=1)
    b, _ = o.render(abi.render_params(32, 32, 4, seed=1), threads=8)
    c, _ = o.render(abi.render_params(32, 32, 4, seed=2))
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)
