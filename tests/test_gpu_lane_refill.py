"""Closest-hit traversal with lane refill (MRAY_RENDER_LANE_REFILL) against
the one-ray-per-lane kernel.

A refilled lane walks its next ray with the same node visits, in the same
order, as the one-ray-per-lane walk, so every hit is the same: the path
schedule (extension, shadow and finished path counts) and the traversal
counters (nodes visited, triangles tested) must be exactly equal, and the
films equal up to the order of the film's float atomics.  Cases: a terrain
scene of one accelerator, 4-wide, spectral; a batch of a Cornell box and a
terrain with uneven views; the Cornell workload against the oracle.
"""
import numpy as np
import pytest

from mray_amd import abi, scenes
from oracle import refcpu
from tests.conftest import film_image, relmse

pytestmark = pytest.mark.gpu

REFILL = abi.RENDER_LANE_REFILL


def _single(acc, params):
    from mray_amd import native
    r = native.Renderer(acc, params)
    r.set_measurement(counting=True)
    st = r.render_all()
    film = r.film()
    r.close()
    return film, st


def _same_schedule(a, b):
    for k in ("dead_paths", "ext_rays", "shadow_rays", "node_visits", "tri_tests",
              "shadow_node_visits", "shadow_tri_tests"):
        assert a[k] == b[k], (k, a[k], b[k])


def test_refill_terrain_single(torch_gpu):
    from mray_amd import native
    scene = scenes.terrain_spheres(grid=256, spheres=300, width=320, height=180)
    acc = native.Accel(scene)
    base = abi.render_params(320, 180, 4, rr=(10, 10), mode=abi.SAMPLE_NEE_MIS, spectral=True,
                             seed=2)
    refill = abi.render_params(320, 180, 4, rr=(10, 10), mode=abi.SAMPLE_NEE_MIS, spectral=True,
                               seed=2, flags=REFILL)
    fa, sa = _single(acc, base)
    fb, sb = _single(acc, refill)
    acc.close()
    print(f"terrain 320x180 x 4: plain {sa} refill {sb}", flush=True)
    assert sa["node_visits"] > 0
    _same_schedule(sa, sb)
    assert np.isfinite(fb).all()
    assert relmse(film_image(fb), film_image(fa)) <= 1e-10


def test_refill_batch(torch_gpu):
    from mray_amd import native
    views = [{"scene": 0, "width": 200, "height": 120, "spp": 3},
             {"scene": 1, "width": 96, "height": 160, "spp": 5}]
    accs = [native.Accel(scenes.cornell(aspect=200 / 120, tess=4, seed=31)),
            native.Accel(scenes.terrain_spheres(grid=128, spheres=80, width=96, height=160))]
    bs = native.BatchScene(accs)
    for a in accs:
        a.close()
    out = []
    for flags in (0, REFILL):
        p = abi.render_params(1, 1, 1, rr=(8, 8), mode=abi.SAMPLE_NEE_MIS, spectral=True,
                              seed=5, max_ray_count=1 << 15, flags=flags)
        r = native.BatchRenderer(bs, views, p)
        r.set_measurement(counting=True)
        st = r.render_all()
        out.append((r.film(), st))
        r.close()
    bs.close()
    (fa, sa), (fb, sb) = out
    print(f"batch: plain {sa} refill {sb}", flush=True)
    _same_schedule(sa, sb)
    for a, b in zip(fa, fb):
        assert np.isfinite(b).all()
        assert relmse(film_image(b), film_image(a)) <= 1e-10


def test_refill_cornell_oracle(torch_gpu):
    from mray_amd import native
    scene = scenes.cornell()
    params = abi.render_params(256, 256, 16, rr=(8, 8), mode=abi.SAMPLE_NEE_MIS, seed=0,
                               flags=REFILL)
    acc = native.Accel(scene)
    r = native.Renderer(acc, params)
    st = r.render_all()
    film = r.film()
    r.close()
    acc.close()
    ref, rst = refcpu.OracleScene(scene).render(params)
    assert st["dead_paths"] == rst["dead_paths"] == 256 * 256 * 16
    assert st["ext_rays"] == rst["ext_rays"]
    e = relmse(film_image(film), film_image(ref))
    print(f"cornell 256^2 x 16 with lane refill: relMSE={e:.3e}", flush=True)
    assert e <= 1e-3
