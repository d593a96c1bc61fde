"""Batch renderer (mray_batch_create) on the GPU against the oracle's batch
render at equal samples, and against the GPU one-scene renderer.

Both sides run the same schedule (one slot pool, global path indices view by
view, stable dead-slot refill, per-slot PCG32 states), so each view's image
matches the oracle's up to transcendental-function ulps and float-atomic
order: relMSE <= 1e-3 per view (eps 1e-2) and exact path counts.
"""
import math

import numpy as np
import pytest

from mray_amd import abi, scenes
from oracle import refcpu
from tests.conftest import film_image, relmse

pytestmark = pytest.mark.gpu

CAM = {"position": (0.3, 1.1, 3.0), "gaze": (0.0, 0.8, 0.0), "up": (0.0, 1.0, 0.0),
       "fov": (math.radians(45),) * 2, "nearFar": (0.01, 100.0)}


def _scenes():
    return [scenes.cornell(),
            scenes.terrain_spheres(grid=96, spheres=60, width=80, height=45),
            scenes.terrain_spheres(grid=48, spheres=20, width=40, height=30, seed=7)]


VIEWS = [dict(scene=1, width=80, height=45, spp=3),
         dict(scene=0, width=48, height=48, spp=4),
         dict(scene=2, width=40, height=30, spp=2),
         dict(scene=0, width=32, height=24, spp=2, camera=CAM),
         dict(scene=1, width=17, height=9, spp=5)]


def _gpu_batch(scs, views, params, width=4, tpb=None):
    from mray_amd import native
    accs = [native.Accel(s, width=width) for s in scs]
    bs = native.BatchScene(accs)
    for a in accs:
        a.close()      # the merged scene holds its own copies
    r = native.BatchRenderer(bs, views, params)
    if tpb:
        r.set_trace_tpb(tpb)
    st = r.render_all()
    films = r.film()
    times = r.view_times()
    r.close()
    bs.close()
    return films, st, times


def _check_against_oracle(scs, views, params, **kw):
    films, st, times = _gpu_batch(scs, views, params, **kw)
    ofilms, ost = refcpu.render_batch([refcpu.OracleScene(s) for s in scs], views, params)
    print(f"gpu={st} cpu={ost}", flush=True)
    assert st["dead_paths"] == ost["dead_paths"] == sum(v["width"] * v["height"] * v["spp"]
                                                        for v in views)
    for k, (f, o) in enumerate(zip(films, ofilms)):
        assert np.array_equal(f[3], o[3]) or np.allclose(f[3], o[3], rtol=1e-5)
        e = relmse(film_image(f), film_image(o))
        print(f"view {k}: relMSE={e:.3e}", flush=True)
        assert e <= 1e-3
    assert np.all(times >= 0.0)
    return films, times


def test_batch_rgb_matches_oracle(torch_gpu):
    params = abi.render_params(1, 1, 1, rr=(3, 8), seed=4, max_ray_count=5000)
    _check_against_oracle(_scenes(), VIEWS, params)


def test_batch_spectral_matches_oracle(torch_gpu):
    from mray_amd import native
    refcpu.spectral_load(native.spectral_tables())
    params = abi.render_params(1, 1, 1, rr=(3, 8), seed=6, spectral=True, max_ray_count=4096)
    _check_against_oracle(_scenes(), VIEWS, params)


@pytest.mark.parametrize("width,tpb,flags", [(8, None, 0), (4, 128, 0),
                                             (4, None, abi.RENDER_MATERIAL_QUEUES)])
def test_batch_settings_match_oracle(torch_gpu, width, tpb, flags):
    """Compressed 8-wide trees (relocated CNode8 children), 128-thread
    traversal workgroups and material-keyed queues: the same samples."""
    params = abi.render_params(1, 1, 1, rr=(3, 8), seed=8, max_ray_count=3000, flags=flags)
    _check_against_oracle(_scenes(), VIEWS[:3], params, width=width, tpb=tpb)


def test_first_view_equals_single_renderer(torch_gpu):
    """View 0's paths are assigned exactly as in a one-scene render with the
    same slot pool."""
    from mray_amd import native
    scs = _scenes()
    params = abi.render_params(80, 45, 3, rr=(3, 8), seed=2, max_ray_count=2000)
    films, st, _ = _gpu_batch(scs, VIEWS, params)
    acc = native.Accel(scs[1])
    r = native.Renderer(acc, params)
    r.render_all()
    single = r.film()
    r.close()
    acc.close()
    e = relmse(film_image(films[0]), film_image(single))
    print(f"view 0 vs single: relMSE={e:.3e}", flush=True)
    assert np.allclose(films[0][3], single[3])
    assert e <= 1e-6


def test_view_times_and_errors(torch_gpu):
    from mray_amd import native
    scs = _scenes()[:2]
    accs = [native.Accel(s) for s in scs]
    bs = native.BatchScene(accs)
    info = bs.info()
    assert info["scenes"] == 2 and info["triangles"] == sum(s.triangle_count for s in scs)
    params = abi.render_params(1, 1, 1, rr=(3, 8), seed=1)
    r = native.BatchRenderer(bs, VIEWS[:2], params)
    assert np.all(r.view_times() < 0)          # nothing finished yet
    st = r.render_all()
    t = r.view_times()
    print(f"view times ms: {t}", flush=True)
    assert st["finished"] and np.all(t >= 0)
    r.close()
    with pytest.raises(native.MRayError, match="names scene 5"):
        native.BatchRenderer(bs, [dict(scene=5, width=8, height=8, spp=1)], params)
    with pytest.raises(native.MRayError, match="non-zero"):
        native.BatchRenderer(bs, [dict(scene=0, width=0, height=8, spp=1)], params)
    bs.close()
    a8 = native.Accel(scs[0], width=8)
    with pytest.raises(native.MRayError, match="same BVH width"):
        native.BatchScene([accs[0], a8])
    a8.close()
    for a in accs:
        a.close()


def test_autotune_small_batch(torch_gpu):
    from mray_amd import native
    scs = _scenes()
    a4 = [native.Accel(s, width=4) for s in scs]
    a8 = [native.Accel(s, width=8) for s in scs]
    b4, b8 = native.BatchScene(a4), native.BatchScene(a8)
    params = abi.render_params(1, 1, 1, rr=(3, 8), seed=1)
    best, trials = native.batch_autotune(b4, b8, VIEWS, params)
    print(f"best={best}", flush=True)
    for t in trials:
        print(t, flush=True)
    # 2 widths x 2 queue modes x 1 pool size (all pixels fit) x 2 workgroup sizes
    assert len(trials) == 8
    assert best["mpaths_per_s"] == max(t["mpaths_per_s"] for t in trials) > 0
    assert best["width"] in (4, 8) and best["trace_tpb"] in (128, 256)
    for x in a4 + a8:
        x.close()
    b4.close()
    b8.close()
