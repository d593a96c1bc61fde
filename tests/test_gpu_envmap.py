"""Textured Skysphere on the GPU against the oracle.

The luminance distribution is built by the same host restatement on both
sides, so the CDF tables read back from the device equal the oracle's bit
for bit.  Renders at equal samples (one-scene, batch, instanced) use the same
sampling and the same bilinear texture reads, so images match to the
renderer's usual bound: relMSE <= 1e-3 (eps 1e-2), exact path counts.
"""
import numpy as np
import pytest

from mray_amd import abi, scenes
from oracle import refcpu
from tests.conftest import film_image, relmse

pytestmark = pytest.mark.gpu


def _scene(env=None, seed=2, width=64, height=48):
    sc = scenes.terrain_spheres(grid=40, spheres=12, width=width, height=height, seed=seed)
    sc.env_map = scenes.sky_env(128, 96) if env is None else env
    return sc


def test_cdfs_equal_oracle(torch_gpu):
    from mray_amd import native
    sc = _scene()
    acc = native.Accel(sc)
    cx, cy = acc.environment_cdfs(96, 128)
    acc.close()
    ox, oy = refcpu.OracleScene(sc).env_cdfs()
    assert np.array_equal(cx, ox) and np.array_equal(cy, oy)


@pytest.mark.parametrize("spectral", [False, True])
def test_render_matches_oracle(torch_gpu, spectral):
    from mray_amd import native
    if spectral:
        refcpu.spectral_load(native.spectral_tables())
    sc = _scene()
    params = abi.render_params(64, 48, 4, rr=(3, 8), seed=3, spectral=spectral, max_ray_count=3000)
    acc = native.Accel(sc)
    r = native.Renderer(acc, params)
    st = r.render_all()
    film = r.film()
    r.close()
    acc.close()
    ofilm, ost = refcpu.OracleScene(sc).render(params)
    assert st["dead_paths"] == ost["dead_paths"]
    assert np.allclose(film[3], ofilm[3], rtol=1e-5)
    e = relmse(film_image(film), film_image(ofilm))
    print(f"spectral={spectral}: relMSE={e:.3e}", flush=True)
    assert e <= 1e-3


def test_batch_views_with_their_own_skies(torch_gpu):
    from mray_amd import native
    scs = [_scene(seed=2), scenes.cornell(),
           _scene(env=scenes.sky_env(32, 32, sun=(-0.3, 0.5, 0.6), sun_radiance=20.0), seed=5)]
    views = [dict(scene=0, width=40, height=30, spp=3), dict(scene=1, width=24, height=24, spp=2),
             dict(scene=2, width=32, height=20, spp=4)]
    params = abi.render_params(1, 1, 1, rr=(3, 8), seed=9, max_ray_count=2500)
    accs = [native.Accel(s) for s in scs]
    bs = native.BatchScene(accs)
    for a in accs:
        a.close()      # the merged scene keeps its own copy of each sky
    r = native.BatchRenderer(bs, views, params)
    st = r.render_all()
    films = r.film()
    r.close()
    bs.close()
    ofilms, ost = refcpu.render_batch([refcpu.OracleScene(s) for s in scs], views, params)
    assert st["dead_paths"] == ost["dead_paths"]
    for k, (f, o) in enumerate(zip(films, ofilms)):
        e = relmse(film_image(f), film_image(o))
        print(f"view {k}: relMSE={e:.3e}", flush=True)
        assert e <= 1e-3


def test_instanced_sky(torch_gpu):
    from mray_amd import native
    meshes, inst, env, _ = scenes.instanced_test_scene(width=48, height=36)
    env["env_map"] = scenes.sky_env(64, 64, sun=(0.2, 0.8, -0.4))
    flat = scenes.flatten_instances(meshes, inst, env)
    params = abi.render_params(48, 36, 4, rr=(3, 8), seed=4, max_ray_count=2000)
    accs = [native.Accel(m) for m in meshes]
    isc = native.InstancedScene(accs, inst, env)
    for a in accs:
        a.close()
    r = native.InstancedRenderer(isc, params)
    r.render_all()
    film = r.film()
    r.close()
    isc.close()
    ofilm, _ = refcpu.OracleScene(flat).render(params)
    e = relmse(film_image(film), film_image(ofilm))
    print(f"instanced sky: relMSE={e:.3e}", flush=True)
    assert e <= 1e-3


def test_errors(torch_gpu):
    from mray_amd import native
    acc = native.Accel(scenes.cornell())
    with pytest.raises(native.MRayError, match="not a Skysphere"):
        acc.set_environment(np.ones((4, 4, 3), np.float32))
    acc.close()
    acc = native.Accel(scenes.terrain_spheres(grid=8, spheres=1, width=8, height=8))
    with pytest.raises(native.MRayError, match="zero luminance"):
        acc.set_environment(np.zeros((4, 4, 3), np.float32))
    bad = np.ones((4, 4, 3), np.float32)
    bad[1, 2, 0] = np.nan
    with pytest.raises(native.MRayError, match="non-finite"):
        acc.set_environment(bad)
    acc.close()


@pytest.mark.parametrize("textured", [False, True])
def test_spherical_skysphere_matches_oracle(torch_gpu, textured):
    """The latitude-longitude Skysphere, constant and textured."""
    from mray_amd import native
    sc = _scene(env=scenes.sky_env(96, 48, mapping="spherical") if textured else None)
    if not textured:
        sc.env_map = None
    sc.boundary_type = abi.BOUNDARY_SKYSPHERE_SPHERICAL
    params = abi.render_params(48, 32, 4, rr=(3, 8), seed=11, max_ray_count=1500)
    acc = native.Accel(sc)
    r = native.Renderer(acc, params)
    st = r.render_all()
    film = r.film()
    r.close()
    acc.close()
    ofilm, ost = refcpu.OracleScene(sc).render(params)
    assert st["dead_paths"] == ost["dead_paths"]
    e = relmse(film_image(film), film_image(ofilm))
    print(f"spherical textured={textured}: relMSE={e:.3e}", flush=True)
    assert e <= 1e-3


def test_spherical_skysphere_batch_and_instanced(torch_gpu):
    """The latitude-longitude Skysphere in a batch beside a concentric-
    octahedral one, and under an instanced scene."""
    from mray_amd import native
    a = _scene(env=scenes.sky_env(64, 32, mapping="spherical"), seed=3)
    a.boundary_type = abi.BOUNDARY_SKYSPHERE_SPHERICAL
    scs = [a, _scene(seed=4, env=scenes.sky_env(32, 32))]
    views = [dict(scene=0, width=32, height=24, spp=3), dict(scene=1, width=24, height=16, spp=2)]
    params = abi.render_params(1, 1, 1, rr=(3, 8), seed=12, max_ray_count=1500)
    accs = [native.Accel(s) for s in scs]
    bs = native.BatchScene(accs)
    for acc in accs:
        acc.close()
    r = native.BatchRenderer(bs, views, params)
    st = r.render_all()
    films = r.film()
    r.close()
    bs.close()
    ofilms, ost = refcpu.render_batch([refcpu.OracleScene(s) for s in scs], views, params)
    assert st["dead_paths"] == ost["dead_paths"]
    for k, (f, o) in enumerate(zip(films, ofilms)):
        e = relmse(film_image(f), film_image(o))
        print(f"batch view {k}: relMSE={e:.3e}", flush=True)
        assert e <= 1e-3
    meshes, inst, env, _ = scenes.instanced_test_scene(width=40, height=30)
    env["boundary_type"] = abi.BOUNDARY_SKYSPHERE_SPHERICAL
    env["env_map"] = scenes.sky_env(48, 24, sun=(0.2, 0.8, -0.4), mapping="spherical")
    flat = scenes.flatten_instances(meshes, inst, env)
    assert flat.boundary_type == abi.BOUNDARY_SKYSPHERE_SPHERICAL
    ip = abi.render_params(40, 30, 4, rr=(3, 8), seed=13, max_ray_count=1200)
    accs = [native.Accel(m) for m in meshes]
    isc = native.InstancedScene(accs, inst, env)
    for acc in accs:
        acc.close()
    ir = native.InstancedRenderer(isc, ip)
    ir.render_all()
    film = ir.film()
    ir.close()
    isc.close()
    ofilm, _ = refcpu.OracleScene(flat).render(ip)
    e = relmse(film_image(film), film_image(ofilm))
    print(f"instanced spherical sky: relMSE={e:.3e}", flush=True)
    assert e <= 1e-3
