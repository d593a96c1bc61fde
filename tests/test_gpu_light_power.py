"""NEE light selection by power (MRAY_RENDER_LIGHT_POWER) on the GPU against
the oracle: the alias tables built for a renderer equal the oracle's, and
renders at equal samples (one-scene, batch, instanced) match to the usual
bound, relMSE <= 1e-3 (eps 1e-2) with exact path counts.
"""
import numpy as np
import pytest

from mray_amd import abi, scenes
from oracle import refcpu
from tests.conftest import film_image, relmse
from tests.test_light_power_cpu import _lit_floor, _uneven

pytestmark = pytest.mark.gpu

POWER = abi.RENDER_LIGHT_POWER


def test_tables_equal_oracle(torch_gpu):
    from mray_amd import native
    sc = _uneven()
    acc = native.Accel(sc)
    r = native.Renderer(acc, abi.render_params(16, 12, 1, rr=(3, 6), flags=POWER))
    prob, alias, pmf = r.light_table(len(sc.light_tris))
    r.close()
    acc.close()
    oprob, oalias, opmf = refcpu.OracleScene(sc).light_power_table()
    assert np.array_equal(alias, oalias)
    np.testing.assert_allclose(prob, oprob, rtol=1e-6)
    np.testing.assert_allclose(pmf, opmf, rtol=1e-6)
    with pytest.raises(native.MRayError, match="uniformly"):
        acc2 = native.Accel(sc)
        r2 = native.Renderer(acc2, abi.render_params(16, 12, 1, rr=(3, 6)))
        try:
            r2.light_table(len(sc.light_tris))
        finally:
            r2.close()
            acc2.close()


@pytest.mark.parametrize("which,spectral", [("floor", False), ("terrain", False), ("terrain", True)])
def test_render_matches_oracle(torch_gpu, which, spectral):
    from mray_amd import native
    if spectral:
        refcpu.spectral_load(native.spectral_tables())
    sc = _lit_floor() if which == "floor" else _uneven(width=48, height=32)
    w, h = (24, 16) if which == "floor" else (48, 32)
    params = abi.render_params(w, h, 8, rr=(3, 6), seed=5, spectral=spectral, flags=POWER,
                               max_ray_count=2000)
    acc = native.Accel(sc)
    r = native.Renderer(acc, params)
    st = r.render_all()
    film = r.film()
    r.close()
    acc.close()
    ofilm, ost = refcpu.OracleScene(sc).render(params)
    assert st["dead_paths"] == ost["dead_paths"]
    e = relmse(film_image(film), film_image(ofilm))
    print(f"{which} spectral={spectral}: relMSE={e:.3e}", flush=True)
    assert e <= 1e-3


def test_batch_and_instanced_match_oracle(torch_gpu):
    from mray_amd import native
    scs = [_lit_floor(), _uneven(width=32, height=24), scenes.cornell()]
    views = [dict(scene=0, width=24, height=16, spp=4), dict(scene=1, width=32, height=24, spp=3),
             dict(scene=2, width=16, height=16, spp=4)]
    params = abi.render_params(1, 1, 1, rr=(3, 6), seed=8, flags=POWER, max_ray_count=1500)
    accs = [native.Accel(s) for s in scs]
    bs = native.BatchScene(accs)
    for a in accs:
        a.close()
    r = native.BatchRenderer(bs, views, params)
    r.render_all()
    films = r.film()
    r.close()
    bs.close()
    ofilms, _ = refcpu.render_batch([refcpu.OracleScene(s) for s in scs], views, params)
    for k, (f, o) in enumerate(zip(films, ofilms)):
        e = relmse(film_image(f), film_image(o))
        print(f"batch view {k}: relMSE={e:.3e}", flush=True)
        assert e <= 1e-3
    meshes, inst, env, flat = scenes.instanced_test_scene(width=40, height=30)
    ip = abi.render_params(40, 30, 4, rr=(3, 6), seed=9, flags=POWER, max_ray_count=1200)
    accs = [native.Accel(m) for m in meshes]
    isc = native.InstancedScene(accs, inst, env)
    for a in accs:
        a.close()
    ir = native.InstancedRenderer(isc, ip)
    ir.render_all()
    film = ir.film()
    ir.close()
    isc.close()
    ofilm, _ = refcpu.OracleScene(flat).render(ip)
    e = relmse(film_image(film), film_image(ofilm))
    print(f"instanced: relMSE={e:.3e}", flush=True)
    assert e <= 1e-3
