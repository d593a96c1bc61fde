"""Per-viewport radiance caches in one batched renderer.

Instanced two-level scene (the batched + cache workload at reduced size):
spectral (4 wavelengths), depth 12, NEE+MIS, lights chosen by power, three
viewports of uneven size rendered frame by frame in one renderer, one cache
per view, trained online after every frame.

* Frame order itself: a batch rendered by frames puts exactly one unit of
  filter weight per frame into every pixel of every view, and each view's
  image agrees in mean with the same view rendered in view order.
* The caches on that scene: queries and training pairs flow for every view,
  terminated paths cut the extension rays, the image stays within a small
  factor of plain path tracing.  The terrain is open to a constant sky, so
  most bounce rays leave the scene before a second surface hit and the
  caches carry little of the light there; the trained-against-frozen test
  runs on closed scenes instead:
* Three Cornell boxes (different box angles and tessellations) batched with
  uneven views, spectral: every view's trained cache beats the same schedule
  with frozen (lr = 0) caches against a high-spp plain judge.
"""
import numpy as np
import pytest

from mray_amd import abi, scenes
from tests.conftest import film_image, relmse

pytestmark = pytest.mark.gpu

VIEWS = [(320, 180), (256, 256), (192, 108)]
FRAMES = 16


@pytest.fixture(scope="module")
def world(torch_gpu):
    from mray_amd import native
    meshes, inst, env = scenes.instanced_workload(mesh_count=3, instance_count=60,
                                                  tri_range=(2000, 6000), emissive_tris=400,
                                                  width=320, height=180)
    accs = [native.Accel(m) for m in meshes]
    isc = native.InstancedScene(accs, inst, env)
    for a in accs:
        a.close()
    views = scenes.instanced_views(spp=FRAMES, count=len(VIEWS))
    for v, (w, h) in zip(views, VIEWS):
        v["width"], v["height"] = w, h
        v["camera"]["fov"] = (v["camera"]["fov"][0], scenes.fov_y_from_x(v["camera"]["fov"][0], w / h))
    yield isc, views
    isc.close()


def _params(slots=1 << 18):
    return abi.render_params(1, 1, 1, rr=(12, 12), mode=abi.SAMPLE_NEE_MIS, spectral=True,
                             seed=3, max_ray_count=slots, flags=abi.RENDER_LIGHT_POWER)


def _frames(r, n, spp=1):
    ext = 0
    for _ in range(n):
        ext += r.render_frame(spp)["ext_rays"]
    return ext


def test_frame_order_weights_and_means(world):
    from mray_amd import native
    isc, views = world
    r = native.BatchRenderer(isc, views, _params())
    _frames(r, 8)
    frame_films = r.film()
    r.reset()                      # back to view order
    r.render_all()                 # FRAMES spp per view
    view_films = r.film()
    r.close()
    for k, (f, g) in enumerate(zip(frame_films, view_films)):
        assert np.isfinite(f).all()
        assert np.allclose(f[3], 8.0, rtol=1e-5), k
        assert np.allclose(g[3], float(FRAMES), rtol=1e-5), k
        a = film_image(f).mean()
        b = film_image(g).mean()
        print(f"view {k}: frame order mean {a:.5f}, view order mean {b:.5f}", flush=True)
        assert abs(a - b) <= 0.06 * b


def _cached(bscene, views, lr, frames, **kw):
    from mray_amd import native
    r = native.BatchRenderer(bscene, views, _params())
    caches = [native.Nrc(lr=lr, seed=7 + k, lazy_adam=True) for k in range(len(views))]
    r.attach_nrcs(caches, **kw)
    ext = _frames(r, frames)
    cs = r.cache_stats()
    films = r.film()
    r.close()
    for c in caches:
        c.close()
    return films, ext, cs


def _judge(bscene, views, spp):
    from mray_amd import native
    r = native.BatchRenderer(bscene, views, _params())
    for _ in range(spp // 64):
        r.render_frame(64)
    ref = [film_image(f) for f in r.film()]
    r.close()
    return ref


def _plain(bscene, views, frames):
    from mray_amd import native
    r = native.BatchRenderer(bscene, views, _params())
    ext = _frames(r, frames)
    films = r.film()
    r.close()
    return films, ext


def test_instanced_caches_run(world):
    isc, views = world
    ref = _judge(isc, views, 256)
    plain, ext_plain = _plain(isc, views, FRAMES)
    kw = dict(train_shift=3, spread=0.01, batch=1 << 12, steps=4, min_depth=1, nee_at_end=False)
    trained, ext_t, cs = _cached(isc, views, 1e-2, FRAMES, **kw)
    assert cs["queries"] > 0 and cs["training_pairs"] > 0
    assert ext_t < ext_plain
    print(f"instanced: ext rays cached {ext_t} plain {ext_plain}; cache {cs}", flush=True)
    for k in range(len(views)):
        assert np.isfinite(trained[k]).all()
        assert np.allclose(trained[k][3], float(FRAMES), rtol=1e-5)
        et = relmse(film_image(trained[k]), ref[k])
        ep = relmse(film_image(plain[k]), ref[k])
        print(f"view {k} {VIEWS[k]}: relMSE trained {et:.4e} plain {ep:.4e}", flush=True)
        assert et < 1.5 * ep


CORNELL_VIEWS = [(192, 108), (128, 128), (96, 96)]
CORNELL_FRAMES = 48


@pytest.fixture(scope="module")
def cornell_batch(torch_gpu):
    from mray_amd import native
    accs = []
    views = []
    for k, ((w, h), tess) in enumerate(zip(CORNELL_VIEWS, (1, 3, 2))):
        accs.append(native.Accel(scenes.cornell(aspect=w / h, tess=tess, seed=200 + k)))
        views.append({"scene": k, "width": w, "height": h, "spp": CORNELL_FRAMES})
    bs = native.BatchScene(accs)
    for a in accs:
        a.close()
    yield bs, views
    bs.close()


def test_per_view_caches(cornell_batch):
    bs, views = cornell_batch
    ref = _judge(bs, views, 1024)
    plain, ext_plain = _plain(bs, views, CORNELL_FRAMES)
    kw = dict(train_shift=4, spread=0.01, batch=1 << 12, steps=4, min_depth=1, nee_at_end=False)
    trained, ext_t, cs = _cached(bs, views, 1e-2, CORNELL_FRAMES, **kw)
    frozen, _, _ = _cached(bs, views, 0.0, CORNELL_FRAMES, **kw)
    assert cs["queries"] > 0 and cs["training_pairs"] > 0
    assert ext_t < ext_plain
    print(f"cornell batch: ext rays cached {ext_t} plain {ext_plain}; cache {cs}", flush=True)
    for k in range(len(views)):
        assert np.isfinite(trained[k]).all()
        assert np.allclose(trained[k][3], float(CORNELL_FRAMES), rtol=1e-5)
        et = relmse(film_image(trained[k]), ref[k])
        ef = relmse(film_image(frozen[k]), ref[k])
        ep = relmse(film_image(plain[k]), ref[k])
        print(f"view {k} {CORNELL_VIEWS[k]}: relMSE trained {et:.4e} frozen {ef:.4e} "
              f"plain {ep:.4e}", flush=True)
        assert et < 0.5 * ef
