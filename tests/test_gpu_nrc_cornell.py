"""The radiance-cache workload on its own scene: the Cornell box at
1920 x 1080, 1 spp per frame, the cache of the specified architecture
(16 levels x 2^19 entries x 2 features, 5 x 64 ReLU MLP, Adam lr 1e-2,
4 x 2^14 training pairs per frame), against plain path tracing with the same
frames and against a high-spp plain judge.

Checks: the trained cache beats the same schedule with a frozen (lr = 0)
cache, its error stays within a small factor of plain path tracing at equal
frames (the cache trades variance for bias; the equal-time figures are the
benchmark's), terminated paths cut the extension rays, and the image is
finite with one unit of filter weight per frame."""
import time

import numpy as np
import pytest

from mray_amd import abi, scenes
from tests.conftest import film_image, relmse

pytestmark = pytest.mark.gpu

W, H = 1920, 1080
FRAMES = 16


def _plain(acc, spp, seed):
    from mray_amd import native
    r = native.Renderer(acc, abi.render_params(W, H, spp, rr=(8, 8), seed=seed))
    t = time.perf_counter()
    done = 0
    st = None
    while done < spp:
        step = min(256, spp - done)
        st = r.render_frame(step)
        done += step
    dt = time.perf_counter() - t
    film = r.film()
    r.close()
    return film, dt


def _cached(acc, lr, seed):
    from mray_amd import native
    nrc = native.Nrc(lr=lr, seed=seed, lazy_adam=True)
    r = native.Renderer(acc, abi.render_params(W, H, FRAMES, rr=(8, 8), seed=seed))
    r.attach_nrc(nrc, train_shift=4, spread=1.0, batch=1 << 14, steps=4, min_depth=2,
                 nee_at_end=True)
    ext = 0
    t = time.perf_counter()
    for _ in range(FRAMES):
        ext += r.render_frame(1)["ext_rays"]
    dt = time.perf_counter() - t
    cs = r.cache_stats()
    film = r.film()
    r.close()
    nrc.close()
    return film, ext, cs, dt


def test_cornell_cache_workload(torch_gpu):
    from mray_amd import native
    scene = scenes.cornell(aspect=W / H)
    acc = native.Accel(scene)
    ref, _ = _plain(acc, 1024, seed=99)
    ref_img = film_image(ref)
    plain, t_plain = _plain(acc, FRAMES, seed=1)
    trained, ext_t, cs, t_cached = _cached(acc, 1e-2, seed=1)
    frozen, _, _, _ = _cached(acc, 0.0, seed=1)
    # plain with the same frames, counted in extension rays
    r = native.Renderer(acc, abi.render_params(W, H, FRAMES, rr=(8, 8), seed=1))
    ext_p = sum(r.render_frame(1)["ext_rays"] for _ in range(FRAMES))
    r.close()
    acc.close()
    e_p = relmse(film_image(plain), ref_img)
    e_t = relmse(film_image(trained), ref_img)
    e_f = relmse(film_image(frozen), ref_img)
    print(f"cornell 1080p x {FRAMES} frames: relMSE plain {e_p:.4e} trained {e_t:.4e} "
          f"frozen {e_f:.4e}; time plain {t_plain:.3f} s cached {t_cached:.3f} s; "
          f"ext rays cached {ext_t} plain {ext_p}; cache {cs}", flush=True)
    assert np.isfinite(trained).all()
    assert np.allclose(trained[3], FRAMES, rtol=1e-5)
    assert cs["queries"] > 0 and cs["training_pairs"] > 0
    assert ext_t < ext_p
    assert e_t < 0.8 * e_f
    assert e_t < 2.0 * e_p
