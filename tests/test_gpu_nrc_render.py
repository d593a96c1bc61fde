"""Path tracing with the neural radiance cache (Cornell box, RGB).

Paths terminate into the cache once their spread is large; 1 in 16 paths is
a full-length training path whose vertices train the cache after every frame.
Reference image: a 4096-spp plain render.  The trained cache must beat the same
schedule with a frozen (untrained, lr = 0) cache by a wide margin, i.e. the
network's training is what carries the estimate, and terminated paths must
have cut the extension rays against plain path tracing at equal samples."""
import numpy as np
import pytest

from mray_amd import abi, scenes
from tests.conftest import film_image, relmse

pytestmark = pytest.mark.gpu

W = H = 128


def _plain(scene, spp, seed):
    from mray_amd import native
    acc = native.Accel(scene)
    r = native.Renderer(acc, abi.render_params(W, H, spp, rr=(8, 8), seed=seed))
    st = r.render_all()
    film = r.film()
    r.close()
    acc.close()
    return film, st


def _cached(scene, frames, lr, seed):
    from mray_amd import native
    acc = native.Accel(scene)
    nrc = native.Nrc(lr=lr, seed=seed)
    r = native.Renderer(acc, abi.render_params(W, H, frames, rr=(8, 8), seed=seed))
    r.attach_nrc(nrc, train_shift=4, spread=0.01, batch=1 << 12, steps=4)
    ext = 0
    losses = []
    for f in range(frames):
        st = r.render_frame(1)
        ext += st["ext_rays"]
        if f % 8 == 7:
            losses.append(nrc.loss())
    cs = r.cache_stats()
    film = r.film()
    r.close()
    nrc.close()
    acc.close()
    return film, ext, cs, losses


def test_trained_cache_beats_untrained(torch_gpu):
    scene = scenes.cornell()
    ref, _ = _plain(scene, 4096, seed=11)
    ref_img = film_image(ref)
    frames = 48
    a, ext_a, cs_a, losses = _cached(scene, frames, lr=1e-2, seed=1)
    b, ext_b, cs_b, _ = _cached(scene, frames, lr=0.0, seed=1)
    c, st_c = _plain(scene, frames, seed=1)
    ea = relmse(film_image(a), ref_img)
    eb = relmse(film_image(b), ref_img)
    ec = relmse(film_image(c), ref_img)
    print(f"relMSE trained {ea:.4e}  frozen {eb:.4e}  plain {ec:.4e}; ext rays cached "
          f"{ext_a} plain {st_c['ext_rays']}; cache {cs_a}; losses {losses}", flush=True)
    assert np.isfinite(a).all()
    assert np.allclose(a[3], frames) and np.allclose(c[3], frames)
    assert cs_a["queries"] > 0 and cs_a["training_pairs"] > 0
    assert ext_a < st_c["ext_rays"]
    assert ea < 0.5 * eb
