"""TracerI front-end on the GPU: a scene fed through ConstructTracer /
Create*Group / Push*Attribute / CreateSurface / CommitSurfaces /
CreateRenderer / StartRender / DoRenderWork renders the same image as the
same scene handed to the C-ABI directly (same kernels, same samples; only the
order of the film's float atomics may differ)."""
import os
import subprocess

import numpy as np
import pytest

from mray_amd import abi, scenes
from tests.conftest import film_image, relmse
from tests.test_tracer_api import DRIVER, TRACER_LIB, _built, write_scene_file

pytestmark = pytest.mark.gpu


def _driver_render(scene, tmp_path, w, h, spp, spectral, mode, rr, seed, exr_prefix=None,
                   render_mode="Throughput", burst=4, hint=0):
    """The driver plays the run command: it accumulates the image sections
    handed over through the TimelineSemaphore into means + weights.  Returned
    as a film (sums = mean x weight) for comparison with the C-ABI's.
    hint = TracerParameters::parallelizationHint (0: width x height)."""
    _built()
    src = str(tmp_path / "scene.bin")
    out = str(tmp_path / "film.bin")
    write_scene_file(scene, src)
    cmd = [DRIVER, TRACER_LIB, src, "render", out, str(w), str(h), str(spp), str(int(spectral)),
           str(mode), str(rr[0]), str(rr[1]), str(seed), str(hint), render_mode, str(burst)]
    if exr_prefix:
        cmd.append(exr_prefix)
    print("run:", " ".join(cmd[2:]), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    assert "1 save request" in r.stdout
    a = np.fromfile(out, np.float32).reshape(4, h, w)
    return np.concatenate([a[:3] * a[3:4], a[3:4]])


def _cabi_render(scene, params):
    from mray_amd import native
    acc = native.Accel(scene)
    rnd = native.Renderer(acc, params)
    rnd.render_all()
    film = rnd.film()
    rnd.close()
    acc.close()
    return film


@pytest.mark.parametrize("spectral", [False, True])
def test_tracer_api_matches_cabi_cornell(torch_gpu, tmp_path, spectral):
    scene = scenes.cornell()
    params = abi.render_params(64, 64, 4, rr=(8, 8), mode=abi.SAMPLE_NEE_MIS, spectral=spectral,
                               seed=3)
    a = _driver_render(scene, tmp_path, 64, 64, 4, spectral, abi.SAMPLE_NEE_MIS, (8, 8), 3)
    b = _cabi_render(scene, params)
    assert np.allclose(a[3], 4.0) and np.allclose(b[3], 4.0)
    e = relmse(film_image(a), film_image(b))
    print(f"relMSE TracerI vs C-ABI = {e:.3e}", flush=True)
    assert e < 1e-8


@pytest.mark.parametrize("sky", [abi.BOUNDARY_SKYSPHERE_COOCTA, abi.BOUNDARY_SKYSPHERE_SPHERICAL])
def test_tracer_api_matches_cabi_terrain(torch_gpu, tmp_path, sky):
    """Under a (L)Skysphere_CoOcta and a (L)Skysphere_Spherical boundary."""
    scene = scenes.terrain_spheres(grid=64, spheres=60, width=80, height=45)
    scene.boundary_type = sky
    params = abi.render_params(80, 45, 2, rr=(10, 10), mode=abi.SAMPLE_NEE_MIS, seed=1)
    a = _driver_render(scene, tmp_path, 80, 45, 2, False, abi.SAMPLE_NEE_MIS, (10, 10), 1)
    b = _cabi_render(scene, params)
    e = relmse(film_image(a), film_image(b))
    print(f"relMSE TracerI vs C-ABI = {e:.3e}", flush=True)
    assert e < 1e-8


def test_json_scene_renders_like_cabi(torch_gpu, tmp_path):
    """Cornell box written as an MRay JSON scene file, loaded by LoadSceneMRay
    and rendered through TracerI, against the same scene through the C-ABI
    (TBN quaternions are rebuilt from normals, so samples agree up to float
    rounding: the equal-spp image bar applies)."""
    import subprocess
    from tests.test_tracer_api import DRIVER, TRACER_LIB
    scene = scenes.cornell()
    src = str(tmp_path / "cornell.json")
    scenes.to_mray_json(scene, src)
    out = str(tmp_path / "film.bin")
    cmd = [DRIVER, TRACER_LIB, src, "render", out, "64", "64", "4", "0", str(abi.SAMPLE_NEE_MIS),
           "8", "8", "5", "0", "Throughput", "4"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    a = np.fromfile(out, np.float32).reshape(4, 64, 64)      # consumer means + weights
    a = np.concatenate([a[:3] * a[3:4], a[3:4]])
    b = _cabi_render(scene, abi.render_params(64, 64, 4, rr=(8, 8), mode=abi.SAMPLE_NEE_MIS, seed=5))
    e = relmse(film_image(a), film_image(b))
    print(f"relMSE JSON/TracerI vs C-ABI = {e:.3e}", flush=True)
    assert e <= 1e-3


def test_tracer_saves_exr_like_the_run_command(torch_gpu, tmp_path):
    """triggerSave once at the sample limit; the consumer's accumulated image
    saved as <prefix>_<wpp>wpp.exr equals the C-ABI render's means, top row
    first."""
    from tests.exr_read import read_exr
    scene = scenes.cornell()
    prefix = str(tmp_path / "cornell")
    a = _driver_render(scene, tmp_path, 48, 32, 4, False, abi.SAMPLE_NEE_MIS, (8, 8), 3,
                       exr_prefix=prefix)
    path = prefix + "_00000004wpp.exr"
    assert os.path.exists(path), os.listdir(tmp_path)
    _, ch = read_exr(path)
    img = film_image(a)
    for k, c in enumerate("RGB"):
        assert np.allclose(ch[c], img[k][::-1], rtol=1e-5, atol=1e-7), c
    b = _cabi_render(scene, abi.render_params(48, 32, 4, rr=(8, 8), mode=abi.SAMPLE_NEE_MIS,
                                              seed=3))
    assert relmse(img, film_image(b)) < 1e-6


def _tiled_cabi(scene, params, passes):
    from mray_amd import native
    acc = native.Accel(scene)
    rnd = native.Renderer(acc, params)
    for n in passes:
        rnd.render_tile(n)
    film = rnd.film()
    tiles = rnd.tiling()
    rnd.close()
    acc.close()
    return film, tiles


@pytest.mark.parametrize("render_mode,burst", [("Latency", 4), ("Throughput", 2)])
def test_tracer_api_tile_by_tile(torch_gpu, tmp_path, render_mode, burst):
    """A parallelization hint below the image size splits it into tiles
    (ImageTiler); Latency renders one sample of one tile per DoRenderWork,
    Throughput burstSize samples (PathTracerRendererBase::DoRender).  The
    consumer's image, assembled from the tile sections, equals the C-ABI's
    tiled renderer driven with the same calls."""
    W, H, spp, hint = 90, 50, 4, 1200
    scene = scenes.cornell(aspect=W / H)
    a = _driver_render(scene, tmp_path, W, H, spp, False, abi.SAMPLE_NEE_MIS, (8, 8), 2,
                       render_mode=render_mode, burst=burst, hint=hint)
    params = abi.render_params(W, H, spp, rr=(8, 8), mode=abi.SAMPLE_NEE_MIS, seed=2,
                               max_ray_count=hint, flags=abi.RENDER_TILED)
    per_call = 1 if render_mode == "Latency" else burst
    from mray_amd import native
    acc = native.Accel(scene)
    rnd = native.Renderer(acc, params)
    (tw, th), (cx, cy) = rnd.tiling()
    rnd.close()
    acc.close()
    assert cx * cy > 1
    passes = [per_call] * (cx * cy * (spp // per_call))
    b, _ = _tiled_cabi(scene, params, passes)
    assert np.allclose(a[3], float(spp)) and np.allclose(b[3], float(spp))
    e = relmse(film_image(a), film_image(b))
    print(f"{render_mode} burst {burst}: tiles {tw}x{th} x {cx}x{cy}, relMSE TracerI vs C-ABI "
          f"{e:.3e}", flush=True)
    assert e < 1e-8
