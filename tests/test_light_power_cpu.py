"""NEE light selection by power (MRAY_RENDER_LIGHT_POWER) in the oracle.

The alias table is pinned by what it must encode: bucket shares and
aliases add up to each light's probability n/(n+1) * w_i / sum(w) with
w = Y(radiance) * area * (2 if two-sided), equal lights give the uniform
sampler's 1/(n+1), and a render with power selection has the same
expectation as the uniform sampler's (image means agree) while its noise
on a scene of very uneven lights is lower.
"""
import numpy as np

from mray_amd import abi, scenes
from oracle import refcpu
from tests.conftest import film_image, relmse

Y_ROW = np.array([0.2126390058715103, 0.7151686787677559, 0.0721923153607337])


def _uneven(seed=3, width=16, height=12):
    sc = scenes.terrain_spheres(grid=24, spheres=6, width=width, height=height, light_grid=4,
                                seed=2)
    rng = np.random.default_rng(seed)
    per_quad = rng.uniform(0.0, 1.0, len(sc.light_tris) // 2) ** 4 * 40.0
    sc.light_radiance = (sc.light_radiance * np.repeat(per_quad, 2)[:, None]).astype(np.float32)
    two = np.zeros(len(sc.light_tris), np.uint8)
    two[::3] = 1
    sc.light_two_sided = two
    return sc


def _areas(sc):
    p = sc.positions[sc.indices[sc.light_tris]].astype(np.float64)
    return 0.5 * np.linalg.norm(np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0]), axis=1)


def test_table_encodes_the_power_distribution():
    sc = _uneven()
    o = refcpu.OracleScene(sc)
    prob, alias, pmf = o.light_power_table()
    n = len(sc.light_tris)
    w = np.abs(sc.light_radiance.astype(np.float64) @ Y_ROW) * _areas(sc) * \
        np.where(sc.light_two_sided != 0, 2.0, 1.0)
    np.testing.assert_allclose(pmf, n / (n + 1) * w / w.sum(), rtol=1e-6)
    implied = np.zeros(n)
    for i in range(n):
        implied[i] += prob[i]
        implied[alias[i]] += 1.0 - prob[i]
    np.testing.assert_allclose(implied / (n + 1), pmf, rtol=1e-5, atol=1e-9)
    assert np.all((prob >= 0) & (prob <= 1)) and np.all(alias < n)


def test_equal_lights_reduce_to_uniform():
    sc = scenes.terrain_spheres(grid=24, spheres=6, width=16, height=12, light_grid=4, seed=2)
    prob, alias, pmf = refcpu.OracleScene(sc).light_power_table()
    n = len(sc.light_tris)
    np.testing.assert_allclose(pmf, 1.0 / (n + 1), rtol=1e-5)
    np.testing.assert_allclose(prob, 1.0, atol=1e-5)


def _lit_floor(seed=5):
    """A grey Lambert floor under 16 small downward quad lights whose
    radiance spans four orders of magnitude; no boundary light."""
    b = scenes._Builder()
    b.add_mesh(*scenes._quad((0.0, 0.0, 0.0), (0.0, 0.0, 2.0), (2.0, 0.0, 0.0)), key=0, cull=0)
    pos, tbn, idx, key, cull = b.finish()
    floor = scenes.Scene(positions=pos, tbn=tbn, indices=idx, surface_keys=key, cull=cull,
                         materials=[{"type": abi.MAT_LAMBERT, "albedo": (0.6, 0.6, 0.6)}],
                         light_tris=np.zeros(0, np.uint32), light_radiance=np.zeros((0, 3), np.float32),
                         light_two_sided=np.zeros(0, np.uint8), camera={})
    panel = scenes.light_panel(16, 0.1, (1.0, 1.0, 1.0), seed=seed, span=1.5)
    cam = {"position": (0.0, 2.5, 2.5), "gaze": (0.0, 0.0, 0.0), "up": (0.0, 1.0, 0.0),
           "fov": (np.radians(50.0),) * 2, "nearFar": (0.01, 100.0)}
    env = {"boundary_type": abi.BOUNDARY_NULL, "boundary_radiance": (0, 0, 0),
           "scene_diameter": 6.0, "camera": cam}
    ident = (1.0, 0.0, 0.0, 0.0)
    sc = scenes.flatten_instances([floor, panel], [
        {"mesh": 0, "rotation": ident, "scale": 1.0, "translation": (0, 0, 0)},
        {"mesh": 1, "rotation": ident, "scale": 1.0, "translation": (0, 1.0, 0)}], env)
    rng = np.random.default_rng(seed)
    power = 10.0 ** rng.uniform(-2.0, 2.0, 16)
    sc.light_radiance = (np.repeat(power, 2)[:, None] * np.ones(3)).astype(np.float32)
    return sc


def test_power_selection_is_unbiased_and_less_noisy():
    """On the floor under uneven lights: the power-sampled image mean equals
    the uniform sampler's converged mean (within 1 %), and at equal samples
    its relMSE is less than half the uniform sampler's."""
    sc = _lit_floor()
    o = refcpu.OracleScene(sc)

    def render(flags, spp, seed):
        p = abi.render_params(24, 16, spp, rr=(3, 4), seed=seed, flags=flags)
        return film_image(o.render(p)[0]).astype(np.float64)

    ref = render(0, 4096, 99)
    pw_long = render(abi.RENDER_LIGHT_POWER, 1024, 2)
    rel = np.abs(pw_long.mean(axis=(1, 2)) - ref.mean(axis=(1, 2))) / ref.mean(axis=(1, 2))
    assert np.all(rel < 0.01), rel
    e_uni = relmse(render(0, 32, 1), ref)
    e_pw = relmse(render(abi.RENDER_LIGHT_POWER, 32, 1), ref)
    assert e_pw < 0.5 * e_uni, (e_pw, e_uni)
