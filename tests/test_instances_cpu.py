"""Flattening of instanced scenes (scenes.flatten_instances), the oracle's
view of a two-level scene, pinned by what geometry fixes:

* an identity instance is the mesh itself;
* by default a transformed instance keeps its mesh's shading frames, as the
  reference's transformed triangles do (PrimitiveDefaultTriangle.hpp:611-614);
* with rotated frames, a similarity (rotation, uniform scale, translation)
  applied to a whole scene and its camera leaves the image unchanged: radiance is invariant
  under it, and the camera rays, path indices and PCG32 streams are the
  same, so only rounding moves paths.  A mistake in the positions' or the
  tangent frames' transform (wrong quaternion order, a missing conjugate,
  scale applied to frames) turns the surfaces or the sampled directions and
  fails this;
* instance order fixes the light order: lights appear once per instance,
  each block in its mesh's order.
"""
import math

import numpy as np

from mray_amd import abi, scenes
from oracle import refcpu
from tests.conftest import film_image, relmse


def _env_of(sc):
    return {"boundary_type": sc.boundary_type, "boundary_radiance": sc.boundary_radiance,
            "scene_diameter": sc.diameter, "camera": sc.camera}


def test_identity_instance_is_the_mesh():
    sc = scenes.cornell()
    flat = scenes.flatten_instances([sc], [{"mesh": 0, "rotation": (1, 0, 0, 0), "scale": 1.0,
                                            "translation": (0, 0, 0)}], _env_of(sc))
    assert np.array_equal(flat.positions, sc.positions)
    assert np.allclose(flat.tbn, sc.tbn, atol=1e-7)
    assert np.array_equal(flat.indices, sc.indices)
    assert np.array_equal(flat.surface_keys, sc.surface_keys)
    assert np.array_equal(flat.light_tris, sc.light_tris)


def test_light_order_follows_instances():
    a = scenes.light_panel(2, 0.1, (1.0, 2.0, 3.0), seed=1)
    b = scenes.light_panel(3, 0.1, (4.0, 5.0, 6.0), seed=2)
    ident = (1.0, 0.0, 0.0, 0.0)
    inst = [{"mesh": 1, "rotation": ident, "scale": 1.0, "translation": (0, 0, 0)},
            {"mesh": 0, "rotation": ident, "scale": 1.0, "translation": (5, 0, 0)},
            {"mesh": 1, "rotation": ident, "scale": 2.0, "translation": (0, 5, 0)}]
    env = {"boundary_type": abi.BOUNDARY_NULL, "boundary_radiance": (0, 0, 0), "camera": {}}
    flat = scenes.flatten_instances([a, b], inst, env)
    rad = flat.light_radiance
    assert len(rad) == 6 + 4 + 6
    assert np.all(rad[:6] == (4, 5, 6)) and np.all(rad[6:10] == (1, 2, 3)) and np.all(rad[10:] == (4, 5, 6))
    # each light key names its own light, whose triangle carries that key
    lk = flat.surface_keys[flat.light_tris]
    assert np.array_equal(lk & ~np.uint32(abi.LIGHT_FLAG), np.arange(16, dtype=np.uint32))
    # the third instance's triangles: mesh 1's scaled by 2 and lifted by 5
    n1 = len(b.positions)
    off = n1 + len(a.positions)
    assert np.allclose(flat.positions[off:off + n1], b.positions * 2 + np.float32([0, 5, 0]))


def _similar(sc, q, s, t):
    """The scene as one instance under (q, s, t), with the camera carried along."""
    qa = np.asarray(q, np.float64) / np.linalg.norm(q)
    cam = dict(sc.camera)
    cam["position"] = tuple(scenes.quat_rotate(qa, cam["position"]) * s + t)
    cam["gaze"] = tuple(scenes.quat_rotate(qa, cam["gaze"]) * s + t)
    cam["up"] = tuple(scenes.quat_rotate(qa, cam["up"]))
    cam["nearFar"] = (cam["nearFar"][0] * s, cam["nearFar"][1] * s)
    env = {"boundary_type": sc.boundary_type, "boundary_radiance": sc.boundary_radiance,
           "scene_diameter": sc.diameter * s, "camera": cam, "rotate_frames": True}
    return scenes.flatten_instances([sc], [{"mesh": 0, "rotation": q, "scale": s, "translation": t}], env)


def test_default_frames_stay_in_mesh_space():
    sc = scenes.cornell()
    q = (np.cos(0.4), 0.0, np.sin(0.4), 0.0)
    inst = [{"mesh": 0, "rotation": q, "scale": 1.5, "translation": (1.0, 0.0, 0.0)}]
    flat = scenes.flatten_instances([sc], inst, _env_of(sc))
    assert np.array_equal(flat.tbn, sc.tbn)
    rot = scenes.flatten_instances([sc], inst, dict(_env_of(sc), rotate_frames=True))
    assert not np.allclose(rot.tbn, sc.tbn, atol=1e-3)


def test_similarity_leaves_the_image_unchanged():
    """Cornell box (every pixel sees a shaded surface, no boundary light):
    the moved scene's image is the original's up to rounding (relMSE about
    5e-5 here); frames left unrotated or composed in the other order
    (r * tbn) give 0.07 and more."""
    sc = scenes.cornell()
    p = abi.render_params(32, 32, 8, rr=(3, 8), seed=7, max_ray_count=2400)
    ref, rst = refcpu.OracleScene(sc).render(p)
    ang = 1.1
    ax = np.array([0.3, 0.8, -0.5]) / np.linalg.norm([0.3, 0.8, -0.5])
    q = (math.cos(ang / 2), *(ax * math.sin(ang / 2)))
    moved = _similar(sc, q, 1.7, (3.0, -2.0, 0.5))
    img, st = refcpu.OracleScene(moved).render(p)
    e = relmse(film_image(img), film_image(ref))
    assert st["dead_paths"] == rst["dead_paths"]
    assert e <= 1e-3, e
    for how in ("unrotated", "reversed"):
        wrong = _similar(sc, q, 1.7, (3.0, -2.0, 0.5))
        if how == "unrotated":
            wrong.tbn = sc.tbn.copy()
        else:
            tb = scenes.quat_mul(np.asarray(q) / np.linalg.norm(q), sc.tbn.astype(np.float64))
            wrong.tbn = (tb / np.linalg.norm(tb, axis=1, keepdims=True)).astype(np.float32)
        bad, _ = refcpu.OracleScene(wrong).render(p)
        assert relmse(film_image(bad), film_image(ref)) > 2e-2, how
