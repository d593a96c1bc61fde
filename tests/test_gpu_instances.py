"""Two-level scenes (mray_instanced_scene_create) on the GPU.

An instanced scene renders the world its instances describe, so its image at
equal samples matches the oracle's render of the flattened scene (every
instance's triangles moved to world space, lights once per instance, in
instance order): the same slot pool, path indices and PCG32 streams run on
both sides, and only the ray's trip into object space and back adds float
rounding, which moves a few paths.  The bound is the batch renderer's:
relMSE <= 1e-3 (eps 1e-2) with exact path counts and filter weights.

A single identity instance walks the same mesh tree with the same ray (the
inverse rows are exactly the identity), so it equals the one-level renderer
up to float-atomic order.
"""
import math

import numpy as np
import pytest

from mray_amd import abi, scenes
from oracle import refcpu
from tests.conftest import film_image, relmse

pytestmark = pytest.mark.gpu


def _gpu_instanced(meshes, instances, env, params):
    from mray_amd import native
    accs = [native.Accel(m) for m in meshes]
    isc = native.InstancedScene(accs, instances, env)
    for a in accs:
        a.close()      # the instanced scene holds its own copies
    info = isc.info()
    r = native.InstancedRenderer(isc, params)
    st = r.render_all()
    film = r.film()
    r.close()
    isc.close()
    return film, st, info


@pytest.mark.parametrize("spectral,rotate", [(False, False), (True, False), (False, True)])
def test_instanced_matches_flattened_oracle(torch_gpu, spectral, rotate):
    """Shading frames kept in mesh space (the reference's behaviour) or
    rotated with each instance (MRAY_INST_ROTATE_FRAMES)."""
    from mray_amd import native
    if spectral:
        refcpu.spectral_load(native.spectral_tables())
    meshes, inst, env, flat = scenes.instanced_test_scene(width=64, height=48, rotate_frames=rotate)
    params = abi.render_params(64, 48, 4, rr=(3, 8), seed=5, spectral=spectral, max_ray_count=4096)
    film, st, info = _gpu_instanced(meshes, inst, env, params)
    assert info["instances"] == len(inst)
    assert info["effective_triangles"] == flat.triangle_count
    assert info["lights"] == len(flat.light_tris)
    ofilm, ost = refcpu.OracleScene(flat).render(params)
    print(f"gpu={st} cpu={ost}", flush=True)
    assert st["dead_paths"] == ost["dead_paths"] == 64 * 48 * 4
    assert np.allclose(film[3], ofilm[3], rtol=1e-5)
    e = relmse(film_image(film), film_image(ofilm))
    print(f"spectral={spectral} rotate={rotate}: relMSE={e:.3e}", flush=True)
    assert e <= 1e-3


def test_instanced_matches_flattened_gpu(torch_gpu):
    """Against the one-level GPU renderer of the flattened scene at a larger
    image: the same samples through the two traversals.  Frames rotate with
    the instances here: with the reference's mesh-space frames, an
    upside-down instance shades around a normal that points into its own
    surface, and such paths turn the last-bit differences of the two
    geometric normals (rotated vs. from world vertices) into different
    paths, so that mode is compared with the oracle at a smaller size."""
    from mray_amd import native
    meshes, inst, env, flat = scenes.instanced_test_scene(width=160, height=120, rotate_frames=True)
    params = abi.render_params(160, 120, 8, rr=(3, 8), seed=9)
    film, st, _ = _gpu_instanced(meshes, inst, env, params)
    acc = native.Accel(flat)
    r = native.Renderer(acc, params)
    fst = r.render_all()
    ffilm = r.film()
    r.close()
    acc.close()
    assert st["dead_paths"] == fst["dead_paths"]
    e = relmse(film_image(film), film_image(ffilm))
    print(f"instanced vs flat GPU: relMSE={e:.3e} ext {st['ext_rays']} vs {fst['ext_rays']}", flush=True)
    assert e <= 1e-3
    assert abs(st["ext_rays"] - fst["ext_rays"]) <= 1e-3 * fst["ext_rays"]


def test_identity_instance_equals_one_level(torch_gpu):
    from mray_amd import native
    sc = scenes.terrain_spheres(grid=48, spheres=20, width=64, height=40, seed=7)
    env = {"boundary_type": sc.boundary_type, "boundary_radiance": sc.boundary_radiance,
           "scene_diameter": sc.diameter, "camera": sc.camera}
    ident = [{"mesh": 0, "rotation": (1.0, 0.0, 0.0, 0.0), "scale": 1.0,
              "translation": (0.0, 0.0, 0.0)}]
    params = abi.render_params(64, 40, 4, rr=(3, 8), seed=3, max_ray_count=2000)
    film, st, info = _gpu_instanced([sc], ident, env, params)
    assert info["tlas_nodes"] == 1
    acc = native.Accel(sc)
    r = native.Renderer(acc, params)
    fst = r.render_all()
    ffilm = r.film()
    r.close()
    acc.close()
    assert st["ext_rays"] == fst["ext_rays"] and st["shadow_rays"] == fst["shadow_rays"]
    assert np.array_equal(film[3], ffilm[3])
    e = relmse(film_image(film), film_image(ffilm))
    print(f"identity instance vs one level: relMSE={e:.3e}", flush=True)
    assert e <= 1e-8


def test_errors(torch_gpu):
    from mray_amd import native
    meshes, inst, env, _ = scenes.instanced_test_scene(width=16, height=16)
    accs = [native.Accel(m) for m in meshes]
    bad_mesh = [dict(inst[0], mesh=7)]
    with pytest.raises(native.MRayError, match="missing mesh"):
        native.InstancedScene(accs, bad_mesh, env)
    with pytest.raises(native.MRayError, match="non-positive scale"):
        native.InstancedScene(accs, [dict(inst[0], scale=0.0)], env)
    with pytest.raises(native.MRayError, match="zero rotation"):
        native.InstancedScene(accs, [dict(inst[0], rotation=(0.0, 0.0, 0.0, 0.0))], env)
    a8 = native.Accel(meshes[0], width=8)
    with pytest.raises(native.MRayError, match="4-wide"):
        native.InstancedScene([a8], [dict(inst[0], mesh=0)], env)
    a8.close()
    isc = native.InstancedScene(accs, inst, env)
    q = abi.render_params(16, 16, 1, rr=(3, 8), seed=1, flags=abi.RENDER_MATERIAL_QUEUES)
    with pytest.raises(native.MRayError, match="material queues"):
        native.InstancedRenderer(isc, q)
    isc.close()
    for a in accs:
        a.close()


def test_many_instances_of_one_mesh(torch_gpu):
    """A 10 x 10 grid of one mesh, rotated and scaled: the instance tree has
    several levels, and the image matches the flattened oracle's."""
    mesh = scenes.terrain_spheres(grid=12, spheres=3, width=48, height=32, light_grid=0, seed=4)
    lights = scenes.light_panel(4, 0.3, (6.0, 6.0, 6.0), seed=1, span=2.0)
    inst = []
    for i in range(10):
        for j in range(10):
            ang = 0.37 * (i * 10 + j)
            inst.append({"mesh": 0, "rotation": (math.cos(ang / 2), 0.0, math.sin(ang / 2), 0.0),
                         "scale": 0.25 + 0.01 * j, "translation": (-2.5 + 0.5 * i, 0.0, -2.5 + 0.5 * j)})
    inst.append({"mesh": 1, "rotation": (1.0, 0.0, 0.0, 0.0), "scale": 1.0,
                 "translation": (0.0, 2.0, 0.0)})
    fov = math.radians(55.0)
    cam = {"position": (0.0, 3.0, 4.5), "gaze": (0.0, 0.0, 0.0), "up": (0.0, 1.0, 0.0),
           "fov": (fov, scenes.fov_y_from_x(fov, 48 / 32)), "nearFar": (0.01, 100.0)}
    env = {"boundary_type": abi.BOUNDARY_SKYSPHERE_COOCTA, "boundary_radiance": (0.2, 0.25, 0.3),
           "scene_diameter": 8.0, "camera": cam}
    flat = scenes.flatten_instances([mesh, lights], inst, env)
    params = abi.render_params(48, 32, 4, rr=(3, 8), seed=12, max_ray_count=3000)
    film, st, info = _gpu_instanced([mesh, lights], inst, env, params)
    assert info["tlas_nodes"] > 1
    ofilm, ost = refcpu.OracleScene(flat).render(params)
    assert st["dead_paths"] == ost["dead_paths"]
    e = relmse(film_image(film), film_image(ofilm))
    print(f"100 instances: relMSE={e:.3e}", flush=True)
    assert e <= 1e-3


def _views():
    cams = []
    for pos, fov in (((2.0, 3.0, 5.0), 50.0), ((-4.0, 2.5, 1.0), 70.0),
                     ((0.5, 6.0, 0.5), 60.0)):
        cams.append({"position": pos, "gaze": (0.0, 0.0, 0.0), "up": (0.0, 1.0, 0.0),
                     "fov": (math.radians(fov),) * 2, "nearFar": (0.01, 100.0)})
    return [dict(scene=0, width=48, height=36, spp=3),
            dict(scene=0, width=40, height=40, spp=2, camera=cams[0]),
            dict(scene=0, width=17, height=9, spp=5, camera=cams[1]),
            dict(scene=0, width=32, height=24, spp=4, camera=cams[2])]


@pytest.mark.parametrize("spectral,rotate", [(False, False), (True, True)])
def test_instanced_batch_matches_flattened_oracle(torch_gpu, spectral, rotate):
    """Several views of one instanced scene in one slot pool, against the
    oracle's batch render of the flattened scene."""
    from mray_amd import native
    if spectral:
        refcpu.spectral_load(native.spectral_tables())
    meshes, inst, env, flat = scenes.instanced_test_scene(width=48, height=36, rotate_frames=rotate)
    views = _views()
    params = abi.render_params(1, 1, 1, rr=(3, 8), seed=13, spectral=spectral, max_ray_count=3000)
    accs = [native.Accel(m) for m in meshes]
    isc = native.InstancedScene(accs, inst, env)
    for a in accs:
        a.close()
    r = native.BatchRenderer(isc, views, params)
    st = r.render_all()
    films = r.film()
    times = r.view_times()
    r.close()
    isc.close()
    ofilms, ost = refcpu.render_batch([refcpu.OracleScene(flat)], views, params)
    print(f"gpu={st} cpu={ost}", flush=True)
    assert st["dead_paths"] == ost["dead_paths"] == sum(v["width"] * v["height"] * v["spp"]
                                                        for v in views)
    for k, (f, o) in enumerate(zip(films, ofilms)):
        assert np.allclose(f[3], o[3], rtol=1e-5)
        e = relmse(film_image(f), film_image(o))
        print(f"view {k}: relMSE={e:.3e}", flush=True)
        assert e <= 1e-3
    assert np.all(times >= 0.0)


def test_instanced_batch_errors(torch_gpu):
    from mray_amd import native
    meshes, inst, env, _ = scenes.instanced_test_scene(width=16, height=16)
    accs = [native.Accel(m) for m in meshes]
    isc = native.InstancedScene(accs, inst, env)
    params = abi.render_params(1, 1, 1, rr=(3, 8), seed=1)
    with pytest.raises(native.MRayError, match="scene 0"):
        native.BatchRenderer(isc, [dict(scene=1, width=8, height=8, spp=1)], params)
    q = abi.render_params(1, 1, 1, rr=(3, 8), seed=1, flags=abi.RENDER_MATERIAL_QUEUES)
    with pytest.raises(native.MRayError, match="material queues"):
        native.BatchRenderer(isc, [dict(scene=0, width=8, height=8, spp=1)], q)
    r = native.BatchRenderer(isc, [dict(scene=0, width=8, height=8, spp=1)], params)
    with pytest.raises(native.MRayError, match="merged scene"):
        r.set_trace_tpb(128)
    r.close()
    isc.close()
    for a in accs:
        a.close()


def test_identity_instance_deep_stack(torch_gpu):
    """The two-level walk with stacks past the 12 LDS levels: one identity
    instance of 4096 parallel planes seen edge-on to their stack (camera on
    the stack's axis, constant sky).  Its walks hold the instance tree's
    entries, the return mark and 18 or more mesh-tree entries, so they
    spill to the global area; the image and counts equal the one-level
    renderer's on the same mesh."""
    from mray_amd import native
    from tests.test_gpu_traversal import _plane_stack
    sc = _plane_stack(4096)
    sc.boundary_type = abi.BOUNDARY_SKYSPHERE_COOCTA
    sc.boundary_radiance = (0.5, 0.5, 0.5)
    env = {"boundary_type": sc.boundary_type, "boundary_radiance": sc.boundary_radiance,
           "scene_diameter": sc.diameter, "camera": sc.camera}
    ident = [{"mesh": 0, "rotation": (1.0, 0.0, 0.0, 0.0), "scale": 1.0,
              "translation": (0.0, 0.0, 0.0)}]
    params = abi.render_params(48, 32, 2, rr=(3, 6), seed=4, max_ray_count=1536)
    film, st, info = _gpu_instanced([sc], ident, env, params)
    acc = native.Accel(sc)
    r = native.Renderer(acc, params)
    fst = r.render_all()
    ffilm = r.film()
    r.close()
    acc.close()
    print(f"deep stack: instanced {st} one level {fst}", flush=True)
    assert st["dead_paths"] == fst["dead_paths"] == 48 * 32 * 2
    assert st["ext_rays"] == fst["ext_rays"] and st["shadow_rays"] == fst["shadow_rays"]
    assert film[:3].any()
    assert np.array_equal(film[3], ffilm[3])
    e = relmse(film_image(film), film_image(ffilm))
    print(f"deep stack identity instance vs one level: relMSE={e:.3e}", flush=True)
    assert e <= 1e-8
