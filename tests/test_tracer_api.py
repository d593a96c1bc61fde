"""TracerI host layer (include/mray_tracer.hpp, mray_amd/host/tracer.cpp).

A C++ front-end (tests/cpp/tracer_driver) loads libmray_tracer.so with dlopen,
constructs the tracer through ConstructTracer and feeds it a scene with the
calls the reference's scene loader makes (groups, reservations, commits,
TransientData attribute pushes, surfaces, light surfaces, boundary light,
camera surface).  Without a GPU the driver stops after CommitSurfaces' host
half and writes the flattened scene, which must reproduce the input scene
triangle for triangle.  The error checks mirror the reference's MRayError
behaviour (TracerBase.cpp:1713-1717, PathTracerRenderer.cu:139-140).
"""
import os
import subprocess

import numpy as np
import pytest

from mray_amd import abi, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "tracer_driver")
TRACER_LIB = os.path.join(ROOT, "mray_amd", "libmray_tracer.so")


def _built():
    if not (os.path.exists(DRIVER) and os.path.exists(TRACER_LIB)):
        pytest.fail("libmray_tracer.so / tests/cpp/tracer_driver are not built; "
                    "run __graft_entry__.build()")


def write_scene_file(scene, path):
    """The driver's input: the MRaySceneDesc arrays plus the camera."""
    d = scene.desc()
    mats = np.frombuffer(bytes(memoryview(scene._keep[1]))[:52 * len(scene.materials)], np.uint8)
    cam = scene.camera
    with open(path, "wb") as f:
        f.write(np.array([len(scene.positions), scene.triangle_count, len(scene.materials),
                          len(scene.light_tris), scene.boundary_type], np.uint32).tobytes())
        f.write(np.array(list(scene.boundary_radiance) + [d.sceneDiameter], np.float32).tobytes())
        f.write(np.ascontiguousarray(scene.positions, np.float32).tobytes())
        f.write(np.ascontiguousarray(scene.tbn, np.float32).tobytes())
        f.write(np.ascontiguousarray(scene.indices, np.uint32).tobytes())
        f.write(np.ascontiguousarray(scene.surface_keys, np.uint32).tobytes())
        f.write(np.ascontiguousarray(scene.cull, np.uint8).tobytes())
        f.write(mats.tobytes())
        f.write(np.ascontiguousarray(scene.light_tris, np.uint32).tobytes())
        f.write(np.ascontiguousarray(scene.light_radiance, np.float32).tobytes())
        f.write(np.ascontiguousarray(scene.light_two_sided, np.uint8).tobytes())
        f.write(np.array(list(cam["position"]) + list(cam["gaze"]) + list(cam["up"]) +
                         list(cam["fov"]) + list(cam["nearFar"]), np.float32).tobytes())


def read_flat_file(path):
    b = open(path, "rb").read()
    o = 0

    def take(dtype, n):
        nonlocal o
        a = np.frombuffer(b, dtype, count=n, offset=o)
        o += a.nbytes
        return a
    V, T, M, L, btype = take(np.uint32, 5)
    out = {"btype": int(btype), "brad": take(np.float32, 3), "diam": take(np.float32, 1)[0]}
    out["pos"] = take(np.float32, V * 3).reshape(-1, 3)
    out["tbn"] = take(np.float32, V * 4).reshape(-1, 4)
    out["idx"] = take(np.uint32, T * 3).reshape(-1, 3)
    out["keys"] = take(np.uint32, T)
    out["cull"] = take(np.uint8, T)
    out["mats"] = take(np.uint8, M * 52)
    out["ltri"] = take(np.uint32, L)
    out["lrad"] = take(np.float32, L * 3).reshape(-1, 3)
    out["l2s"] = take(np.uint8, L)
    assert o == len(b)
    return out


def _flatten(scene, tmp_path):
    _built()
    src = str(tmp_path / "scene.bin")
    dst = str(tmp_path / "flat.bin")
    write_scene_file(scene, src)
    r = subprocess.run([DRIVER, TRACER_LIB, src, "flatten", dst], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    return read_flat_file(dst)


def _check_same(scene, f):
    idx = np.asarray(scene.indices)
    assert np.array_equal(f["pos"][f["idx"]], np.asarray(scene.positions, np.float32)[idx])
    assert np.array_equal(f["tbn"][f["idx"]], np.asarray(scene.tbn, np.float32)[idx])
    assert np.array_equal(f["keys"], scene.surface_keys)
    assert np.array_equal(f["cull"], scene.cull)
    mats = np.frombuffer(bytes(memoryview(scene._keep[1]))[:52 * len(scene.materials)], np.uint8)
    assert np.array_equal(f["mats"], mats)
    assert np.array_equal(f["ltri"], scene.light_tris)
    assert np.array_equal(f["lrad"], np.asarray(scene.light_radiance, np.float32))
    assert np.array_equal(f["l2s"], scene.light_two_sided)
    assert f["btype"] == scene.boundary_type
    assert np.array_equal(f["brad"], np.asarray(scene.boundary_radiance, np.float32))
    # the boundary light's diameter follows TracerBase.cpp:1651-1664
    assert f["diam"] == np.float32(scene.diameter)


def test_flatten_cornell(tmp_path):
    s = scenes.cornell()
    _check_same(s, _flatten(s, tmp_path))


def test_flatten_terrain_with_skysphere(tmp_path):
    s = scenes.terrain_spheres(grid=48, spheres=40, width=64, height=36)
    assert s.boundary_type == abi.BOUNDARY_SKYSPHERE_COOCTA
    _check_same(s, _flatten(s, tmp_path))


def test_flatten_spherical_skysphere(tmp_path):
    """A (L)Skysphere_Spherical boundary flattens to the latitude-longitude
    boundary type, through direct TracerI calls and through a JSON file."""
    s = scenes.terrain_spheres(grid=16, spheres=4, width=32, height=18)
    s.boundary_type = abi.BOUNDARY_SKYSPHERE_SPHERICAL
    _check_same(s, _flatten(s, tmp_path))
    (tmp_path / "json").mkdir()
    _check_same_json(s, _flatten_json(s, tmp_path / "json"))


def test_error_behaviour():
    _built()
    r = subprocess.run([DRIVER, TRACER_LIB, "errors"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout


def _flatten_json(scene, tmp_path):
    _built()
    src = str(tmp_path / "scene.json")
    dst = str(tmp_path / "flat.bin")
    scenes.to_mray_json(scene, src)
    r = subprocess.run([DRIVER, TRACER_LIB, src, "flatten", dst], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stderr
    return read_flat_file(dst)


def _check_same_json(scene, f):
    """As _check_same, with TBN quaternions rebuilt from the written normals
    (so equal up to float rounding) instead of copied."""
    scene.desc()
    idx = np.asarray(scene.indices)
    assert np.array_equal(f["pos"][f["idx"]], np.asarray(scene.positions, np.float32)[idx])
    assert np.allclose(f["tbn"][f["idx"]], np.asarray(scene.tbn, np.float32)[idx], atol=2e-6)
    assert np.array_equal(f["keys"], scene.surface_keys)
    assert np.array_equal(f["cull"], scene.cull)
    mats = np.frombuffer(bytes(memoryview(scene._keep[1]))[:52 * len(scene.materials)], np.uint8)
    got = f["mats"].view(np.uint32).reshape(-1, 13)
    want = mats.view(np.uint32).reshape(-1, 13)
    assert np.array_equal(got[:, 0], want[:, 0])                       # types
    assert np.allclose(got[:, 1:].view(np.float32), want[:, 1:].view(np.float32), rtol=1e-6)
    assert np.array_equal(f["ltri"], scene.light_tris)
    assert np.allclose(f["lrad"], np.asarray(scene.light_radiance, np.float32))
    assert np.array_equal(f["l2s"], scene.light_two_sided)
    assert f["btype"] == scene.boundary_type
    assert np.allclose(f["brad"], np.asarray(scene.boundary_radiance, np.float32))


def test_json_scene_cornell(tmp_path):
    """MRay JSON scene file -> LoadSceneMRay -> TracerI -> the same scene."""
    s = scenes.cornell()
    _check_same_json(s, _flatten_json(s, tmp_path))


def test_json_scene_terrain(tmp_path):
    s = scenes.terrain_spheres(grid=40, spheres=30, width=64, height=36)
    _check_same_json(s, _flatten_json(s, tmp_path))


def test_json_scene_transforms_and_errors(tmp_path):
    """A hand-written file: a unit quad placed by a 'trs' transform and by a
    'matrix' transform, both of which the flattened positions must show; and
    an unknown material type must throw."""
    import json
    quad = {"id": 0, "type": "Triangle", "tag": "nodeTriangleIndexed",
            "position": [[-0.5, -0.5, 0], [0.5, -0.5, 0], [0.5, 0.5, 0], [-0.5, 0.5, 0]],
            "index": [[0, 1, 2], [0, 2, 3]]}
    doc = {
        "Cameras": [{"id": 0, "type": "Pinhole", "isFovX": True, "fov": 40.0, "aspect": 1.0,
                     "planes": [0.1, 100.0], "gaze": [0, 0, 0], "position": [0, 0, 4], "up": [0, 1, 0]}],
        "Lights": [{"id": 0, "type": "Null"}],
        "Mediums": [], "Textures": [],
        "Transforms": [{"id": 0, "type": "Identity"},
                       {"id": 1, "type": "Single", "layout": "trs", "translate": [1, 2, 3],
                        "rotate": [0, 0, 90], "scale": [2, 2, 2]},
                       {"id": 2, "type": "Single", "layout": "matrix",
                        "matrix": [1, 0, 0, 5, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1]},
                       {"id": 3, "type": "Single", "layout": "trs", "translate": [0, 0, 0],
                        "rotate": [0, 0, 0], "scale": [1, 2, 3]}],
        "Materials": [{"id": [0, 1], "type": "Lambert", "albedo": [[0.5, 0.5, 0.5], [0.2, 0.3, 0.4]]}],
        "Primitives": [quad],
        "Surfaces": [{"transform": 1, "material": 0, "primitive": 0},
                     {"transform": 2, "material": 1, "primitive": 0, "cullBackFace": False},
                     {"transform": 3, "material": 0, "primitive": 0}],
        "LightSurfaces": [], "CameraSurfaces": [{"camera": 0}],
        "Boundary": {"light": 0},
    }
    src = str(tmp_path / "t.json")
    with open(src, "w") as fh:
        json.dump(doc, fh)
    dst = str(tmp_path / "flat.bin")
    r = subprocess.run([DRIVER, TRACER_LIB, src, "flatten", dst], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    f = read_flat_file(dst)
    p = f["pos"][f["idx"]]
    # trs: scale 2, rotate 90 deg about Z, then translate (1,2,3): (-0.5,-0.5,0) -> (2,1,3)
    assert np.allclose(p[0, 0], [2.0, 1.0, 3.0], atol=1e-5)
    # matrix: translate x by 5
    assert np.allclose(p[2, 0], [4.5, -0.5, 0.0], atol=1e-6)
    # trs with a non-uniform scale moves positions only; shading frames stay
    # the mesh's, as for every surface transform
    assert np.allclose(p[4, 0], [-0.5, -1.0, 0.0], atol=1e-6)
    assert np.allclose(f["tbn"][f["idx"][4]], f["tbn"][f["idx"][2]], atol=1e-7)
    assert list(f["cull"]) == [1, 1, 0, 0, 1, 1]
    assert list(f["keys"]) == [0, 0, 1, 1, 0, 0]
    doc["Materials"][0]["type"] = "Glass"
    with open(src, "w") as fh:
        json.dump(doc, fh)
    r = subprocess.run([DRIVER, TRACER_LIB, src, "flatten", dst], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0 and "Unknown material type" in r.stderr


GOLDEN_SCENES = os.path.join(ROOT, "tests", "golden", "scenes")


def _load_fixture(name, tmp_path):
    _built()
    dst = str(tmp_path / f"{name}.bin")
    r = subprocess.run([DRIVER, TRACER_LIB, os.path.join(GOLDEN_SCENES, f"{name}.json"),
                        "flatten", dst], capture_output=True, text=True, timeout=120)
    return r, dst


def test_reference_scene_fixtures(tmp_path):
    """The reference loader test's scenes (T_SceneLoaderMRay.cpp:58-88): the
    empty scene fails to load; MinimalValidScene and BasicScene load, with
    the surfaces they define; BasicScene's one triangle arrives unchanged."""
    r, _ = _load_fixture("empty", tmp_path)
    assert r.returncode != 0 and "MRayError" in (r.stdout + r.stderr)
    r, _ = _load_fixture("minimal_valid", tmp_path)
    assert "loaded: 0 surfaces, 0 light surfaces, 1 camera surfaces" in r.stdout, r.stdout + r.stderr
    r, dst = _load_fixture("basic", tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "loaded: 1 surfaces, 0 light surfaces, 1 camera surfaces" in r.stdout
    f = read_flat_file(dst)
    tri = f["pos"][f["idx"]].reshape(-1, 3, 3)
    assert tri.shape == (1, 3, 3)
    assert np.array_equal(tri[0], np.array([[0, 0.5, 0], [-0.5, -0.5, 0], [0.5, -0.5, 0]],
                                           np.float32))
    mats = f["mats"].view(np.uint32).reshape(-1, 13)
    assert mats[0, 0] == abi.MAT_LAMBERT
    assert np.allclose(mats[0, 1:4].view(np.float32), [0, 1, 0])
    # normal +Z at every vertex: the TBN quaternion takes +Z to the tangent-space z
    from mray_amd import scenes as sc
    want = sc.quat_world_to_tangent(np.array([[0.0, 0.0, 1.0]]))[0]
    for q in f["tbn"][f["idx"]].reshape(-1, 4):
        assert np.allclose(np.abs(q), np.abs(want), atol=1e-6)
