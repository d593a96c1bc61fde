"""Closest-hit and visibility casting on the GPU vs the CPU oracle.

Hit triangle ids, t and barycentrics must be bit-identical for identical rays
(Moller-Trumbore as Ray.hpp:121-167; closest-t, then smaller triangle index).
"""
import math

import numpy as np
import pytest

from mray_amd import abi, scenes
from oracle import refcpu

pytestmark = pytest.mark.gpu


def _random_rays(n, seed, origin_scale=2.0):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-origin_scale, origin_scale, size=(n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3] = 1e-4
    rays[:, 4:7] = d
    rays[:, 7] = 3.4e38
    # a few short rays and axis-aligned directions
    rays[: n // 10, 7] = rng.uniform(0.1, 2.0, size=n // 10)
    rays[n // 10: n // 10 + 6, 4:7] = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0],
                                                [0, 0, 1], [0, 0, -1]], np.float32)
    return rays


def _compare(acc, oscene, rays, brute):
    import torch
    d = torch.from_numpy(rays).cuda()
    prim, t, bary = acc.cast_closest(d)
    prim = prim.cpu().numpy().view(np.uint32)
    t = t.cpu().numpy()
    bary = bary.cpu().numpy()
    rp, rt, rb = oscene.cast_closest(rays, brute=brute)
    hit = rp != abi.INVALID_PRIM
    print(f"rays={len(rays)} hits={int(hit.sum())}", flush=True)
    assert np.array_equal(prim, rp)
    assert np.array_equal(t.view(np.uint32), rt.view(np.uint32))
    assert np.array_equal(bary[hit].view(np.uint32), rb[hit].view(np.uint32))
    vis = acc.cast_visibility(d).cpu().numpy()
    rvis = oscene.cast_visibility(rays)
    assert np.array_equal(vis.astype(np.uint32), rvis)
    # any-hit agrees with closest hit
    assert np.array_equal(vis == 1, ~hit)


@pytest.mark.parametrize("builder,width", [("sah", 4), ("lbvh", 4), ("sah", 8), ("lbvh", 8)])
def test_closest_soup_vs_bruteforce(torch_gpu, builder, width):
    from mray_amd import native
    s = scenes.random_triangles(3000, seed=9)
    acc = native.Accel(s, builder=builder, width=width)
    _compare(acc, refcpu.OracleScene(s), _random_rays(20000, 1), brute=True)


@pytest.mark.parametrize("builder,width", [("sah", 4), ("lbvh", 4), ("sah", 8)])
def test_closest_soup_large(torch_gpu, builder, width):
    from mray_amd import native
    s = scenes.random_triangles(200000, seed=4)
    acc = native.Accel(s, builder=builder, width=width)
    _compare(acc, refcpu.OracleScene(s), _random_rays(100000, 2), brute=False)


def test_primary_rays_cornell(torch_gpu):
    """Primary camera rays of the Cornell workload (identical rays, from the oracle)."""
    from mray_amd import native
    s = scenes.cornell()
    acc = native.Accel(s)
    o = refcpu.OracleScene(s)
    params = abi.render_params(256, 256, 1)
    rays, _ = o.camera_rays(params, 256 * 256)
    _compare(acc, o, rays, brute=True)


@pytest.mark.parametrize("builder,width", [("sah", 4), ("lbvh", 4), ("sah", 8)])
def test_primary_rays_terrain(torch_gpu, builder, width):
    from mray_amd import native
    s = scenes.terrain_spheres(grid=256, spheres=300, width=320, height=180)
    acc = native.Accel(s, builder=builder, width=width)
    o = refcpu.OracleScene(s)
    params = abi.render_params(320, 180, 1)
    rays, _ = o.camera_rays(params, 320 * 180)
    _compare(acc, o, rays, brute=False)


def test_single_triangle_and_edges(torch_gpu):
    """One-triangle scene (single-leaf LBVH case) and rays through edges/vertices."""
    from mray_amd import native
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    s = scenes.Scene(positions=pos, tbn=scenes.quat_world_to_tangent(np.tile([[0, 0, 1.0]], (3, 1))),
                     indices=np.array([[0, 1, 2]], np.uint32),
                     surface_keys=np.zeros(1, np.uint32), cull=np.zeros(1, np.uint8),
                     materials=[{"type": abi.MAT_LAMBERT, "albedo": (0.5, 0.5, 0.5)}],
                     light_tris=np.zeros(0, np.uint32), light_radiance=np.zeros((0, 3), np.float32),
                     light_two_sided=np.zeros(0, np.uint8),
                     camera={"position": (0, 0, 2), "gaze": (0, 0, 0), "up": (0, 1, 0),
                             "fov": (math.radians(40),) * 2, "nearFar": (0.01, 100.0)})
    acc = native.Accel(s)
    xs = np.linspace(-0.25, 1.25, 61, dtype=np.float32)
    g = np.array([[x, y] for x in xs for y in xs], np.float32)
    rays = np.zeros((len(g), 8), np.float32)
    rays[:, 0:2] = g
    rays[:, 2] = 1.0
    rays[:, 3] = 0.0
    rays[:, 6] = -1.0
    rays[:, 7] = 10.0
    _compare(acc, refcpu.OracleScene(s), rays, brute=True)


@pytest.mark.parametrize("width", [4, 8])
def test_rays_in_box_planes(torch_gpu, width):
    """Axis-parallel rays (+0 and -0 direction components) whose origins lie
    on the BVH box planes of a grid of axis-aligned quads."""
    from mray_amd import native
    pos, idx = [], []
    for i in range(4):
        for j in range(4):
            b = len(pos)
            x0, y0 = i / 4, j / 4
            pos += [[x0, y0, 0], [x0 + 0.25, y0, 0], [x0 + 0.25, y0 + 0.25, 0], [x0, y0 + 0.25, 0]]
            idx += [[b, b + 1, b + 2], [b, b + 2, b + 3]]
    pos = np.array(pos, np.float32)
    idx = np.array(idx, np.uint32)
    n = len(idx)
    s = scenes.Scene(positions=pos, tbn=np.tile(np.array([[1, 0, 0, 0]], np.float32), (len(pos), 1)),
                     indices=idx, surface_keys=np.zeros(n, np.uint32), cull=np.zeros(n, np.uint8),
                     materials=[{"type": abi.MAT_LAMBERT, "albedo": (1, 1, 1)}],
                     light_tris=np.zeros(0, np.uint32), light_radiance=np.zeros((0, 3), np.float32),
                     light_two_sided=np.zeros(0, np.uint8),
                     camera={"position": (0, 0, 1), "gaze": (0, 0, 0), "up": (0, 1, 0),
                             "fov": (1.0, 1.0), "nearFar": (0.1, 10.0)})
    acc = native.Accel(s, width=width)
    o = refcpu.OracleScene(s)
    g = np.arange(0, 17, dtype=np.float32) / 16
    xy = np.array([[x, y] for x in g for y in g], np.float32)
    rays = np.zeros((len(xy), 8), np.float32)
    rays[:, 0:2] = xy
    rays[:, 2] = 1.0
    rays[:, 6] = -1.0
    rays[:, 7] = 10.0
    neg = rays.copy()
    neg[:, 4] = -0.0
    neg[:, 5] = -0.0
    for r in (rays, neg):
        _compare(acc, o, r, brute=True)


def _plane_stack(n_planes):
    """n_planes parallel squares [-1, 1]^2 at z = k / n_planes, two triangles
    each: every box of the tree spans the whole square, so a ray along +z
    enters every child of every node it visits."""
    pos, idx = [], []
    for k in range(n_planes):
        z = k / n_planes
        b = len(pos)
        pos += [[-1, -1, z], [1, -1, z], [1, 1, z], [-1, 1, z]]
        idx += [[b, b + 1, b + 2], [b, b + 2, b + 3]]
    pos = np.array(pos, np.float32)
    idx = np.array(idx, np.uint32)
    n = len(idx)
    return scenes.Scene(positions=pos, tbn=np.tile(np.array([[1, 0, 0, 0]], np.float32), (len(pos), 1)),
                        indices=idx, surface_keys=np.zeros(n, np.uint32), cull=np.zeros(n, np.uint8),
                        materials=[{"type": abi.MAT_LAMBERT, "albedo": (1, 1, 1)}],
                        light_tris=np.zeros(0, np.uint32), light_radiance=np.zeros((0, 3), np.float32),
                        light_two_sided=np.zeros(0, np.uint8),
                        camera={"position": (0, 0, -2), "gaze": (0, 0, 0), "up": (0, 1, 0),
                                "fov": (1.0, 1.0), "nearFar": (0.1, 10.0)})


def _first_descent_depth(nodes, x, y):
    """Stack depth of a +z ray at (x, y) when its near-first walk of the
    4-wide tree reaches its first leaf: every node pushes its hit inner
    children but the nearest (words 0-23: lo/hi x, y, z of the 4 children,
    24-27: child words, leaf flag 0x80000000)."""
    node, depth = 0, 0
    while True:
        w = nodes[node]
        inner = []
        for k in range(4):
            lx, hx = w[0 + k].view(np.float32), w[4 + k].view(np.float32)
            ly, hy = w[8 + k].view(np.float32), w[12 + k].view(np.float32)
            lz = w[16 + k].view(np.float32)
            c = int(w[24 + k])
            if not (lx <= x <= hx and ly <= y <= hy) or (c & 0x80000000):
                continue
            inner.append((float(lz), c))
        if not inner:
            return depth
        inner.sort()
        depth += len(inner) - 1
        node = inner[0][1]


@pytest.mark.parametrize("builder", ["sah", "lbvh"])
def test_deep_stack_spills_past_lds(torch_gpu, builder):
    """A walk whose stack outgrows the 12 LDS levels: the deeper entries go
    to the lane's global spill area, and waves mix lanes on either side of
    the boundary.  Hits, t, barycentrics and visibility equal brute force."""
    from mray_amd import native
    s = _plane_stack(4096)
    acc = native.Accel(s, builder=builder, width=4)
    nodes, _ = acc.read_wide()
    depth = _first_descent_depth(nodes, np.float32(0.1), np.float32(-0.2))
    print(f"{builder}: {len(nodes)} nodes, first-descent stack depth {depth}", flush=True)
    assert depth > 12
    rng = np.random.default_rng(3)
    n = 4096
    rays = np.zeros((n, 8), np.float32)
    # +z rays from below (deep stacks), oblique rays and rays from between
    # the planes (shallow stacks) interleaved within every wave
    rays[:, 0:2] = rng.uniform(-0.9, 0.9, size=(n, 2))
    rays[:, 2] = -1.0
    rays[:, 6] = 1.0
    odd = np.arange(n) % 3 == 1
    d = rng.normal(size=(int(odd.sum()), 3))
    d[:, 2] = np.abs(d[:, 2]) + 0.2
    rays[odd, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    mid = np.arange(n) % 3 == 2
    rays[mid, 2] = rng.uniform(0.0, 1.0, size=int(mid.sum()))
    rays[mid, 6] = -1.0
    rays[:, 3] = 1e-4
    rays[:, 7] = 3.4e38
    _compare(acc, refcpu.OracleScene(s), rays, brute=True)
