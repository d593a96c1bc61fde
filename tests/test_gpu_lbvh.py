"""LBVH build on the GPU vs the CPU oracle: bit-exact integer and box outputs.

Reference: KCGenMortonCode (AcceleratorLBVH.cu:96-168), the CUB pair sort
(AcceleratorLBVH.hpp:828-838), KCConstructLBVHInternalNodes
(AcceleratorLBVH.cu:170-299) and KCUnionLBVHBoundingBoxes (:301-435).
"""
import numpy as np
import pytest

from mray_amd import scenes
from oracle import refcpu

pytestmark = pytest.mark.gpu


def _scenes():
    return [scenes.cornell(), scenes.random_triangles(20000, seed=3),
            scenes.terrain_spheres(grid=96, spheres=40, width=64, height=36)]


@pytest.mark.parametrize("idx", [0, 1, 2])
def test_accel_lbvh_matches_oracle(torch_gpu, idx):
    from mray_amd import native
    scene = _scenes()[idx]
    print(f"scene {scene.name} tris={scene.triangle_count}", flush=True)
    acc = native.Accel(scene)
    codes, order, nodes, boxes = acc.read_lbvh()
    o = refcpu.OracleScene(scene)
    rcodes, rorder, rnodes, rboxes, _ = o.lbvh()
    assert np.array_equal(codes, rcodes), "Morton codes differ"
    assert np.array_equal(order, rorder), "sorted leaf order differs"
    assert np.array_equal(nodes, rnodes), "Karras topology differs"
    assert np.array_equal(boxes.view(np.uint32), rboxes.view(np.uint32)), "node boxes differ"
    info = acc.info()
    assert info["triangles"] == scene.triangle_count
    acc.close()


def test_stage_morton_codes(torch_gpu):
    from mray_amd import native
    torch = torch_gpu
    rng = np.random.default_rng(11)
    for n in (1, 7, 1000, 300001):
        centers = rng.uniform(-3.0, 5.0, size=(n, 3)).astype(np.float32)
        # include points on the box faces and corners
        centers[: min(n, 4)] = np.array([[-3, -3, -3], [5, 5, 5], [-3, 5, 1], [5, -3, 0]],
                                        np.float32)[: min(n, 4)]
        box = np.array([-3, -3, -3, 5, 5, 5.5], np.float32)
        d_c = torch.from_numpy(centers).cuda()
        d_b = torch.from_numpy(box).cuda()
        g63 = native.morton63(d_c, d_b).cpu().numpy().view(np.uint64)
        g30 = native.morton30(d_c, d_b).cpu().numpy().view(np.uint32)
        assert np.array_equal(g63, refcpu.morton63_codes(centers, box))
        assert np.array_equal(g30, refcpu.morton30_codes(centers, box))


def test_stage_centers_aabbs_reduce(torch_gpu):
    from mray_amd import native
    torch = torch_gpu
    s = scenes.random_triangles(5000, seed=5)
    d_pos = torch.from_numpy(s.positions).cuda()
    d_idx = torch.from_numpy(s.indices.astype(np.int32)).cuda()
    c, a = native.prim_centers_aabbs(d_pos, d_idx)
    rc, ra = refcpu.prim_centers_aabbs(s.positions, s.indices)
    assert np.array_equal(c.cpu().numpy().view(np.uint32), rc.view(np.uint32))
    assert np.array_equal(a.cpu().numpy().view(np.uint32), ra.view(np.uint32))
    box = native.reduce_aabb(a).cpu().numpy()
    assert np.array_equal(box, np.concatenate([ra[:, :3].min(0), ra[:, 3:].max(0)]))


@pytest.mark.parametrize("n", [1, 2, 4095, 4096, 4097, 250000])
def test_stage_radix_sort_stable(torch_gpu, n):
    from mray_amd import native
    torch = torch_gpu
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 1 << 63, size=n, dtype=np.uint64)
    keys[::3] = keys[0]                     # many duplicates: stability matters
    keys[1::7] = np.uint64((1 << 63) - 1)
    vals = np.arange(n, dtype=np.uint32)
    dk = torch.from_numpy(keys.view(np.int64).copy()).cuda()
    dv = torch.from_numpy(vals.view(np.int32).copy()).cuda()
    native.radix_sort_u64_u32(dk, dv)
    gk = dk.cpu().numpy().view(np.uint64)
    gv = dv.cpu().numpy().view(np.uint32)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(gk, keys[order])
    assert np.array_equal(gv, vals[order])


def _lbvh_inputs(codes, seed):
    rng = np.random.default_rng(seed)
    n = len(codes)
    order = rng.permutation(n).astype(np.uint32)
    lo = rng.uniform(-1, 0, size=(n, 3)).astype(np.float32)
    aabbs = np.concatenate([lo, lo + 0.1], axis=1).astype(np.float32)
    return order, aabbs


def _gpu_lbvh(codes, order, aabbs):
    import torch
    from mray_amd import native
    nodes, boxes = native.lbvh_build(torch.from_numpy(codes.view(np.int64).copy()).cuda(),
                                     torch.from_numpy(order.view(np.int32).copy()).cuda(),
                                     torch.from_numpy(aabbs).cuda())
    return nodes.cpu().numpy().view(np.uint32), boxes.cpu().numpy()


def test_stage_lbvh_build_duplicate_runs(torch_gpu):
    """63-bit codes with long runs of equal codes (index-XOR fallback of Delta)."""
    rng = np.random.default_rng(2)
    n = 5000
    codes = rng.integers(0, 1 << 63, size=n, dtype=np.uint64)
    codes[100:700] = codes[100]
    codes[2000:2003] = codes[2000]
    codes = np.sort(codes)
    order, aabbs = _lbvh_inputs(codes, 3)
    assert refcpu.lbvh_tree_valid(codes, order)
    nodes, boxes = _gpu_lbvh(codes, order, aabbs)
    rnodes, rboxes = refcpu.lbvh_from_sorted(codes, order, aabbs)
    assert np.array_equal(nodes, rnodes)
    assert np.array_equal(boxes, rboxes)


def test_stage_lbvh_build_degenerate_codes(torch_gpu):
    """Codes that differ only in low bits next to equal-code runs: the
    reference's Delta yields a malformed tree.  The kernels stay in bounds and
    produce the same (malformed) child links as the oracle; parent links and
    box values of a malformed tree are order-dependent and not compared."""
    rng = np.random.default_rng(2)
    n = 3000
    codes = np.sort(rng.integers(0, 64, size=n).astype(np.uint64))
    order, aabbs = _lbvh_inputs(codes, 4)
    assert not refcpu.lbvh_tree_valid(codes, order)
    assert refcpu.lbvh_tree_valid(codes, order, strict=True)
    nodes, _ = _gpu_lbvh(codes, order, aabbs)
    rnodes, _ = refcpu.lbvh_from_sorted(codes, order, aabbs)
    # children are written by their own node's thread (deterministic); in a
    # malformed tree several nodes claim one child, so parent words race
    assert np.array_equal(nodes[:, :2], rnodes[:, :2])


def _canonical(nodes, prims):
    """The 4-wide tree as nested tuples from the root: each slot's box bits
    and either its leaf word or its subtree (node numbering removed)."""
    out = []
    stack = [0]
    seen = 0
    while stack:
        v = stack.pop()
        seen += 1
        w = nodes[v]
        row = [tuple(w[:24].tolist())]
        for k in range(4):
            c = int(w[24 + k])
            if c == 0xFFFFFFFF or c & 0x80000000:
                row.append(c)
            else:
                row.append("node")
                stack.append(c)
        out.append(tuple(row))
    return out, seen


@pytest.mark.parametrize("idx", [0, 1, 2, 3])
def test_device_collapse_is_host_collapse(torch_gpu, idx):
    """The LBVH builder's 4-wide tree and triangle soup built on the GPU
    (collapse.hip) against the same collapse on the host: the same nodes in
    the same depth-first order of their children, the same leaf words and
    the same soup."""
    from mray_amd import native
    scene = (_scenes() + [scenes.random_triangles(300001, seed=9)])[idx]
    dev = native.Accel(scene, builder="lbvh")
    host = native.Accel(scene, builder="lbvh", host_collapse=True)
    dn, dp = dev.read_wide()
    hn, hp = host.read_wide()
    di, hi = dev.info(), host.info()
    dev.close()
    host.close()
    print(f"scene {scene.name}: {scene.triangle_count} tris, wide nodes {di['wide_nodes']}", flush=True)
    assert di["wide_nodes"] == hi["wide_nodes"] == len(dn) == len(hn)
    assert np.array_equal(dp, hp)
    assert np.array_equal(np.sort(dp), np.arange(scene.triangle_count, dtype=np.uint32))
    cd, nd = _canonical(dn, dp)
    ch, nh = _canonical(hn, hp)
    assert nd == nh == len(dn)
    assert cd == ch
