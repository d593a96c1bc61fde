"""Oracle LBVH: Karras invariants, brute-force splits and traversal parity.

No reference test pins LBVH topology (there are no LBVH tests upstream), so
the construction is pinned by what Karras 2012 fixes: each internal node
covers a contiguous range of sorted leaves split at the first position where
the highest differing bit of (code, index) changes, every node/leaf has one
parent, and every box is the exact union of its children.  The node
numbering, which the reference's traversal order and node arrays follow
(internal node i has i at one end of its range), is pinned node for node
against a top-down construction written independently here.
"""
import numpy as np
import pytest

from mray_amd import scenes
from oracle import refcpu

LEAF = 0x80000000


def _check_tree(nodes, boxes, codes, order, aabbs):
    nodes = nodes.astype(np.int64)
    n = len(codes)
    m = n - 1
    parent = {}
    leaf_seen = np.zeros(n, bool)
    inv = np.empty(n, np.int64)
    inv[order] = np.arange(n)

    def rng_of(v):
        # contiguous leaf range of a subtree, by DFS
        lo, hi = n, -1
        st = [v]
        while st:
            x = st.pop()
            for c in nodes[x][:2]:
                if c & LEAF:
                    pos = int(inv[c & ~LEAF])
                    lo, hi = min(lo, pos), max(hi, pos)
                else:
                    st.append(int(c))
        return lo, hi

    for v in range(m):
        for side in range(2):
            c = int(nodes[v][side])
            if c & LEAF:
                t = c & ~LEAF
                assert not leaf_seen[t]
                leaf_seen[t] = True
                cb = aabbs[t]
            else:
                assert c not in parent
                parent[c] = v
                assert int(nodes[c][2]) == v, "parent pointer mismatch"
                cb = boxes[c]
            assert np.all(boxes[v][:3] <= cb[:3]) and np.all(boxes[v][3:] >= cb[3:])
    assert leaf_seen.all()
    assert set(parent) == set(range(1, m))
    assert nodes[0][2] == 0xFFFFFFFF
    # boxes are exact unions
    for v in range(m):
        cs = [aabbs[c & ~LEAF] if c & LEAF else boxes[c] for c in nodes[v][:2]]
        assert np.array_equal(boxes[v][:3], np.minimum(cs[0][:3], cs[1][:3]))
        assert np.array_equal(boxes[v][3:], np.maximum(cs[0][3:], cs[1][3:]))
    # Karras split: the children split the node's range where the common
    # prefix of (code, index) is shortest (highest differing bit)
    key = [(int(c) << 32) | i for i, c in enumerate(codes)]
    for v in range(m):
        lo, hi = rng_of(v)
        left = nodes[v][0]
        llo, lhi = (rng_of(int(left)) if not (left & LEAF)
                    else (int(inv[left & ~LEAF]),) * 2)
        assert llo == lo
        split = lhi
        hb = (key[lo] ^ key[hi]).bit_length()
        assert ((key[split] ^ key[split + 1]).bit_length()) == hb


def _karras_top_down(codes, order):
    """Karras 2012's tree built top down, independently of the oracle's
    bottom-up per-node search: a range [lo, hi] of sorted leaves splits after
    the last leaf whose (code, index) key shares the range's highest
    differing bit with key[lo]; the node of a range is numbered by the end
    that is not its parent's split (root 0), so the left child of a split at
    g is internal node g and the right child internal node g + 1, each a leaf
    when its range is one leaf (AcceleratorLBVH.cu:268-291)."""
    n = len(codes)
    key = [(int(c) << 32) | i for i, c in enumerate(codes)]
    nodes = np.zeros((max(n - 1, 1), 3), np.uint32)
    nodes[0][2] = 0xFFFFFFFF
    todo = [(0, 0, n - 1)]
    while todo:
        v, lo, hi = todo.pop()
        hb = (key[lo] ^ key[hi]).bit_length() - 1
        g = lo
        while g + 1 < hi and not ((key[g + 1] ^ key[lo]) >> hb):
            g += 1
        for side, (a, b, idx) in enumerate(((lo, g, g), (g + 1, hi, g + 1))):
            if a == b:
                nodes[v][side] = LEAF | int(order[a])
            else:
                nodes[v][side] = idx
                nodes[idx][2] = v
                todo.append((idx, a, b))
    return nodes


@pytest.mark.parametrize("n,seed,dups", [(2, 0, False), (3, 1, False), (8, 2, False),
                                         (37, 3, False), (200, 4, True), (1000, 5, True)])
def test_karras_invariants(n, seed, dups):
    rng = np.random.default_rng(seed)
    codes = rng.integers(0, 1 << 63, size=n, dtype=np.uint64)
    if dups:
        codes[n // 4: n // 2] = codes[n // 4]
    codes = np.sort(codes)
    order = rng.permutation(n).astype(np.uint32)
    lo = rng.uniform(-1, 1, size=(n, 3)).astype(np.float32)
    aabbs = np.concatenate([lo, lo + rng.uniform(0, 0.2, size=(n, 3)).astype(np.float32)], 1)
    assert refcpu.lbvh_tree_valid(codes, order)
    nodes, boxes = refcpu.lbvh_from_sorted(codes, order, aabbs)
    _check_tree(nodes, boxes, codes, order, aabbs)
    # the node numbering too: node for node, the top-down construction
    assert np.array_equal(nodes, _karras_top_down(codes, order))


def test_karras_numbering_fails_on_planted_mistakes():
    """The top-down comparison catches a swapped child pair and a split moved
    by one leaf (both keep the tree's shape invariants plausible)."""
    rng = np.random.default_rng(9)
    n = 64
    codes = np.sort(rng.integers(0, 1 << 63, size=n, dtype=np.uint64))
    order = rng.permutation(n).astype(np.uint32)
    good = _karras_top_down(codes, order)
    nodes, _ = refcpu.lbvh_from_sorted(codes, order, np.zeros((n, 6), np.float32))
    assert np.array_equal(nodes, good)
    swapped = good.copy()
    swapped[0][[0, 1]] = swapped[0][[1, 0]]
    assert not np.array_equal(nodes, swapped)
    # moving the root's split one leaf left changes the numbering below it
    shifted = codes.copy()
    hb = int(codes[0] ^ codes[-1]).bit_length() - 1
    g = max(i for i in range(n - 1) if not (int(codes[i] ^ codes[0]) >> hb))
    shifted[g] = codes[g + 1]  # leaf g now sorts with the right half
    assert not np.array_equal(_karras_top_down(shifted, order), good)


def test_reference_delta_malformed_on_degenerate_codes():
    """Codes differing only in low bits beside equal-code runs break the
    reference's Delta (clz of the index XOR); the strict (code, index) tie-break
    is always well-formed.  Scene builds fall back to it."""
    rng = np.random.default_rng(2)
    codes = np.sort(rng.integers(0, 64, size=3000).astype(np.uint64))
    order = rng.permutation(3000).astype(np.uint32)
    assert not refcpu.lbvh_tree_valid(codes, order)
    assert refcpu.lbvh_tree_valid(codes, order, strict=True)


def test_scene_lbvh_consistency():
    s = scenes.random_triangles(500, seed=1)
    o = refcpu.OracleScene(s)
    codes, order, nodes, boxes, aabb = o.lbvh()
    _, aabbs = refcpu.prim_centers_aabbs(s.positions, s.indices)
    assert np.array_equal(np.sort(codes), codes[order])
    assert np.array_equal(aabb, np.concatenate([aabbs[:, :3].min(0), aabbs[:, 3:].max(0)]))
    _check_tree(nodes, boxes, codes[order], order, aabbs)
    assert np.array_equal(nodes, _karras_top_down(codes[order], order))


def _rays(n, seed):
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 8), np.float32)
    r[:, :3] = rng.uniform(-2, 2, size=(n, 3))
    d = rng.normal(size=(n, 3))
    r[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    r[:, 3] = 1e-4
    r[:, 7] = np.where(rng.uniform(size=n) < 0.2, rng.uniform(0.1, 2, size=n), 3.4e38)
    return r


def test_traversal_equals_brute_force():
    s = scenes.random_triangles(4000, seed=6)
    o = refcpu.OracleScene(s)
    rays = _rays(20000, 7)
    p0, t0, b0 = o.cast_closest(rays, brute=False)
    p1, t1, b1 = o.cast_closest(rays, brute=True)
    assert np.array_equal(p0, p1)
    assert np.array_equal(t0, t1)
    assert np.array_equal(b0, b1)
    hit = p0 != 0xFFFFFFFF
    assert 0.05 < hit.mean() < 0.95
    vis = o.cast_visibility(rays)
    assert np.array_equal(vis == 1, ~hit)


def test_moller_trumbore_closed_form():
    """Axis-aligned rays onto the unit right triangle: t, barycentrics and the
    reordering (1-u-v, u) are known exactly; culling rejects the back side."""
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    idx = np.array([[0, 1, 2]], np.uint32)
    for cull in (0, 1):
        s = scenes.Scene(positions=pos, tbn=np.tile(np.array([[1, 0, 0, 0]], np.float32), (3, 1)),
                         indices=idx, surface_keys=np.zeros(1, np.uint32),
                         cull=np.array([cull], np.uint8),
                         materials=[{"type": 0, "albedo": (1, 1, 1)}],
                         light_tris=np.zeros(0, np.uint32),
                         light_radiance=np.zeros((0, 3), np.float32),
                         light_two_sided=np.zeros(0, np.uint8),
                         camera={"position": (0, 0, 1), "gaze": (0, 0, 0), "up": (0, 1, 0),
                                 "fov": (1.0, 1.0), "nearFar": (0.1, 10.0)})
        o = refcpu.OracleScene(s)
        rays = np.array([[0.25, 0.5, 2.0, 0.0, 0, 0, -1, 10.0],     # front, hits at t=2
                         [0.25, 0.5, -2.0, 0.0, 0, 0, 1, 10.0],     # from behind
                         [0.75, 0.75, 2.0, 0.0, 0, 0, -1, 10.0],    # outside (u+v>1)
                         [0.25, 0.5, 2.0, 2.5, 0, 0, -1, 10.0]],    # t < tMin
                        np.float32)
        prim, t, bary = o.cast_closest(rays, brute=True)
        assert prim[0] == 0 and t[0] == 2.0
        assert bary[0][0] == np.float32(0.25) and bary[0][1] == np.float32(0.25)
        assert prim[1] == (0xFFFFFFFF if cull else 0)
        assert prim[2] == 0xFFFFFFFF and prim[3] == 0xFFFFFFFF


def test_traversal_rays_in_box_planes():
    """Axis-parallel rays whose origins lie exactly on box face planes (0 * inf
    in the slab test) still reach their triangles: traversal == brute force."""
    # axis-aligned unit quads on a 4 x 4 grid at z = 0, every vertex on x/y = k/4
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
                     materials=[{"type": 0, "albedo": (1, 1, 1)}],
                     light_tris=np.zeros(0, np.uint32), light_radiance=np.zeros((0, 3), np.float32),
                     light_two_sided=np.zeros(0, np.uint8),
                     camera={"position": (0, 0, 1), "gaze": (0, 0, 0), "up": (0, 1, 0),
                             "fov": (1.0, 1.0), "nearFar": (0.1, 10.0)})
    o = refcpu.OracleScene(s)
    g = np.arange(0, 17, dtype=np.float32) / 16          # includes every box plane
    xy = np.array([[x, y] for x in g for y in g], np.float32)
    rays = np.zeros((len(xy), 8), np.float32)
    rays[:, 0:2] = xy
    rays[:, 2] = 1.0
    rays[:, 6] = -1.0          # direction (+0, +0, -1)
    rays[:, 7] = 10.0
    neg = rays.copy()
    neg[:, 4] = -0.0           # direction (-0, -0, -1): the other slab sign
    neg[:, 5] = -0.0
    for r in (rays, neg):
        p0, t0, _ = o.cast_closest(r, brute=False)
        p1, t1, _ = o.cast_closest(r, brute=True)
        assert (p1 != 0xFFFFFFFF).all()
        assert np.array_equal(p0, p1) and np.array_equal(t0, t1)
        assert (o.cast_visibility(r) == 0).all()
