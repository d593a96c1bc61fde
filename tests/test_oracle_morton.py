"""Oracle Morton codes against the reference's golden words and brute force.

Golden words: Tests/Core/T_GraphicsFunctions.cpp:182-305 (Compose3D /
Decompose3D, 32- and 64-bit).  Quantisation (AcceleratorLBVH.cu:135-165) is
checked against an independent float64/decimal restatement on random and
boundary points.
"""
import numpy as np

from oracle import refcpu

ONES10 = 0b1111111111
ONES20 = 0b11111111111111111111


def test_compose3d_64_golden():
    assert refcpu.morton63_compose(ONES20, 0, 0) == 0x249249249249249
    assert refcpu.morton63_compose(0, ONES20, 0) == 0x492492492492492
    assert refcpu.morton63_compose(0, 0, ONES20) == 0x924924924924924


def test_decompose3d_64_golden():
    assert refcpu.morton63_decompose(0x249249249249249) == (ONES20, 0, 0)
    assert refcpu.morton63_decompose(0x492492492492492) == (0, ONES20, 0)
    assert refcpu.morton63_decompose(0x924924924924924) == (0, 0, ONES20)


def test_compose3d_32_golden():
    assert refcpu.morton30_compose(ONES10, 0, 0) == 0b001001001001001001001001001001
    assert refcpu.morton30_compose(0, ONES10, 0) == 0b010010010010010010010010010010
    assert refcpu.morton30_compose(0, 0, ONES10) == 0b100100100100100100100100100100


def test_decompose3d_32_golden():
    assert refcpu.morton30_decompose(0b001001001001001001001001001001) == (ONES10, 0, 0)
    assert refcpu.morton30_decompose(0b010010010010010010010010010010) == (0, ONES10, 0)
    assert refcpu.morton30_decompose(0b100100100100100100100100100100) == (0, 0, ONES10)


def _interleave(x, y, z, bits):
    c = 0
    for b in range(bits):
        c |= ((x >> b) & 1) << (3 * b)
        c |= ((y >> b) & 1) << (3 * b + 1)
        c |= ((z >> b) & 1) << (3 * b + 2)
    return c


def test_compose_brute_force_random():
    rng = np.random.default_rng(0)
    for _ in range(300):
        x, y, z = (int(v) for v in rng.integers(0, 1 << 21, size=3))
        c = refcpu.morton63_compose(x, y, z)
        assert c == _interleave(x, y, z, 21)
        assert refcpu.morton63_decompose(c) == (x, y, z)
        x, y, z = (int(v) for v in rng.integers(0, 1 << 10, size=3))
        c = refcpu.morton30_compose(x, y, z)
        assert c == _interleave(x, y, z, 10)
        assert refcpu.morton30_decompose(c) == (x, y, z)
    # bits above the field width are dropped by the masks
    assert refcpu.morton63_compose((1 << 21) | 5, 0, 0) == _interleave(5, 0, 0, 21)
    assert refcpu.morton30_compose((1 << 10) | 5, 0, 0) == _interleave(5, 0, 0, 10)


def _quantize_ref(c, box, bits):
    """float32 offset, float64 product, float32 round, lround (half away from
    zero), clamp -- restated with numpy/decimal rounding, not the oracle's."""
    c = np.float32(c)
    lo = np.float32(box[0])
    span = np.float32(box[1])
    d = np.float32(c - lo)
    d = max(d, np.float32(0))
    q = np.float32(np.float64(d) * (np.float64(1 << bits) / np.float64(span)))
    r = int(np.floor(np.float64(q) + 0.5)) if q >= 0 else -int(np.floor(-np.float64(q) + 0.5))
    return min(max(r, 0), (1 << bits) - 1)


def test_quantisation_matches_independent_restatement():
    rng = np.random.default_rng(1)
    box = np.array([-2.0, -1.0, 0.5, 3.0, 1.0, 4.5], np.float32)
    span = box[3:] - box[:3]
    max_side = span.max()
    pts = rng.uniform(box[:3], box[3:], size=(4000, 3)).astype(np.float32)
    # boundary points, outside points (clamped) and exact half-slice points
    extra = [box[:3], box[3:], box[:3] - 1, box[3:] + 1]
    half = box[:3] + np.float32(max_side / (1 << 21) * 2.5)
    extra.append(half.astype(np.float32))
    pts = np.concatenate([pts, np.array(extra, np.float32)])
    codes = refcpu.morton63_codes(pts, box)
    codes30 = refcpu.morton30_codes(pts, box)
    for p, c, c30 in zip(pts, codes, codes30):
        q = [_quantize_ref(p[k], (box[k], max_side), 21) for k in range(3)]
        assert int(c) == _interleave(*q, 21)
        q30 = [_quantize_ref(p[k], (box[k], max_side), 10) for k in range(3)]
        assert int(c30) == _interleave(*q30, 10)


def test_centroid_is_uncontracted_left_to_right():
    """Triangle centre = (p0*k + p1*k) + p2*k, k = 0.333333333, each op one
    float32 rounding (PrimitiveDefaultTriangle.hpp:107-115)."""
    rng = np.random.default_rng(3)
    pos = rng.normal(size=(300, 3)).astype(np.float32) * np.float32(7.3)
    idx = np.arange(300, dtype=np.uint32).reshape(-1, 3)
    centers, aabbs = refcpu.prim_centers_aabbs(pos, idx)
    k = np.float32(0.333333333)
    p = pos[idx]
    ref = (p[:, 0] * k + p[:, 1] * k) + p[:, 2] * k
    assert np.array_equal(centers, ref.astype(np.float32))
    assert np.array_equal(aabbs[:, :3], p.min(axis=1))
    assert np.array_equal(aabbs[:, 3:], p.max(axis=1))
