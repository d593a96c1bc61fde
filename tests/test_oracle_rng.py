"""Oracle RNG: std::mt19937 seeding and PCG32 (Tracer/Random.{h,cu}).

Pins: the C++ standard's mt19937 10000th-output value, the reference's own
static_asserts on ToFloat01 (Random.h:101-120), the LCG jump-ahead identity
(Brown 1994, the algorithm of PermutedCG32::Advance) and bijectivity of the
RXS-M-XS output permutation.
"""
import numpy as np

from mray_amd import scenes
from oracle import refcpu

M = 747796405
INC = 2891336453
MASK = 0xFFFFFFFF


def test_mt19937_standard_value():
    r = scenes.MT19937(5489)
    for _ in range(9999):
        r()
    assert r() == 4123659995


def _advance(state, delta):
    """Brown's arbitrary-stride LCG jump (independent of single stepping)."""
    acc_mult, acc_plus, cur_mult, cur_plus = 1, 0, M, INC
    while delta > 0:
        if delta & 1:
            acc_mult = (acc_mult * cur_mult) & MASK
            acc_plus = (acc_plus * cur_mult + cur_plus) & MASK
        cur_plus = ((cur_mult + 1) * cur_plus) & MASK
        cur_mult = (cur_mult * cur_mult) & MASK
        delta >>= 1
    return (acc_mult * state + acc_plus) & MASK


def test_pcg_state_matches_jump_ahead():
    import ctypes
    for s0 in (0, 1, 0xDEADBEEF, 123456789):
        st = np.array([s0], np.uint32)
        L = refcpu.lib()
        for k in range(1, 200):
            L.rc_pcg_next(st.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
            assert int(st[0]) == _advance(s0, k)


def test_seed_states_follow_mt19937():
    """seed64 -> seed32 = hi ^ lo; state_i = GenerateState(mt()_i)
    (Random.cu:680-703); GenerateState(x) = Step(Step(0) + x)."""
    seed = (0x1234 << 32) | (0x1234 ^ 5489)       # hi ^ lo == 5489
    states = refcpu.seed_states(seed, 10000)
    r = scenes.MT19937(5489)
    xs = [r() for _ in range(10000)]
    assert xs[9999] == 4123659995
    step0 = _advance(0, 1)
    for i in (0, 1, 17, 9999):
        assert int(states[i]) == _advance((step0 + xs[i]) & MASK, 1)


def test_output_permutation_is_bijective():
    """2^20 distinct states must give 2^20 distinct outputs; a random map
    would collide ~128 times."""
    import ctypes
    L = refcpu.lib()
    rng = np.random.default_rng(0)
    states = np.unique(rng.integers(0, 1 << 32, size=1 << 20, dtype=np.uint64)).astype(np.uint32)
    outs = np.empty_like(states)
    buf = np.zeros(1, np.uint32)
    ptr = buf.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
    for i in range(0, len(states), 4096):
        for j, s in enumerate(states[i:i + 4096]):
            buf[0] = s
            outs[i + j] = L.rc_pcg_next(ptr)
    assert len(np.unique(outs)) == len(states)


def test_to_float01_reference_asserts():
    # static_asserts of ToFloat01 (Random.h:107-109) and the clamp to the
    # largest float below one
    assert refcpu.to_float01(0) == 0.0
    assert np.float32(0xFFFFFF7F) * np.float32(2.0 ** -32) != np.float32(1.0)
    assert np.float32(0xFFFFFF80) * np.float32(2.0 ** -32) == np.float32(1.0)
    below_one = float(np.nextafter(np.float32(1), np.float32(0)))
    assert refcpu.to_float01(0xFFFFFF80) == below_one
    assert refcpu.to_float01(0xFFFFFFFF) == below_one
    assert refcpu.to_float01(0x80000000) == 0.5


def test_pcg_uniformity():
    seq = refcpu.pcg_sequence(42, 200000)
    u = np.array([refcpu.to_float01(v) for v in seq[:50000]])
    hist, _ = np.histogram(u, bins=50, range=(0, 1))
    exp = len(u) / 50
    chi2 = float(((hist - exp) ** 2 / exp).sum())
    assert chi2 < 100.0          # 49 dof, p ~ 1e-4 bound
    assert 0.49 < u.mean() < 0.51
