"""Pins of the neural-radiance-cache oracle (oracle/nrc_ref.py) by what the
mathematics fixes: finite differences for every gradient, exact trilinear
reproduction of linear fields on dense levels, the spatial hash's published
primes, bf16 round-to-nearest-even on known values, Adam's first step
(lr * sign(g)), and the one-blob kernel shape."""
import numpy as np
import pytest

from oracle import nrc_ref as N


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e38, 1e-40], np.float32)
    r = N.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # tie -> even
    assert r[2] == np.float32(1.0 + 2 ** -7)
    assert r[3] == -2.5
    assert np.isfinite(r[4]) or np.isinf(r[4])
    # every result is representable in bf16: low 16 bits zero
    assert np.all((r.view(np.uint32) & 0xFFFF) == 0)


def test_level_scales():
    s = N.level_scale()
    assert s[0] == pytest.approx(N.N_MIN - 1)
    assert s[-1] == pytest.approx(N.N_MAX - 1, rel=1e-5)
    assert np.all(np.diff(s) > 0)


def test_hash_index_matches_published_primes():
    s = N.level_scale()
    lvl = N.LEVELS - 1                         # finest level is hashed
    assert N.level_res(s)[lvl] ** 3 > (1 << N.LOG2_T)
    cell = np.array([[1, 2, 3], [1000, 7, 65535]], np.int64)
    got = N.grid_index(lvl, cell, s)
    want = [((x * 1) ^ (y * 2654435761) ^ (z * 805459861)) & ((1 << 19) - 1) for x, y, z in cell]
    assert list(got) == want


def test_trilinear_reproduces_linear_field_on_dense_level():
    s = N.level_scale()
    lvl = 0
    res = int(N.level_res(s)[lvl])
    assert res ** 3 <= (1 << N.LOG2_T)
    grid = np.zeros((N.LEVELS, 1 << N.LOG2_T, N.FEATS), np.float32)
    c = np.stack(np.meshgrid(np.arange(res), np.arange(res), np.arange(res), indexing="ij"), -1)
    c = c.reshape(-1, 3)
    idx = N.grid_index(lvl, c, s)
    a = np.array([1.0, 2.0, 3.0])
    grid[lvl, idx, 0] = c @ a
    rng = np.random.default_rng(0)
    pos = rng.uniform(0, 1, size=(200, 3)).astype(np.float32)
    f, _ = N.encode_grid(pos, grid, s)
    p = pos.astype(np.float32) * s[lvl] + np.float32(0.5)
    assert np.allclose(f[:, 0], p.astype(np.float64) @ a, rtol=1e-9, atol=1e-9)


def test_one_blob_peaks():
    k = N.BLOB
    centres = (np.arange(k) + 0.5) / k
    e = N.one_blob(centres)
    assert np.allclose(np.diag(e), 1.0)
    assert np.all(e[0, 1:] < 1.0)
    assert np.allclose(e[:, 0][1], np.exp(-0.5))   # one bin away = one sigma


def _small_net(rng):
    return [rng.normal(scale=np.sqrt(2.0 / i), size=(i, o)).astype(np.float32)
            for i, o in N.weight_shapes()]


def _inputs(rng, n):
    rec = np.zeros((n, 11), np.float32)
    rec[:, 0:8] = rng.uniform(0.05, 0.95, size=(n, 8))
    rec[:, 8:11] = rng.uniform(0, 1, size=(n, 3))
    return rec


def test_mlp_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    W = [w.astype(np.float64) for w in _small_net(rng)]
    n = 8
    x = rng.normal(size=(n, N.IN_PAD))
    t = rng.uniform(0, 1, size=(n, 3))
    y, acts = N.forward(x, W, mixed=False)
    den = y ** 2 + 0.01                      # stop-gradient denominator, kept fixed

    def loss(Ws, xs):
        yy, _ = N.forward(xs, Ws, mixed=False)
        return np.sum((yy - t) ** 2 / den) / (n * 3)
    _, dy = N.loss_grad(y, t, n)
    grads, dx = N.backward(dy, acts, W, mixed=False)
    h = 1e-6
    for li in (0, 2, 5):
        for _ in range(4):
            i, j = rng.integers(0, W[li].shape[0]), rng.integers(0, W[li].shape[1] if li < 5 else 3)
            Wp = [w.copy() for w in W]
            Wm = [w.copy() for w in W]
            Wp[li][i, j] += h
            Wm[li][i, j] -= h
            fd = (loss(Wp, x) - loss(Wm, x)) / (2 * h)
            assert grads[li][i, j] == pytest.approx(fd, rel=1e-5, abs=1e-9)
    for _ in range(6):
        s, k = rng.integers(0, n), rng.integers(0, N.IN_PAD)
        xp, xm = x.copy(), x.copy()
        xp[s, k] += h
        xm[s, k] -= h
        fd = (loss(W, xp) - loss(W, xm)) / (2 * h)
        assert dx[s, k] == pytest.approx(fd, rel=1e-5, abs=1e-9)


def test_grid_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    W = [w.astype(np.float64) for w in _small_net(rng)]
    s = N.level_scale()
    grid = rng.uniform(-1e-1, 1e-1, size=(N.LEVELS, 1 << N.LOG2_T, N.FEATS)).astype(np.float64)
    n = 6
    rec = _inputs(rng, n)
    t = rng.uniform(0, 1, size=(n, 3))
    x, aux = N.encode_inputs(rec, grid, s)
    y, acts = N.forward(x, W, mixed=False)
    den = y ** 2 + 0.01
    _, dy = N.loss_grad(y, t, n)
    _, dx = N.backward(dy, acts, W, mixed=False)
    gg = N.grid_grad(dx, aux, grid.shape)

    def loss(g):
        xx, _ = N.encode_inputs(rec, g, s)
        yy, _ = N.forward(xx, W, mixed=False)
        return np.sum((yy - t) ** 2 / den) / (n * 3)
    h = 1e-6
    # entries actually touched by sample 0 at a dense and at a hashed level
    for lvl in (0, N.LEVELS - 1):
        idx, _ = aux[lvl]
        e = idx[0, 3]
        for f in range(N.FEATS):
            gp, gm = grid.copy(), grid.copy()
            gp[lvl, e, f] += h
            gm[lvl, e, f] -= h
            fd = (loss(gp) - loss(gm)) / (2 * h)
            assert gg[lvl, e, f] == pytest.approx(fd, rel=1e-5, abs=1e-10)


def test_adam_first_step_is_lr_sign():
    g = np.array([0.5, -2.0, 1e-3])
    p, m, v = N.adam(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, lr=0.01)
    assert np.allclose(p, -0.01 * np.sign(g))
