"""CPU oracle for the neural radiance cache (test infrastructure only).

The reference has no neural cache (its nearest analogue is the irradiance
hash grid of Tracer/HashGrid.h:17-476 with an EMA update at
GuidedPTRendererShaders.h:343-359), so this restates the published methods
the GPU implementation follows:

* multiresolution hash encoding -- Mueller et al. 2022, "Instant Neural
  Graphics Primitives" sec. 3: L levels, T entries per level, F features per
  entry, resolution N_l = floor(N_min * b^l); a level whose (res+1)^3 cells fit
  in T is indexed densely, otherwise by the spatial hash
  (x*1 ^ y*2654435761 ^ z*805459861) mod T; trilinear interpolation of the 8
  corner features;
* one-blob encoding -- Mueller et al. 2019, "Neural Importance Sampling":
  k Gaussian kernels (sigma = 1/k) centred at (j + 0.5)/k;
* the network -- Mueller et al. 2021, "Real-time Neural Radiance Caching":
  64-wide fully fused MLP, ReLU hidden layers, no biases, linear output, relative
  L2 loss (y - t)^2 / (sg(y)^2 + 0.01), Adam;
* mixed precision as the GPU runs it: weights and activations rounded to
  bfloat16 (round to nearest even) before every product, products
  accumulated in fp32 (here: float64, the difference is the tolerance of the
  parity tests).

Everything is plain numpy; shapes follow the GPU buffers.
"""
import numpy as np

LEVELS = 16
LOG2_T = 19
FEATS = 2
N_MIN = 16
N_MAX = 2048
WIDTH = 64
HIDDEN = 5            # hidden layers of width 64 (ReLU)
OUT = 3
OUT_PAD = 16
IN_PAD = 64
BLOB = 4
PRIMES = (1, 2654435761, 805459861)


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    rounding = ((u >> 16) & 1) + 0x7FFF
    r = ((u + rounding) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(a), a, out)


def f16(x):
    return np.asarray(x, dtype=np.float32).astype(np.float16).astype(np.float32)


def level_scale(levels=LEVELS, n_min=N_MIN, n_max=N_MAX):
    """Per-level scale s_l (float32): positions map to x * s_l + 0.5."""
    b = np.exp((np.log(n_max) - np.log(n_min)) / (levels - 1))
    return np.array([np.float32(n_min * b ** l - 1.0) for l in range(levels)], np.float32)


def level_res(scale):
    return (np.ceil(scale).astype(np.int64) + 2)


def grid_index(level, cell, scale, log2_t=LOG2_T):
    """Entry index of integer corners cell [n,3] at one level."""
    t = 1 << log2_t
    res = int(level_res(scale)[level])
    c = cell.astype(np.uint64)
    if res ** 3 <= t:
        return (c[:, 0] + c[:, 1] * res + c[:, 2] * res * res).astype(np.int64)
    h = (c[:, 0] * np.uint64(PRIMES[0])) ^ (c[:, 1] * np.uint64(PRIMES[1])) ^ \
        (c[:, 2] * np.uint64(PRIMES[2]))
    return (h & np.uint64(t - 1)).astype(np.int64)


def encode_grid(pos, grid, scale):
    """pos [n,3] in [0,1]; grid [L, T, F] (fp16 values as float32).
    Returns features [n, L*F] and, for the backward pass, per level the corner
    indices [n,8] and weights [n,8]."""
    n = pos.shape[0]
    L = grid.shape[0]
    out = np.zeros((n, L * FEATS), np.float64)
    aux = []
    for l in range(L):
        p = pos.astype(np.float32) * scale[l] + np.float32(0.5)
        i0 = np.floor(p).astype(np.int64)
        w = (p - i0).astype(np.float64)
        idx = np.zeros((n, 8), np.int64)
        wt = np.zeros((n, 8), np.float64)
        for c in range(8):
            o = np.array([(c >> 0) & 1, (c >> 1) & 1, (c >> 2) & 1])
            cell = i0 + o
            wc = np.ones(n)
            for d in range(3):
                wc = wc * (w[:, d] if o[d] else (1.0 - w[:, d]))
            idx[:, c] = grid_index(l, cell, scale, int(np.log2(grid.shape[1])))
            wt[:, c] = wc
            out[:, l * FEATS:(l + 1) * FEATS] += wc[:, None] * grid[l, idx[:, c]].astype(np.float64)
        aux.append((idx, wt))
    return out, aux


def one_blob(s, k=BLOB):
    s = np.asarray(s, np.float64)[:, None]
    c = (np.arange(k) + 0.5) / k
    return np.exp(-((s - c) ** 2) * (k * k) / 2.0)


def encode_inputs(rec, grid, scale):
    """rec [n, 11] = pos(3, [0,1]), dir(2, [0,1]), normal(2, [0,1]),
    roughness(1, [0,1]), albedo(3).  Returns the padded 64-wide input row
    [n, 64] (before bf16 rounding) and the grid aux data."""
    n = rec.shape[0]
    g, aux = encode_grid(rec[:, 0:3], grid, scale)
    x = np.zeros((n, IN_PAD), np.float64)
    x[:, 0:32] = g
    x[:, 32:36] = one_blob(rec[:, 3])
    x[:, 36:40] = one_blob(rec[:, 4])
    x[:, 40:44] = one_blob(rec[:, 5])
    x[:, 44:48] = one_blob(rec[:, 6])
    x[:, 48:52] = one_blob(rec[:, 7])
    x[:, 52:55] = rec[:, 8:11]
    x[:, 55] = 1.0
    return x, aux


def weight_shapes():
    """[in, out] of every layer: 64x64 (input), 4 x 64x64 (hidden), 64x16 (output)."""
    return [(IN_PAD, WIDTH)] + [(WIDTH, WIDTH)] * (HIDDEN - 1) + [(WIDTH, OUT_PAD)]


def forward(x, weights, mixed=True):
    """x [n,64]; weights list of [in,out] float32 (fp32 masters).
    Returns output [n,3] and the per-layer inputs (after rounding) for backward."""
    r = bf16 if mixed else (lambda a: np.asarray(a, np.float64))
    acts = []
    h = r(x)
    for li, w in enumerate(weights):
        acts.append(np.asarray(h, np.float64))
        y = np.asarray(h, np.float64) @ np.asarray(r(w), np.float64)
        if li < len(weights) - 1:
            y = np.maximum(y, 0.0)
            h = r(y.astype(np.float32)) if mixed else y
        else:
            h = y
    return h[:, :OUT], acts


def loss_grad(y, t, batch):
    """Relative L2: mean of (y-t)^2 / (y^2 + 0.01) over batch*3 values; y is
    treated as a constant in the denominator.  Returns (loss, dL/dy)."""
    den = np.asarray(y, np.float64) ** 2 + 0.01
    d = np.asarray(y, np.float64) - np.asarray(t, np.float64)
    loss = float(np.sum(d * d / den) / (batch * OUT))
    return loss, 2.0 * d / den / (batch * OUT)


def backward(dy, acts, weights, mixed=True):
    """Gradients of every weight matrix and of the 64-wide input row.
    dy [n,3]; acts from forward().  The upstream gradient entering each
    product is rounded to bf16 as on the GPU."""
    r = bf16 if mixed else (lambda a: np.asarray(a, np.float64))
    n = dy.shape[0]
    g = np.zeros((n, OUT_PAD), np.float64)
    g[:, :OUT] = dy
    grads = [None] * len(weights)
    for li in range(len(weights) - 1, -1, -1):
        gz = np.asarray(r(g.astype(np.float32)), np.float64)
        grads[li] = acts[li].T @ gz
        gx = gz @ np.asarray(r(weights[li]), np.float64).T
        if li > 0:
            gx = gx * (acts[li] > 0)          # ReLU of the previous layer
        g = gx
    return grads, g


def grid_grad(dx, aux, grid_shape):
    """Scatter the input-row gradient of the 32 grid features to the grid."""
    gg = np.zeros(grid_shape, np.float64)
    for l, (idx, wt) in enumerate(aux):
        for f in range(FEATS):
            gf = dx[:, l * FEATS + f]
            for c in range(8):
                np.add.at(gg[l, :, f], idx[:, c], wt[:, c] * gf)
    return gg


def adam(param, grad, m, v, step, lr=1e-2, b1=0.9, b2=0.99, eps=1e-15):
    """One Adam step (Kingma & Ba 2015) with bias correction; float64."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return param - lr * mh / (np.sqrt(vh) + eps), m, v
