"""FP64 oracle of the learned directional guide — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline may import this
module, as the checker of the GPU implementation; nothing in the product calls it.

The reference has no learned guide (its guiding outline ends at recording PathVertex
data, Renderer.h:398-414), so this restates the published methods the GPU module
implements:
  * multiresolution hash encoding (Mueller et al. 2022, "Instant Neural Graphics
    Primitives"): 16 levels x 2 features, res_l = floor(res_min b^l), trilinear
    interpolation, dense indexing while (res_l+1)^3 fits the table, otherwise the
    xor hash with primes 1, 2654435761, 805459861;
  * one-blob encoding (Mueller et al. 2019, "Neural Importance Sampling"), 8 bins,
    Gaussian kernel of width 1/8;
  * a 80 -> 64 -> 64 -> 64 -> 41 ReLU network;
  * an 8-lobe von Mises-Fisher mixture on the sphere with a one-sample-MIS selection
    probability c = 0.05 + 0.9 sigmoid(s), trained by the weighted negative
    log-likelihood -w log(c q(wi) + (1 - c) p_bsdf(wi)); stable vMF sampling after
    Jakob, "Numerically stable sampling of the von Mises Fisher distribution on S2";
  * Adam (Kingma & Ba 2015) with bias correction.
The analytic backward pass here is checked against central finite differences in
FP64 (tests/test_mlp_oracle.py); `emulate=True` applies the BF16 rounding the GPU
applies to weights, encoded inputs, activations and deltas.
"""
from __future__ import annotations

import numpy as np

LEVELS = 16
INPUTS, INPUTS_USED = 96, 80
HIDDEN = 64
OUTPUTS, OUTPUTS_USED = 48, 41
LOBES = 8
K_LOG_MIN, K_LOG_MAX = -6.0, 9.0
SEL_MIN, SEL_RANGE = 0.05, 0.9

OFF_W1 = 0
OFF_B1 = OFF_W1 + HIDDEN * INPUTS
OFF_W2 = OFF_B1 + HIDDEN
OFF_B2 = OFF_W2 + HIDDEN * HIDDEN
OFF_W3 = OFF_B2 + HIDDEN
OFF_B3 = OFF_W3 + HIDDEN * HIDDEN
OFF_W4 = OFF_B3 + HIDDEN
OFF_B4 = OFF_W4 + OUTPUTS * HIDDEN
N_PARAMS = OFF_B4 + OUTPUTS
LAYERS = [(OFF_W1, OFF_B1, HIDDEN, INPUTS), (OFF_W2, OFF_B2, HIDDEN, HIDDEN),
          (OFF_W3, OFF_B3, HIDDEN, HIDDEN), (OFF_W4, OFF_B4, OUTPUTS, HIDDEN)]


def bf16(x):
    """Round to the nearest BF16 (ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = np.isnan(f)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = r.astype(np.uint32)
    r = np.where(nan, (u.astype(np.uint32) | 0x00400000) & 0xFFFF0000, r).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def levels(res_min=16.0, res_max=1024.0, log2_table=19):
    """Per-level resolution and dense flag (mirrors pg_mlp_init)."""
    T = 1 << log2_table
    b = np.exp((np.log(res_max) - np.log(res_min)) / (LEVELS - 1))
    res = np.array([max(1, int(np.floor(res_min * b ** l + 1e-6))) for l in range(LEVELS)], dtype=np.uint64)
    dense = np.array([(int(r) + 1) ** 3 <= T for r in res])
    return res, dense


def _entry(l, x, y, z, res, dense, T):
    x = x.astype(np.uint64); y = y.astype(np.uint64); z = z.astype(np.uint64)
    if dense[l]:
        n1 = np.uint64(int(res[l]) + 1)
        return x + n1 * (y + n1 * z)
    m = np.uint64(0xFFFFFFFF)
    h = (x * np.uint64(1)) & m
    h ^= (y * np.uint64(2654435761)) & m
    h ^= (z * np.uint64(805459861)) & m
    return h & np.uint64(T - 1)


def unit_position(x, frame):
    """Position in [0,1]^3 of the scene cube, in FP32 as the device computes it."""
    lo = np.asarray(frame[:3], dtype=np.float32)
    inv = np.float32(frame[3])
    p = (np.asarray(x, dtype=np.float32) - lo) * inv
    return np.clip(p, np.float32(0.0), np.float32(1.0))


def hash_cells(p01, res, l):
    """Cell corner and fractional position at level l (FP32 like the device)."""
    r = np.float32(int(res[l]))
    p = p01 * r
    i = np.floor(p).astype(np.int64)
    i = np.clip(i, 0, int(res[l]) - 1)
    f = (p - i.astype(np.float32)).astype(np.float32)
    return i, f.astype(np.float64)


def encode(x, n, wo, grid, res, dense, log2_table, frame, emulate=True):
    """Inputs (N, 96) and the corner cache of the hash features.

    grid: (LEVELS, T, 2) feature values (as the device holds them)."""
    T = 1 << log2_table
    N = len(x)
    feats = np.zeros((N, INPUTS))
    p01 = unit_position(x, frame)
    cache = []
    for l in range(LEVELS):
        i, f = hash_cells(p01, res, l)
        ents, ws = [], []
        for c in range(8):
            cx, cy, cz = c & 1, (c >> 1) & 1, c >> 2
            e = _entry(l, i[:, 0] + cx, i[:, 1] + cy, i[:, 2] + cz, res, dense, T).astype(np.int64)
            w = ((f[:, 0] if cx else 1 - f[:, 0]) * (f[:, 1] if cy else 1 - f[:, 1])) * (f[:, 2] if cz else 1 - f[:, 2])
            feats[:, 2 * l] += w * grid[l, e, 0]
            feats[:, 2 * l + 1] += w * grid[l, e, 1]
            ents.append(e)
            ws.append(w)
        cache.append((np.stack(ents, 1), np.stack(ws, 1)))
    comps = np.concatenate([np.asarray(n, float), np.asarray(wo, float)], 1)
    t = np.arange(8) + 0.5
    for k in range(6):
        s = (comps[:, k] + 1.0) * 0.5
        feats[:, 32 + 8 * k: 40 + 8 * k] = np.exp(-0.5 * (t[None, :] - 8.0 * s[:, None]) ** 2)
    if emulate:
        feats = bf16(feats)
    return feats, cache


def split(params):
    out = []
    for ow, ob, rows, cols in LAYERS:
        out.append((params[ow: ow + rows * cols].reshape(rows, cols), params[ob: ob + rows]))
    return out


def forward(params, feats, emulate=True):
    layers = split(np.asarray(params, float))
    acts = [feats]
    zs = []
    a = feats
    for k, (W, b) in enumerate(layers):
        Wk = bf16(W) if emulate else W
        z = a @ Wk.T + b
        zs.append(z)
        if k < 3:
            a = np.maximum(z, 0.0)
            if emulate:
                a = bf16(a)
            acts.append(a)
    return zs[-1], (acts, zs)


def mixture(out):
    """pi (N,8), mu (N,8,3), kappa (N,8), c (N,) of network outputs (N,48)."""
    a = out[:, :8]
    e = np.exp(a - a.max(1, keepdims=True))
    pi = e / e.sum(1, keepdims=True)
    m = out[:, 8:32].reshape(-1, 8, 3)
    ln = np.linalg.norm(m, axis=2, keepdims=True)
    mu = np.where(ln > 1e-20, m / np.maximum(ln, 1e-300), np.array([0.0, 0.0, 1.0]))
    kappa = np.exp(np.clip(out[:, 32:40], K_LOG_MIN, K_LOG_MAX))
    c = SEL_MIN + SEL_RANGE / (1.0 + np.exp(-out[:, 40]))
    return pi, mu, kappa, c


def vmf_pdf(mu, kappa, w):
    """Density of vMF lobes (..., 3)/(...) at directions w (..., 3)."""
    t = np.sum(mu * w, -1)
    return kappa / (2 * np.pi * (-np.expm1(-2 * kappa))) * np.exp(kappa * (t - 1.0))


def mix_pdf(pi, mu, kappa, w):
    return np.sum(pi * vmf_pdf(mu, kappa, w[:, None, :]), 1)


def frame_from(nv):
    """Orthonormal frame around n (the Frame construction the GPU shares with the
    reference's shading frame, MyMath.h:406-433)."""
    w = nv / np.linalg.norm(nv)
    if abs(w[0]) > abs(w[1]):
        l = 1.0 / np.sqrt(w[0] * w[0] + w[2] * w[2])
        u = np.array([w[2] * l, 0.0, -w[0] * l])
    else:
        l = 1.0 / np.sqrt(w[1] * w[1] + w[2] * w[2])
        u = np.array([0.0, w[2] * l, -w[1] * l])
    v = np.cross(w, u)
    return u, v, w


def vmf_sample(mu, kappa, u1, u2):
    ct = 1.0 + np.log1p((1.0 - u1) * np.expm1(-2.0 * kappa)) / kappa
    ct = np.clip(ct, -1.0, 1.0)
    st = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
    ph = 2 * np.pi * u2
    u, v, w = frame_from(mu)
    d = u * st * np.cos(ph) + v * st * np.sin(ph) + w * ct
    return d / np.linalg.norm(d)


def mix_sample(pi, mu, kappa, u0, u1, u2):
    cdf = np.cumsum(pi)
    j = int(np.sum(u0 >= cdf[:-1]))
    return vmf_sample(mu[j], kappa[j], u1, u2)


def loss_grad_out(out, wi, pb, w, inv_b):
    """Per-sample loss -w log P * inv_b and d loss / d outputs (N, 48)."""
    N = len(out)
    a = out[:, :8]
    amax = a.max(1, keepdims=True)
    e = np.exp(a - amax)
    Z = e.sum(1, keepdims=True)
    pi = e / Z
    m = out[:, 8:32].reshape(N, 8, 3)
    ln = np.linalg.norm(m, axis=2)
    mu = m / np.maximum(ln, 1e-300)[..., None]
    kr = out[:, 32:40]
    kc = np.clip(kr, K_LOG_MIN, K_LOG_MAX)
    clamped = kc != kr
    k = np.exp(kc)
    t = np.einsum("nij,nj->ni", mu, wi)
    lognorm = kc - np.log(2 * np.pi) - np.log(-np.expm1(-2 * k))
    lq = (a - amax - np.log(Z)) + lognorm + k * (t - 1.0)
    M = lq.max(1, keepdims=True)
    S = np.exp(lq - M).sum(1, keepdims=True)
    sg = 1.0 / (1.0 + np.exp(-out[:, 40]))
    cc = SEL_MIN + SEL_RANGE * sg
    la = np.log(cc) + M[:, 0] + np.log(S[:, 0])
    with np.errstate(divide="ignore"):
        lb = np.where(pb > 0, np.log(1 - cc) + np.log(np.maximum(pb, 1e-300)), -np.inf)
    logP = np.logaddexp(la, lb)
    G = (w * inv_b)
    rv = np.exp(la - logP)
    rb = np.where(pb > 0, np.exp(lb - logP), 0.0)
    rj = np.exp(np.log(cc)[:, None] + lq - logP[:, None])
    d = np.zeros((N, OUTPUTS))
    d[:, :8] = -G[:, None] * (rj - pi * rv[:, None])
    with np.errstate(over="ignore"):  # 2k / expm1(2k) -> 0 for sharp lobes
        ratio = 2 * k / np.expm1(2 * k)
    d[:, 32:40] = np.where(clamped, 0.0, -G[:, None] * rj * ((1.0 - ratio) + k * (t - 1.0)))
    coef = -G[:, None] * rj * k / ln
    dm = coef[..., None] * (wi[:, None, :] - t[..., None] * mu)
    d[:, 8:32] = dm.reshape(N, 24)
    d[:, 40] = -G * (rv / cc - rb / (1 - cc)) * (SEL_RANGE * sg * (1 - sg))
    live = w > 0
    d[~live] = 0.0
    loss = np.where(live, -w * logP * inv_b, 0.0)
    return loss, d


def backward(params, cache, dout, emulate=True):
    """Gradients of the parameters (N_PARAMS) and of the inputs (N, 96)."""
    acts, zs = cache
    layers = split(np.asarray(params, float))
    g = np.zeros(N_PARAMS)
    d = bf16(dout) if emulate else dout
    for k in range(3, -1, -1):
        ow, ob, rows, cols = LAYERS[k]
        W = layers[k][0]
        Wk = bf16(W) if emulate else W
        g[ow: ow + rows * cols] = (d.T @ acts[k]).reshape(-1)
        g[ob: ob + rows] = d.sum(0)
        dprev = d @ Wk
        if k > 0:
            dprev = dprev * (zs[k - 1] > 0)
            if emulate:
                dprev = bf16(dprev)
        d = dprev
    g[OFF_W1: OFF_B1].reshape(HIDDEN, INPUTS)[:, INPUTS_USED:] = 0.0
    return g, d


def grid_gradient(dfeat, enc_cache, log2_table):
    T = 1 << log2_table
    gg = np.zeros((LEVELS, T, 2))
    for l in range(LEVELS):
        ents, ws = enc_cache[l]
        for c in range(8):
            np.add.at(gg[l, :, 0], ents[:, c], ws[:, c] * dfeat[:, 2 * l])
            np.add.at(gg[l, :, 1], ents[:, c], ws[:, c] * dfeat[:, 2 * l + 1])
    return gg


def loss_and_grads(params, grid, samples, res, dense, log2_table, frame, emulate=True):
    """Mean loss over the samples (N, 16) and its gradients (params, grid)."""
    s = np.asarray(samples, float).reshape(-1, 16)
    x, w = s[:, 0:3], s[:, 3]
    nrm, pb = s[:, 4:7], s[:, 7]
    wo, wi = s[:, 8:11], s[:, 12:15]
    feats, enc = encode(x, nrm, wo, grid, res, dense, log2_table, frame, emulate)
    out, cache = forward(params, feats, emulate)
    inv_b = 1.0 / len(s)
    loss, dout = loss_grad_out(out, wi, pb, w, inv_b)
    gp, dfeat = backward(params, cache, dout, emulate)
    gg = grid_gradient(dfeat, enc, log2_table)
    return loss.sum(), gp, gg


def adam(theta, m, v, g, t, lr=1e-2, b1=0.9, b2=0.99, eps=1e-15):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    th = theta - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return th, m, v
