"""Pins the learned-guide oracle (oracle/mlp_oracle.py) by mathematics alone: the
analytic backward pass against FP64 central finite differences, the vMF density
against its normalisation integral, the stable vMF sampler against the closed-form
mean cosine, the mixture sampler against its density, and the BF16 rounding against
numpy's float32 bit layout. CPU only."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo

LOG2T = 12


def _random_setup(rng, n=6, scale=0.6):
    params = np.zeros(mo.N_PARAMS)
    for ow, ob, rows, cols in mo.LAYERS:
        used_r = mo.OUTPUTS_USED if rows == mo.OUTPUTS else rows
        used_c = mo.INPUTS_USED if cols == mo.INPUTS else cols
        W = np.zeros((rows, cols))
        W[:used_r, :used_c] = rng.uniform(-scale, scale, (used_r, used_c))
        params[ow: ow + rows * cols] = W.reshape(-1)
        params[ob: ob + used_r] = rng.uniform(-0.2, 0.2, used_r)
    res, dense = mo.levels(4.0, 64.0, LOG2T)
    grid = rng.uniform(-0.5, 0.5, (mo.LEVELS, 1 << LOG2T, 2))
    frame = np.array([-1.0, -1.0, -1.0, 0.5])
    x = rng.uniform(-0.95, 0.95, (n, 3))
    def unit(k):
        v = rng.normal(size=(k, 3))
        return v / np.linalg.norm(v, axis=1, keepdims=True)
    s = np.zeros((n, 16))
    s[:, 0:3] = x
    s[:, 3] = rng.uniform(0.2, 2.0, n)
    s[:, 4:7] = unit(n)
    s[:, 7] = rng.uniform(0.05, 1.0, n)
    s[:, 8:11] = unit(n)
    s[:, 12:15] = unit(n)
    return params, grid, res, dense, frame, s


def test_bf16_rounding_matches_bit_layout():
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 65504.0, 1e-30, np.inf], dtype=np.float32)
    r = mo.bf16(x)
    # every result is representable: low 16 bits of the float32 pattern are zero
    assert np.all((r.astype(np.float32).view(np.uint32) & 0xFFFF) == 0)
    # ties to even: 1 + 2^-8 lies halfway between 1 and 1 + 2^-7 -> 1
    assert r[1] == 1.0
    # above the tie rounds up
    assert r[2] == np.float32(1.0078125)
    assert abs(r[3] - (-3.140625)) < 1e-12
    assert np.isinf(r[6])
    assert np.isnan(mo.bf16(np.float32(np.nan)))


def test_levels_schedule():
    res, dense = mo.levels(16.0, 1024.0, 19)
    assert res[0] == 16 and res[-1] == 1024
    assert np.all(np.diff(res.astype(np.int64)) > 0)
    # dense while (res+1)^3 fits 2^19 entries
    for r, d in zip(res, dense):
        assert d == ((int(r) + 1) ** 3 <= 1 << 19)


@pytest.mark.parametrize("kappa", [0.003, 0.4, 3.0, 40.0, 600.0, 8000.0])
def test_vmf_density_normalised(kappa):
    # integrate over cos(theta) around the axis (the density depends on it alone)
    mu = np.array([0.0, 0.0, 1.0])
    t = np.linspace(-1.0, 1.0, 2_000_001)
    w = np.stack([np.sqrt(1 - t * t), np.zeros_like(t), t], 1)
    p = mo.vmf_pdf(mu, np.float64(kappa), w)
    integral = 2 * np.pi * np.trapezoid(p, t)
    assert abs(integral - 1.0) < 1e-4


@pytest.mark.parametrize("kappa", [0.01, 1.0, 10.0, 300.0])
def test_vmf_sampler_mean_cosine(kappa):
    rng = np.random.default_rng(5)
    mu = np.array([0.3, -0.5, 0.8]); mu /= np.linalg.norm(mu)
    n = 40000
    u1, u2 = rng.random(n), rng.random(n)
    d = np.stack([mo.vmf_sample(mu, kappa, a, b) for a, b in zip(u1[:4000], u2[:4000])])
    # closed form E[mu . w] = coth(kappa) - 1/kappa
    expect = 1.0 / np.tanh(kappa) - 1.0 / kappa
    got = (d @ mu).mean()
    sd = np.sqrt(max((d @ mu).var(), 1e-12) / len(d))
    assert abs(got - expect) < 5 * sd + 1e-6
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-12)


def test_mixture_sampler_matches_density():
    rng = np.random.default_rng(9)
    pi = rng.dirichlet(np.ones(8))
    mu = rng.normal(size=(8, 3)); mu /= np.linalg.norm(mu, axis=1, keepdims=True)
    kappa = np.exp(rng.uniform(-1, 4, 8))
    n = 20000
    u = rng.random((n, 3))
    d = np.stack([mo.mix_sample(pi, mu, kappa, *row) for row in u])
    # chi-square over 12 x 24 equal-area bins of the cylindrical map
    ui = np.clip(((np.arctan2(d[:, 1], d[:, 0]) / (2 * np.pi)) % 1.0 * 24).astype(int), 0, 23)
    vi = np.clip(((d[:, 2] + 1) * 0.5 * 12).astype(int), 0, 11)
    hist = np.zeros((12, 24)); np.add.at(hist, (vi, ui), 1)
    # expected counts by quadrature of the mixture density over each bin
    g = 8
    ph = (np.arange(24 * g) + 0.5) / (24 * g) * 2 * np.pi
    ct = (np.arange(12 * g) + 0.5) / (12 * g) * 2 - 1
    P, C = np.meshgrid(ph, ct)
    S = np.sqrt(1 - C * C)
    w = np.stack([S * np.cos(P), S * np.sin(P), C], -1).reshape(-1, 3)
    dens = np.sum(pi * mo.vmf_pdf(mu[None], kappa[None], w[:, None, :]), 1).reshape(12 * g, 24 * g)
    cell = 4 * np.pi / dens.size
    exp = (dens * cell).reshape(12, g, 24, g).sum((1, 3)) * n
    keep = exp > 5
    chi2 = np.sum((hist[keep] - exp[keep]) ** 2 / exp[keep])
    dof = keep.sum() - 1
    assert chi2 < dof + 6 * np.sqrt(2 * dof), (chi2, dof)


def test_backward_matches_central_differences():
    rng = np.random.default_rng(3)
    params, grid, res, dense, frame, s = _random_setup(rng)

    def total(p, g):
        return mo.loss_and_grads(p, g, s, res, dense, LOG2T, frame, emulate=False)[0]

    loss, gp, gg = mo.loss_and_grads(params, grid, s, res, dense, LOG2T, frame, emulate=False)
    h = 1e-6
    # every layer's weights and biases, sampled
    idx = []
    for ow, ob, rows, cols in mo.LAYERS:
        used_r = mo.OUTPUTS_USED if rows == mo.OUTPUTS else rows
        used_c = mo.INPUTS_USED if cols == mo.INPUTS else cols
        for _ in range(12):
            idx.append(ow + rng.integers(used_r) * cols + rng.integers(used_c))
        for _ in range(4):
            idx.append(ob + rng.integers(used_r))
    # the selection logit, kappa logits, mixture logits and lobe axes of W4
    for o in (40, 32, 35, 0, 7, 8, 20, 31):
        idx.append(mo.OFF_B4 + o)
    worst = 0.0
    for i in idx:
        pp, pm = params.copy(), params.copy()
        pp[i] += h; pm[i] -= h
        fd = (total(pp, grid) - total(pm, grid)) / (2 * h)
        err = abs(fd - gp[i]) / max(1e-6, abs(fd) + abs(gp[i]))
        worst = max(worst, err)
        assert err < 1e-5 or abs(fd - gp[i]) < 1e-9, (i, fd, gp[i])
    # hash-table entries that the samples touch
    touched = np.argwhere(np.abs(gg) > 0)
    for l, e, f in touched[rng.choice(len(touched), 24, replace=False)]:
        gp_, gm_ = grid.copy(), grid.copy()
        gp_[l, e, f] += h; gm_[l, e, f] -= h
        fd = (total(params, gp_) - total(params, gm_)) / (2 * h)
        assert abs(fd - gg[l, e, f]) <= 1e-5 * max(1e-6, abs(fd)) + 1e-9, (l, e, f, fd, gg[l, e, f])
    # a zero-weight sample contributes nothing
    s0 = s.copy(); s0[0, 3] = 0.0
    l0, g0, _ = mo.loss_and_grads(params, grid, s0[:1], res, dense, LOG2T, frame, emulate=False)
    assert l0 == 0.0 and not np.any(g0)


def test_loss_equals_direct_mixture_likelihood():
    """The log-domain loss equals -w log(c q + (1-c) p_bsdf) evaluated directly."""
    rng = np.random.default_rng(4)
    params, grid, res, dense, frame, s = _random_setup(rng, n=32)
    feats, _ = mo.encode(s[:, 0:3], s[:, 4:7], s[:, 8:11], grid, res, dense, LOG2T, frame, emulate=False)
    out, _ = mo.forward(params, feats, emulate=False)
    pi, mu, kappa, c = mo.mixture(out)
    q = mo.mix_pdf(pi, mu, kappa, s[:, 12:15])
    direct = -s[:, 3] * np.log(c * q + (1 - c) * s[:, 7]) / len(s)
    loss, _ = mo.loss_grad_out(out, s[:, 12:15], s[:, 7], s[:, 3], 1.0 / len(s))
    assert np.allclose(loss, direct, rtol=1e-12, atol=0)


def test_adam_first_step_is_lr_times_sign():
    g = np.array([3.0, -0.5, 1e-3])
    th, m, v = mo.adam(np.zeros(3), np.zeros(3), np.zeros(3), g, 1, lr=0.01)
    assert np.allclose(th, -0.01 * np.sign(g), rtol=1e-9)
