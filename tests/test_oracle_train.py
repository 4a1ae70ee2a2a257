"""Pins the oracle's training half: loss gradients and the full backward against
central finite differences (eps = 1e-3 on 256 random parameters), grid gradients
against the forward's linearity, Adam against its closed form for a constant gradient."""
import numpy as np
import pytest

import nrrs_amd as na

N = 96


def setup(ora, n_pixels=N, **kw):
    cfg = na.default_config(n_pixels=n_pixels, grid=dict(log2_table_size=12), **kw)
    d = ora.synth_vertices(N, 5, n_pixels)
    hv = ora.HostVertices(d)
    film = np.random.default_rng(3).lognormal(-1, 1, (3, n_pixels)).astype(np.float32)
    return cfg, hv, film


def fd_check(ora, cfg, net, hv, params, loss_of, grad, n_check=256, eps=1e-3, seed=0):
    """Central differences (step eps) on n_check parameters: half among the touched
    grid entries, half among the MLP weights. A step of 1e-3 can straddle a ReLU
    kink of some sample; such parameters (at most 1/8) are re-checked with a 1e-5
    step. Returns the worst error relative to the largest checked gradient."""
    rng = np.random.default_rng(seed)
    g, m = ora.param_count(cfg, net)
    touched = np.nonzero(grad[:g])[0]
    idx = np.concatenate([rng.choice(touched, n_check // 2, replace=False),
                          g + rng.choice(m, n_check // 2, replace=False)])
    scale = np.max(np.abs(grad[idx])) + 1e-30

    def fd(k, h):
        p1 = params.copy()
        p1[k] += h
        p2 = params.copy()
        p2[k] -= h
        return (loss_of(p1) - loss_of(p2)) / (2 * h)

    worst, kinks = 0.0, 0
    for k in idx:
        err = abs(fd(k, eps) - grad[k]) / scale
        if err > 1e-4:  # central-difference truncation at 1e-3, or a kink
            kinks += 1
            err = abs(fd(k, 1e-5) - grad[k]) / scale
        worst = max(worst, err)
    assert kinks <= n_check // 8, kinks
    return worst


def test_ll2_gradient_fd(ora):
    cfg, hv, film = setup(ora)
    params = ora.init_params(cfg, na.STATNET, 7, 0.3)
    tL = ora.synth_targets(N, 1)
    val, grad, dy = ora.train_grads(cfg, na.STATNET, hv, params, na.LOSS_LL2 if hasattr(na, "LOSS_LL2") else 1,
                                    film=film, target_L=tL)
    y0 = ora.net_forward(cfg, na.STATNET, hv, film, None, params).T.copy()

    def loss_of(p):  # frozen-denominator surrogate of the relative L2
        y = ora.net_forward(cfg, na.STATNET, hv, film, None, p).T
        return ora.loss(cfg, 1, hv, y, target_L=tL, y_den=y0)[0]

    assert val == pytest.approx(ora.loss(cfg, 1, hv, y0, target_L=tL)[0], rel=1e-12)
    assert fd_check(ora, cfg, na.STATNET, hv, params, loss_of, grad) < 1e-4


def rrs_side(n, seed):
    rng = np.random.default_rng(seed)
    qt = rng.uniform(0.05, 5.0, n)
    qt[rng.uniform(size=n) < 0.1] = -1.0  # duplicate nodes
    return dict(ref_mean=None, error=rng.uniform(0.0, 2.0, n), sample_weight=1.0 / rng.integers(1, 4, n),
                q_target=qt, error_sum=float(n) * 0.7)


def test_rrs_gradient_fd(ora):
    cfg, hv, film = setup(ora)
    params = ora.init_params(cfg, na.RRSNET, 8, 0.3)
    stats = np.abs(np.random.default_rng(2).normal(0.5, 0.5, (6, N)))
    side = ora.RrsSide(**rrs_side(N, 4))
    val, grad, dy = ora.train_grads(cfg, na.RRSNET, hv, params, 2, film=film, stats=stats, side=side)
    assert val > 0 and np.count_nonzero(dy) > N // 2

    def loss_of(p):
        y = ora.net_forward(cfg, na.RRSNET, hv, film, stats, p).T
        return ora.loss(cfg, 2, hv, y, film=film, stats=stats, side=side)[0]

    assert fd_check(ora, cfg, na.RRSNET, hv, params, loss_of, grad) < 1e-4


def test_rrs_loss_structure(ora):
    """Duplicates (q_t = -1) and steps < 3 contribute nothing; the loss is linear in the weights."""
    cfg, hv, film = setup(ora)
    stats = np.ones((6, N))
    y = np.random.default_rng(5).normal(0, 1, (1, N))
    s = rrs_side(N, 6)
    s["q_target"][:] = -1.0
    assert ora.loss(cfg, 2, hv, y, film=film, stats=stats, side=ora.RrsSide(**s))[0] == 0.0
    s = rrs_side(N, 6)
    a = ora.loss(cfg, 2, hv, y, film=film, stats=stats, side=ora.RrsSide(**s))[0]
    s2 = dict(s, sample_weight=2 * s["sample_weight"])
    assert ora.loss(cfg, 2, hv, y, film=film, stats=stats, side=ora.RrsSide(**s2))[0] == pytest.approx(2 * a)
    cfg2, _, _ = setup(ora, rrs_step=2)
    assert ora.loss(cfg2, 2, hv, y, film=film, stats=stats, side=ora.RrsSide(**s))[0] == 0.0


def test_ll2_closed_form(ora):
    """y = target gives zero loss/gradient; a unit offset gives 1/(y^2+0.01) per element."""
    cfg, hv, _ = setup(ora, loss_clamp_on=0)
    tL = np.full((3, N), 2.0, np.float32)
    y = np.vstack([np.full((3, N), 2.0), np.full((3, N), 4.0)])
    val, dy = ora.loss(cfg, 1, hv, y, target_L=tL)
    assert val == 0.0 and not dy.any()
    y2 = y + 1.0
    val, dy = ora.loss(cfg, 1, hv, y2, target_L=tL)
    expect = (3 / (9 + 0.01) + 3 / (25 + 0.01)) / 6
    assert val == pytest.approx(expect, rel=1e-12)
    cfg0, _, _ = setup(ora, loss_clamp_on=0, ll2_step=0)
    val0, dy0 = ora.loss(cfg0, 1, hv, y2, target_L=tL)
    assert val0 == pytest.approx(1 / (9 + 0.01), rel=1e-12) and not dy0[3:].any()


def test_grid_gradient_is_linear_adjoint(ora):
    """Forward is linear in the table: <dL/dtable, T> equals <dL/denc_grid, enc_grid(T)>
    for an external dL/dy through a net whose first layer maps grid channels only."""
    cfg, hv, film = setup(ora)
    g, m = ora.param_count(cfg, na.STATNET)
    rng = np.random.default_rng(9)
    params = np.zeros(g + m)
    params[:g] = rng.uniform(-1, 1, g)
    # MLP = identity-ish: W0 picks grid channel k for out k (k < 32), other layers pass through
    W0 = np.zeros((64, 80))
    W0[np.arange(32), np.arange(32)] = 1.0
    Ws = [W0] + [np.eye(64)] * 3
    Wl = np.zeros((6, 64))
    Wl[:, :6] = rng.normal(0, 1, (6, 6))
    params[g:] = np.concatenate([w.ravel() for w in Ws + [Wl]])
    params_pos = params.copy()
    params_pos[:g] = np.abs(params[:g]) + 1.0  # keep every ReLU active
    dy = rng.normal(0, 1, (6, N))
    _, grad, _ = ora.train_grads(cfg, na.STATNET, hv, params_pos, 0, film=film, dy_ext=dy)
    enc = ora.encode(cfg, na.STATNET, hv, film, None, params_pos)
    # L = <dy, Wl enc[:, :64]> -> dL/denc_grid = (Wl^T dy)[:32]
    denc = (Wl.T @ dy)[:32]
    lhs = np.dot(grad[:g], params_pos[:g])
    rhs = np.sum(denc * enc[:, :32].T)
    assert lhs == pytest.approx(rhs, rel=1e-10)


def test_adam_constant_gradient_closed_form(ora):
    """With a constant gradient the bias-corrected moments are exact: every step
    moves each parameter by lr * sign(g) (eps negligible)."""
    p = np.array([1.0, -2.0, 0.5, 3.0])
    m = np.zeros(4)
    v = np.zeros(4)
    g = np.array([0.3, -0.7, 1e-3, 0.0])
    for t in range(1, 4):
        ora.adam(p, m, v, g, 1e-2, 0.9, 0.99, 1e-15, 0.0, t)
    assert np.allclose(p, [1.0 - 0.03, -2.0 + 0.03, 0.5 - 0.03, 3.0], atol=1e-12)
    # l2_reg acts as an extra gradient l2 * p
    p = np.array([2.0])
    m, v = np.zeros(1), np.zeros(1)
    ora.adam(p, m, v, np.array([0.0]), 1e-2, 0.9, 0.99, 1e-15, 0.5, 1)
    assert p[0] == pytest.approx(2.0 - 1e-2)


def test_ema_is_the_debiased_weighted_average(ora):
    """The Ema recursion (2net-simple.json:16-18) equals the explicit debiased average
    sum_k (1 - d) d^(t-k) p_k / (1 - d^t) of the parameter trajectory: its first value is the
    first step's parameters, a constant trajectory stays constant, and for d = 0.95 over 12
    random steps it matches the weighted sum (a missing debias or an undamped old average fails)."""
    rng = np.random.default_rng(0)
    traj = rng.normal(0, 1, (12, 5))
    for d in (0.95, 0.5):
        ema = np.full(5, 123.0)  # the value before the first step does not matter
        for t, p in enumerate(traj, start=1):
            ora.ema_step(ema, p, d, t)
            w = (1.0 - d) * d ** (t - np.arange(1, t + 1))
            want = (w[:, None] * traj[:t]).sum(axis=0) / (1.0 - d ** t)
            assert np.allclose(ema, want, rtol=1e-12, atol=1e-12)
            if t == 1:
                assert np.allclose(ema, traj[0], rtol=1e-15, atol=1e-15)
    ema = np.zeros(3)
    for t in range(1, 6):
        ora.ema_step(ema, np.array([1.5, -2.0, 0.25]), 0.95, t)
    assert np.allclose(ema, [1.5, -2.0, 0.25], rtol=1e-14)
