"""The joint training step (nrrs_train_step with both nets, the reference's NRRSNetwork::train:
network.cpp:196-295 -> myNetworkLL2.cpp:41-89 -> train.cu:155-240) against the FP64 oracle:
StatNet forward -> NRRS_LL2 -> backward -> Adam, and on the same batch RRSNet fed with the
StatNet statistics of that forward -> NRRS_RRS -> backward -> Adam.

The oracle runs the same composition: StatNet statistics from its FP64 forward with the
parameters before the step, both nets' FP64 gradients, then its Adam. With adam_eps far above
every gradient and lr = K eps, the first Adam step moves a parameter by -K g (1 - |g| / eps),
so the parameter change measures the gradient of each net through the whole joint step; with
the default knobs the first step moves each parameter by -lr sign(g), which is checked on every
entry whose gradient stands clear of the rounding noise. A batch at the benchmark size
(2^20 samples, BF16) is checked through properties: finite and decreasing losses, the
paired-table path bitwise equal to the per-net path, and a sane loss-scale state."""
import numpy as np
import pytest

import nrrs_amd as na

pytestmark = pytest.mark.gpu

# Gradient error relative to the largest gradient entry (MLP, entry-wise) and norm-wise error of
# the table gradient, per precision of the MFMA operands (as for the single-net backward).
GRAD_TOL = {na.FP32: 2e-4, na.FP16: 1e-2, na.BF16: 5e-2}
GRID_NORM_TOL = {na.FP32: 2e-4, na.FP16: 5e-2, na.BF16: 1.5e-1}
PRECISIONS = [na.FP32, na.FP16, na.BF16]
# RRSNet's gradients are small (its loss is O(0.1)): its larger ratio keeps the parameter change
# far above the FP32 rounding of the master weights.
LINEAR_ADAM = dict(lr_statnet=1e4, lr_rrsnet=1e7, adam_eps=1e4)


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    assert torch.cuda.is_available()
    return torch


def _batch(torch, ora, ctx, n, seed):
    rng = np.random.default_rng(seed)
    for net in (na.STATNET, na.RRSNET):
        p = ctx.params_get(net)
        g, _ = ctx.param_count(net)
        p[:g] = rng.uniform(-0.3, 0.3, g).astype(np.float32)
        ctx.params_set(net, p)
    vb = na.synth_vertices(n, 500 + seed, ctx.cfg.n_pixels)
    film_np = rng.lognormal(-1, 1, (3, ctx.cfg.n_pixels)).astype(np.float32)
    tL = ora.synth_targets(n, seed + 1)
    qt = rng.uniform(0.05, 5.0, n).astype(np.float32)
    qt[rng.uniform(size=n) < 0.1] = -1.0  # duplicate nodes
    side = dict(error=rng.uniform(0, 2, n).astype(np.float32),
                sample_weight=(1.0 / rng.integers(1, 4, n)).astype(np.float32), q_target=qt,
                error_sum=float(ctx.cfg.n_pixels) * 0.7)
    return vb, film_np, tL, side


def _device_step(torch, ctx, vb, film_np, tL, side, mask=3):
    film = torch.from_numpy(film_np).cuda()
    keep = {k: torch.from_numpy(v).cuda() for k, v in side.items() if k != "error_sum"}
    sd = ctx.side(error=keep["error"], sample_weight=keep["sample_weight"], q_target=keep["q_target"],
                  error_sum=side["error_sum"])
    loss = torch.zeros(2, device="cuda")
    na.check(ctx.train_step(vb, mask, film=film, target_L=torch.from_numpy(tL).cuda(), side=sd, loss=loss),
             "train_step")
    torch.cuda.synchronize()
    return loss.cpu().numpy().astype(np.float64)


def _oracle_step(ora, ctx, vb, film_np, tL, side, p_stat, p_rrs):
    """FP64 joint step: (losses, gradients, parameters after Adam) of both nets."""
    cfg = ctx.cfg
    hv = ora.HostVertices(vb.to_numpy())
    stats = ora.net_forward(cfg, na.STATNET, hv, film_np, None, p_stat).T.copy()  # [6][n]
    l0, g0, _ = ora.train_grads(cfg, na.STATNET, hv, p_stat, na.LOSS_LL2, film=film_np, target_L=tL)
    rs = ora.RrsSide(None, side["error"], side["sample_weight"], side["q_target"], side["error_sum"])
    l1, g1, _ = ora.train_grads(cfg, na.RRSNET, hv, p_rrs, na.LOSS_RRS, film=film_np, stats=stats, side=rs)
    out = []
    for p, g, lr in ((p_stat, g0, cfg.lr_statnet), (p_rrs, g1, cfg.lr_rrsnet)):
        q = p.copy()
        ora.adam(q, np.zeros_like(q), np.zeros_like(q), g, lr, cfg.adam_beta1, cfg.adam_beta2, cfg.adam_eps,
                 cfg.adam_l2_reg, 1)
        out.append(q)
    return (l0, l1), (g0, g1), out


def _rel_max(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


def _rel_norm(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("n", [4096, 3001])
@pytest.mark.parametrize("paired", [1, 0])
def test_joint_step_matches_fp64(torch_cuda, ora, prec, n, paired):
    if prec == na.FP32 and paired == 0:
        pytest.skip("FP32 contexts have no paired table")
    _joint_step_vs_fp64(torch_cuda, ora, prec, n, paired)


@pytest.mark.parametrize("prec", PRECISIONS)
def test_joint_step_train_sigma_matches_fp64(torch_cuda, ora, prec):
    """train_sigma = 1 (network.h:158): StatNet's outputs 3-5 are sigmas (NRRS_LL2 scores
    sigma^2 + L^2 against L^2); RRSNet's training input takes activationSigma(max(L2, 0))
    (train.cu:196-200) and NRRS_RRS the second moments sigma^2 + max(L, 0)^2."""
    _joint_step_vs_fp64(torch_cuda, ora, prec, 4096, 1, train_sigma=1)


def _joint_step_vs_fp64(torch_cuda, ora, prec, n, paired, config=None, **kw):
    torch = torch_cuda
    config = config or na.default_config
    ctx = na.Context(config(precision=prec, n_pixels=n, max_queue=n, max_train=n, paired_tables=paired,
                            **LINEAR_ADAM, **kw))
    vb, film_np, tL, side = _batch(torch, ora, ctx, n, seed=n % 97)
    p0 = [ctx.params_get(net).astype(np.float64) for net in (na.STATNET, na.RRSNET)]
    loss = _device_step(torch, ctx, vb, film_np, tL, side)
    p1 = [ctx.params_get(net).astype(np.float64) for net in (na.STATNET, na.RRSNET)]
    (l0, l1), _, p_ref = _oracle_step(ora, ctx, vb, film_np, tL, side, *p0)
    for net, name in ((na.STATNET, "StatNet"), (na.RRSNET, "RRSNet")):
        g, _ = ctx.param_count(net)
        d_dev, d_ref = p1[net] - p0[net], p_ref[net] - p0[net]
        # (a net without a grid has no table part)
        e_mlp, e_grid = _rel_max(d_dev[g:], d_ref[g:]), _rel_norm(d_dev[:g], d_ref[:g]) if g else 0.0
        e_loss = abs(loss[net] - (l0, l1)[net]) / abs((l0, l1)[net])
        print(f"prec {prec} n {n} paired {paired} {name}: loss {loss[net]:.6g} vs {(l0, l1)[net]:.6g} "
              f"(rel {e_loss:.2e}); step MLP err {e_mlp:.2e}, grid norm err {e_grid:.2e}")
        assert e_loss < GRAD_TOL[prec]
        assert e_mlp < GRAD_TOL[prec]
        assert e_grid < GRID_NORM_TOL[prec]
        if prec == na.FP32 and g:
            assert _rel_max(d_dev[:g], d_ref[:g]) < GRAD_TOL[prec]
    for net in (na.STATNET, na.RRSNET):
        st = ctx.train_state(net)
        assert st["step"] == 1 and st["skipped"] == 0


@pytest.mark.parametrize("prec", PRECISIONS)
def test_joint_step_default_adam_signs(torch_cuda, ora, prec):
    """Default knobs (lr 1e-2, eps 1e-15): the first Adam step moves every parameter by
    -lr sign(g); entries whose FP64 gradient exceeds the rounding noise of the precision
    (relative to the largest gradient of their part, MLP or table) move in its direction."""
    torch = torch_cuda
    n = 4096
    ctx = na.Context(na.default_config(precision=prec, n_pixels=n, max_queue=n, max_train=n))
    vb, film_np, tL, side = _batch(torch, ora, ctx, n, seed=3)
    p0 = [ctx.params_get(net).astype(np.float64) for net in (na.STATNET, na.RRSNET)]
    _device_step(torch, ctx, vb, film_np, tL, side)
    p1 = [ctx.params_get(net).astype(np.float64) for net in (na.STATNET, na.RRSNET)]
    _, grads, _ = _oracle_step(ora, ctx, vb, film_np, tL, side, *p0)
    lr = (ctx.cfg.lr_statnet, ctx.cfg.lr_rrsnet)
    for net in (na.STATNET, na.RRSNET):
        g, _ = ctx.param_count(net)
        step = p1[net] - p0[net]
        gr = grads[net]
        for part in (slice(0, g), slice(g, None)):
            clear = np.abs(gr[part]) > 5 * GRAD_TOL[prec] * np.max(np.abs(gr[part]))
            moved = step[part][clear]
            want = -lr[net] * np.sign(gr[part][clear])
            print(f"prec {prec} net {net}: {int(clear.sum())} entries clear of the noise")
            assert clear.sum() > 0
            assert np.allclose(moved, want, rtol=1e-3, atol=1e-6)


def test_joint_step_2p20_bf16_properties(torch_cuda, ora):
    """Benchmark size: 2^20 samples in BF16, with the paired table (the joint record pass) and with
    per-net tables, from the same parameters on the same batch: five joint steps (StatNet's loss
    falls at every step; RRSNet's inputs, the StatNet statistics, move meanwhile, so its loss is
    not stationary), then five RRSNet steps on the frozen StatNet (its loss falls). Both paths
    agree bitwise throughout, and the loss-scale state is that of ten finite steps."""
    torch = torch_cuda
    n = 1 << 20
    ctxs = [na.Context(na.default_config(precision=na.BF16, n_pixels=n, max_queue=n, max_train=n, lr_rrsnet=1e-3,
                                         paired_tables=paired)) for paired in (1, 0)]
    p_init = [ctxs[0].params_get(net) for net in (na.STATNET, na.RRSNET)]
    for c in ctxs[1:]:
        for net in (na.STATNET, na.RRSNET):
            c.params_set(net, p_init[net])
    vb = na.synth_vertices(n, 4242, n)
    rng = np.random.default_rng(8)
    film_np = rng.lognormal(-1, 1, (3, n)).astype(np.float32)
    tL = ora.synth_targets(n, 9)
    side = dict(error=rng.uniform(0, 2, n).astype(np.float32),
                sample_weight=(1.0 / rng.integers(1, 4, n)).astype(np.float32),
                q_target=rng.uniform(0.05, 5.0, n).astype(np.float32), error_sum=0.7 * n)
    hist = []
    for c in ctxs:
        h = [_device_step(torch, c, vb, film_np, tL, side, mask=3) for _ in range(5)]
        h += [_device_step(torch, c, vb, film_np, tL, side, mask=2) for _ in range(5)]
        hist.append(np.array(h))
        print("losses per step:", np.array2string(hist[-1], precision=5))
    for h in hist:
        assert np.all(np.isfinite(h[:, 1])) and np.all(np.isfinite(h[:5, 0]))
        assert np.all(np.diff(h[:5, 0]) < 0)  # StatNet loss falls at every joint step
        assert h[-1, 1] < h[5, 1]  # RRSNet loss falls on the frozen StatNet
    assert np.array_equal(hist[0], hist[1])
    for net in (na.STATNET, na.RRSNET):
        assert np.array_equal(ctxs[0].params_get(net), ctxs[1].params_get(net)), net
        for c in ctxs:
            st = c.train_state(net)
            want = 5 if net == na.STATNET else 10
            assert st == dict(step=want, loss_scale=32768.0, skipped=0), st


def test_nonfinite_loss_reported(torch_cuda, ora):
    """A NaN sample weight makes the RRSNet batch loss NaN: nrrs_train_state reports it once
    with NRRS_ERR_NONFINITE (the reference stops on a NaN loss, integrator.cpp:2650-2656), and a
    finite step afterwards leaves no report."""
    torch = torch_cuda
    n = 4096
    for prec in PRECISIONS:
        ctx = na.Context(na.default_config(precision=prec, n_pixels=n, max_queue=n, max_train=n))
        vb, film_np, tL, side = _batch(torch, ora, ctx, n, seed=5)
        bad = dict(side)
        bad["q_target"] = np.abs(side["q_target"])
        bad["sample_weight"] = side["sample_weight"].copy()
        bad["sample_weight"][17] = np.nan
        loss = _device_step(torch, ctx, vb, film_np, tL, bad, mask=2)
        assert np.isnan(loss[1])
        with pytest.raises(na.NrrsError) as ei:
            ctx.train_state(na.RRSNET)
        assert ei.value.code == na.ERR_NONFINITE
        ctx.train_state(na.RRSNET)  # reported once
        ctx.train_state(na.STATNET)
        loss = _device_step(torch, ctx, vb, film_np, tL, side, mask=3)
        assert np.all(np.isfinite(loss))
        for net in (na.STATNET, na.RRSNET):
            ctx.train_state(net)
