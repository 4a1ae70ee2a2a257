"""AllSelf StatNets on the device (ll2_two_head 3; twn_AllSelf.cpp: two nets, each with its own
encoding, optimizer and loss) against the FP64 oracle, which evaluates each as a network of its
own:

- the statistics (nrrs_statnet_infer: L from NRRS_STATNET, L^2 from NRRS_STATNET2) and the
  chained RRS stage on them;
- the joint training step: each StatNet on its head's half of NRRS_LL2 with Adam at
  lr_statnet / 5 (twn_AllSelf.cpp:9-16), RRSNet on the statistics of both, in the linear-Adam
  regime of the joint-step test (RRSNet against the oracle fed the device's statistics);
- the output rows of the other head: zero after nrrs_params_set whatever the caller passes, and
  still zero after training steps with the default knobs;
- the frame kernel refuses AllSelf."""
import numpy as np
import pytest

import nrrs_amd as na
from test_gpu_joint import GRAD_TOL, GRID_NORM_TOL, LINEAR_ADAM, _batch, _device_step, _rel_max, _rel_norm

pytestmark = pytest.mark.gpu

PRECISIONS = [na.FP32, na.FP16, na.BF16]
OUT_TOL = {na.FP32: 1e-5, na.FP16: 5e-3, na.BF16: 4e-2}  # of the output range (test_gpu_rrs)
Q_TOL = {na.FP32: 1e-4, na.FP16: 5e-3, na.BF16: 3e-2}
STATNETS = (na.STATNET, na.STATNET2)


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    assert torch.cuda.is_available()
    return torch


def _cfg(prec, n, **kw):
    base = dict(precision=prec, n_pixels=n, max_queue=n, max_train=n, ll2_two_head=3,
                statnet=dict(n_hidden_layers=2, n_outputs=6))
    base.update(kw)
    return na.default_config(**base)


def _other_rows(ctx, net):
    """Indices (into a StatNet parameter vector) of the output weights of the other head."""
    g, m = ctx.param_count(net)
    out0 = g + m - 6 * 64
    head = 0 if net == na.STATNET else 1
    rows = [r for r in range(6) if r // 3 != head]
    return np.concatenate([out0 + 64 * r + np.arange(64) for r in rows])


def _stats64(ora, cfg, hv, film, p_a, p_b):
    ya = ora.net_forward(cfg, na.STATNET, hv, film, None, p_a).T
    yb = ora.net_forward(cfg, na.STATNET2, hv, film, None, p_b).T
    return np.vstack([ya[:3], yb[3:]])


def _random_grid(ctx, net, rng):
    p = ctx.params_get(net)
    g, _ = ctx.param_count(net)
    p[:g] = rng.uniform(-0.3, 0.3, g).astype(np.float32)
    ctx.params_set(net, p)


@pytest.mark.parametrize("prec", PRECISIONS)
def test_all_self_inference_matches_fp64(torch_cuda, ora, prec):
    torch = torch_cuda
    n = 4096
    ctx = na.Context(_cfg(prec, n))
    rng = np.random.default_rng(5)
    for net in STATNETS:
        _random_grid(ctx, net, rng)
    vb = na.synth_vertices(n, 33, n)
    film_np = rng.lognormal(-1, 1, (3, n)).astype(np.float32)
    film = torch.from_numpy(film_np).cuda()
    hv = ora.HostVertices(vb.to_numpy())
    ps = [ctx.params_get(net).astype(np.float64) for net in STATNETS]
    pr = ctx.params_get(na.RRSNET).astype(np.float64)
    st = ctx.statnet_infer(vb).cpu().numpy()
    ref = _stats64(ora, ctx.cfg, hv, film_np, *ps)
    e_st = np.max(np.abs(st - ref)) / np.max(np.abs(ref))
    q = torch.empty(n, dtype=torch.float32, device="cuda")
    surv = torch.empty(8 * n, dtype=torch.int32, device="cuda")
    ns = torch.zeros(1, dtype=torch.int32, device="cuda")
    args = ctx.decide_args(vb, film, q, surv, ns, q_mode=na.Q_RRSNET_CHAINED, seed=3)
    na.check(ctx.rrs_infer_decide(args), "rrs_infer_decide")
    raw = ora.net_forward(ctx.cfg, na.RRSNET, hv, film_np, ref, pr)[:, 0]
    q64 = ora.q_f64(ctx.cfg, na.Q_RRSNET_CHAINED, hv, raw, film_np)
    e_q = float(np.max(np.abs(q.cpu().numpy() - q64) / q64))
    print(f"AllSelf prec {prec}: statnet err {e_st:.2e}, chained q err {e_q:.2e}")
    assert e_st < OUT_TOL[prec]
    assert e_q < Q_TOL[prec]


@pytest.mark.parametrize("prec", PRECISIONS)
def test_all_self_joint_step_matches_fp64(torch_cuda, ora, prec):
    torch = torch_cuda
    n = 4096
    ctx = na.Context(_cfg(prec, n, **LINEAR_ADAM))
    vb, film_np, tL, side = _batch(torch, ora, ctx, n, seed=23)
    _random_grid(ctx, na.STATNET2, np.random.default_rng(24))
    nets = (na.STATNET, na.STATNET2, na.RRSNET)
    st_dev = ctx.statnet_infer(vb).cpu().numpy().astype(np.float64)  # RRSNet's input and loss statistics
    p0 = {net: ctx.params_get(net).astype(np.float64) for net in nets}
    loss = _device_step(torch, ctx, vb, film_np, tL, side)
    p1 = {net: ctx.params_get(net).astype(np.float64) for net in nets}
    cfg = ctx.cfg
    hv = ora.HostVertices(vb.to_numpy())
    rs = ora.RrsSide(None, side["error"], side["sample_weight"], side["q_target"], side["error_sum"])
    ref, losses = {}, {}
    for net in nets:
        if net == na.RRSNET:
            l, g, _ = ora.train_grads(cfg, net, hv, p0[net], na.LOSS_RRS, film=film_np, stats=st_dev, side=rs)
            lr = cfg.lr_rrsnet
        else:
            l, g, _ = ora.train_grads(cfg, net, hv, p0[net], na.LOSS_LL2, film=film_np, target_L=tL)
            lr = cfg.lr_statnet / 5.0
        q = p0[net].copy()
        ora.adam(q, np.zeros_like(q), np.zeros_like(q), g, lr, cfg.adam_beta1, cfg.adam_beta2, cfg.adam_eps,
                 cfg.adam_l2_reg, 1)
        ref[net], losses[net] = q, l
    l_stat = losses[na.STATNET] + losses[na.STATNET2]  # loss_dev[0]: (loss1 + loss2) / 2 of NRRS_LL2
    for net, name, lv, lr_ in ((na.STATNET, "StatNet L", l_stat, loss[0]), (na.STATNET2, "StatNet L^2", l_stat, loss[0]),
                               (na.RRSNET, "RRSNet", losses[na.RRSNET], loss[1])):
        g, _ = ctx.param_count(net)
        d_dev, d_ref = p1[net] - p0[net], ref[net] - p0[net]
        e_mlp, e_grid = _rel_max(d_dev[g:], d_ref[g:]), _rel_norm(d_dev[:g], d_ref[:g])
        e_loss = abs(lr_ - lv) / abs(lv)
        print(f"AllSelf prec {prec} {name}: loss {lr_:.6g} vs {lv:.6g} (rel {e_loss:.2e}); step MLP err "
              f"{e_mlp:.2e}, grid norm err {e_grid:.2e}")
        assert e_loss < GRAD_TOL[prec]
        assert e_mlp < GRAD_TOL[prec]
        assert e_grid < GRID_NORM_TOL[prec]
    for net in STATNETS:
        assert np.all(p1[net][_other_rows(ctx, net)] == 0.0)
        st = ctx.train_state(net)
        assert st["step"] == 1 and st["skipped"] == 0


def test_all_self_other_rows_stay_zero(torch_cuda, ora):
    torch = torch_cuda
    n = 4096
    ctx = na.Context(_cfg(na.BF16, n))
    for net in STATNETS:
        rows = _other_rows(ctx, net)
        p = ctx.params_get(net)
        assert np.all(p[rows] == 0.0)
        p[rows] = 0.5  # nrrs_params_set puts them back to zero
        ctx.params_set(net, p)
        assert np.all(ctx.params_get(net)[rows] == 0.0)
    hist = []
    for k in range(3):
        vb, film_np, tL, side = _batch(torch, ora, ctx, n, seed=60 + k)
        hist.append(_device_step(torch, ctx, vb, film_np, tL, side))
    assert np.all(np.isfinite(hist))
    for net in STATNETS:
        assert np.all(ctx.params_get(net)[_other_rows(ctx, net)] == 0.0)
        assert ctx.train_state(net)["step"] == 3


def test_all_self_frame_refused(torch_cuda):
    n = 4096
    ctx = na.Context(_cfg(na.BF16, n, max_queue=2 * n))
    a, _ = ctx.frame_buffers(n, 2, capacity=2 * n)
    film = torch_cuda.ones(3, n, device="cuda")
    a.film_dev = film.data_ptr()
    assert ctx.frame_run(a) == na.ERR_UNSUPPORTED
