"""The hand-written backward kernels behind a torch.autograd.Function (tests/nrrs_autograd.py):
torch.autograd.gradcheck by central differences of the device forward (FP32), torch-composed
losses against the FP64 oracle gradient (FP32, FP16, BF16), and central differences on 256
parameters of the FP32 training step at 2^18 samples (NRRS_LL2, the configuration of the
training benchmark), plus 64 parameters in FP16 and BF16 against central differences of the
FP64 loss."""
import numpy as np
import pytest

import nrrs_amd as na
import nrrs_autograd as nag

pytestmark = pytest.mark.gpu

GRAD_TOL = {na.FP32: 2e-4, na.FP16: 1e-2, na.BF16: 5e-2}
SCALE = {na.FP32: 1.0, na.FP16: 4096.0, na.BF16: 4096.0}


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    assert torch.cuda.is_available()
    return torch


def _setup(torch, prec, n, seed):
    ctx = na.Context(na.default_config(precision=prec, n_pixels=n, max_queue=n, max_train=n))
    rng = np.random.default_rng(seed)
    base = []
    for net in (na.STATNET, na.RRSNET):
        p = ctx.params_get(net)
        g, _ = ctx.param_count(net)
        p[:g] = rng.uniform(-0.3, 0.3, g).astype(np.float32)
        ctx.params_set(net, p)
        base.append(p)
    vb = na.synth_vertices(n, 900 + seed, n)
    film_np = rng.lognormal(-1, 1, (3, n)).astype(np.float32)
    stats_np = np.abs(rng.normal(0.5, 0.5, (6, n))).astype(np.float32)
    return ctx, base, vb, film_np, torch.from_numpy(film_np).cuda(), stats_np, torch.from_numpy(stats_np).cuda()


def _touched(ctx, net, vb, film, stats, n_out, n):
    """Grid entries the batch reaches (nonzero gradient of a generic dL/dy)."""
    import torch

    ctx.grads_zero(net)
    dy = torch.ones(n_out, n, device="cuda")
    na.check(ctx.backward(net, vb, na.LOSS_EXTERNAL, film=film, stats=stats, dy=dy), "backward")
    g, _ = ctx.param_count(net)
    gr = ctx.grads_get(net)
    ctx.grads_zero(net)
    return np.nonzero(np.abs(gr[:g]) > 1e-6 * np.abs(gr[:g]).max())[0]


@pytest.mark.parametrize("net", [na.STATNET, na.RRSNET])
def test_gradcheck_fp32(torch_cuda, net):
    """torch.autograd.gradcheck (random-projection form) on 32 parameters (16 touched grid
    entries, 16 MLP weights) of the FP32 net: v^T J u by central differences of the device
    forward against the backward kernels."""
    torch = torch_cuda
    n = 2048
    ctx, base, vb, _, film, _, stats = _setup(torch, na.FP32, n, 1)
    n_out = 6 if net == na.STATNET else 1
    rng = np.random.default_rng(2)
    g, m = ctx.param_count(net)
    sel = np.concatenate([rng.choice(_touched(ctx, net, vb, film, stats, n_out, n), 16, replace=False),
                          g + rng.choice(m, 16, replace=False)])
    x = torch.tensor(base[net][sel].astype(np.float64), device="cuda", requires_grad=True)

    def f(v):
        return nag.net_outputs(v, ctx, net, vb, film, stats if net == na.RRSNET else None, base[net], sel)

    assert torch.autograd.gradcheck(f, (x,), eps=1e-3, atol=2e-3, rtol=2e-2, fast_mode=True)


@pytest.mark.parametrize("prec", [na.FP32, na.FP16, na.BF16])
@pytest.mark.parametrize("net", [na.STATNET, na.RRSNET])
def test_autograd_loss_matches_fp64(torch_cuda, ora, prec, net):
    """A loss written in torch (weighted sum of squares of the outputs) differentiated through
    the autograd function, against the FP64 oracle gradient of the same loss."""
    torch = torch_cuda
    n = 4096
    ctx, base, vb, film_np, film, stats_np, stats = _setup(torch, prec, n, 3)
    n_out = 6 if net == na.STATNET else 1
    g, m = ctx.param_count(net)
    sel = np.arange(g + m)  # every parameter
    x = torch.tensor(base[net].astype(np.float64), device="cuda", requires_grad=True)
    w = torch.from_numpy(np.random.default_rng(4).uniform(0.5, 1.5, (n_out, n))).cuda()
    y = nag.net_outputs(x, ctx, net, vb, film, stats if net == na.RRSNET else None, base[net], sel, SCALE[prec])
    loss = 1e-3 * (w * y * y).sum() / 2
    loss.backward()
    gd = x.grad.cpu().numpy()
    hv = ora.HostVertices(vb.to_numpy())
    st64 = stats_np.astype(np.float64) if net == na.RRSNET else None
    y64 = ora.net_forward(ctx.cfg, net, hv, film_np, st64, base[net].astype(np.float64)).T
    dy = 1e-3 * w.cpu().numpy() * y64  # dL/dy of the same loss at the FP64 outputs
    _, g_ref, _ = ora.train_grads(ctx.cfg, net, hv, base[net].astype(np.float64), na.LOSS_EXTERNAL, film=film_np,
                                  stats=st64, dy_ext=dy)
    e_mlp = np.max(np.abs(gd[g:] - g_ref[g:])) / np.max(np.abs(g_ref[g:]))
    e_grid = np.linalg.norm(gd[:g] - g_ref[:g]) / np.linalg.norm(g_ref[:g])
    print(f"prec {prec} net {net}: MLP grad err {e_mlp:.2e}, grid norm err {e_grid:.2e}")
    assert e_mlp < GRAD_TOL[prec]
    assert e_grid < (GRAD_TOL[prec] if prec == na.FP32 else 3 * GRAD_TOL[prec])


def test_central_differences_fp32_2p18_256_params(torch_cuda, ora):
    """The FP32 StatNet training step at 2^18 samples (NRRS_LL2 on log-normal radiance with a 5%
    firefly tail): the backward's gradient on 256 random parameters (128 touched grid entries,
    128 MLP weights) against central differences (eps 1e-3) of the device forward through the
    loss with its denominator frozen (the relative-L2 convention of the gradient), and the whole
    gradient against the FP64 oracle."""
    torch = torch_cuda
    n = 1 << 18
    ctx, base, vb, film_np, film, _, _ = _setup(torch, na.FP32, n, 5)
    tL = ora.synth_targets(n, 6)
    hv = ora.HostVertices(vb.to_numpy())
    ctx.grads_zero(na.STATNET)
    na.check(ctx.backward(na.STATNET, vb, na.LOSS_LL2, target_L=torch.from_numpy(tL).cuda()), "ll2")
    grad = ctx.grads_get(na.STATNET).astype(np.float64)
    p0 = base[na.STATNET]
    y0 = ctx.statnet_infer(vb).cpu().numpy().astype(np.float64)
    g, m = ctx.param_count(na.STATNET)
    rng = np.random.default_rng(7)
    touched = np.nonzero(np.abs(grad[:g]) > 1e-3 * np.abs(grad[:g]).max())[0]
    idx = np.concatenate([rng.choice(touched, 128, replace=False), g + rng.choice(m, 128, replace=False)])

    def loss_at(p):
        ctx.params_set(na.STATNET, p)
        y = ctx.statnet_infer(vb).cpu().numpy().astype(np.float64)
        return ora.loss(ctx.cfg, na.LOSS_LL2, hv, y, target_L=tL, y_den=y0)[0]

    scale = np.abs(grad[idx]).max()
    errs = []
    for k in idx:
        p1, p2 = p0.copy(), p0.copy()
        p1[k] += 1e-3
        p2[k] -= 1e-3
        errs.append(abs((loss_at(p1) - loss_at(p2)) / 2e-3 - grad[k]) / scale)
    ctx.params_set(na.STATNET, p0)
    errs = np.array(errs)
    print(f"2^18 FD on 256 parameters: median {np.median(errs):.2e}, 90th pct {np.percentile(errs, 90):.2e}, "
          f"max {errs.max():.2e}")
    # FP32 forward noise over 2 eps and ReLU kinks straddled by a 1e-3 step: the bulk is tight
    assert np.median(errs) < 1e-3 and np.mean(errs < 1e-2) >= 0.9
    _, g_ref, _ = ora.train_grads(ctx.cfg, na.STATNET, hv, p0.astype(np.float64), na.LOSS_LL2, film=film_np,
                                  target_L=tL)
    e = np.max(np.abs(grad - g_ref)) / np.max(np.abs(g_ref))
    print(f"2^18 gradient against FP64: {e:.2e}")
    assert e < GRAD_TOL[na.FP32]


@pytest.mark.parametrize("prec", [na.FP16, na.BF16])
def test_central_differences_low_precision(torch_cuda, ora, prec):
    """The FP16/BF16 backward (loss-scaled, FP16 table gradient) on 64 parameters against central
    differences (eps 1e-3) of the FP64 loss of the same parameters: the low-precision forward
    itself is too coarse for a 1e-3 step, so the differences are taken on the function the
    backward differentiates."""
    torch = torch_cuda
    n = 4096
    ctx, base, vb, film_np, film, _, _ = _setup(torch, prec, n, 8)
    tL = ora.synth_targets(n, 9)
    hv = ora.HostVertices(vb.to_numpy())
    ctx.grads_zero(na.STATNET)
    na.check(ctx.backward(na.STATNET, vb, na.LOSS_LL2, target_L=torch.from_numpy(tL).cuda(), loss_scale=SCALE[prec]),
             "ll2")
    grad = ctx.grads_get(na.STATNET).astype(np.float64) / SCALE[prec]
    p0 = base[na.STATNET].astype(np.float64)
    y0 = ora.net_forward(ctx.cfg, na.STATNET, hv, film_np, None, p0).T.copy()
    g, m = ctx.param_count(na.STATNET)
    rng = np.random.default_rng(10)
    touched = np.nonzero(np.abs(grad[:g]) > 1e-2 * np.abs(grad[:g]).max())[0]
    idx = np.concatenate([rng.choice(touched, 32, replace=False), g + rng.choice(m, 32, replace=False)])

    def loss_at(p):
        y = ora.net_forward(ctx.cfg, na.STATNET, hv, film_np, None, p).T
        return ora.loss(ctx.cfg, na.LOSS_LL2, hv, y, target_L=tL, y_den=y0)[0]

    fd = []
    for k in idx:
        p1, p2 = p0.copy(), p0.copy()
        p1[k] += 1e-3
        p2[k] -= 1e-3
        fd.append((loss_at(p1) - loss_at(p2)) / 2e-3)
    fd = np.array(fd)
    mlp = idx >= g
    e_mlp = np.max(np.abs(grad[idx][mlp] - fd[mlp])) / np.max(np.abs(fd[mlp]))
    e_grid = np.linalg.norm(grad[idx][~mlp] - fd[~mlp]) / np.linalg.norm(fd[~mlp])
    print(f"prec {prec}: FD on 64 parameters, MLP err {e_mlp:.2e}, grid norm err {e_grid:.2e}")
    assert e_mlp < 2 * GRAD_TOL[prec]
    assert e_grid < 3 * GRAD_TOL[prec]
