"""Training parity: hand-written backward kernels (encoding, MLP, both losses) and
Adam against the FP64 oracle, central finite differences of the device forward,
loss-scale handling, and that training reduces the loss."""
import numpy as np
import pytest

import nrrs_amd as na

pytestmark = pytest.mark.gpu

N = 4096
# Gradient error relative to the largest gradient entry, per operand precision
# (FP32 MFMA, FP16 / BF16 MFMA with FP32 accumulation and FP32 gradient buffers).
GRAD_TOL = {na.FP32: 2e-4, na.FP16: 1e-2, na.BF16: 5e-2}
# Norm-wise error of the table gradient in low precision (see the external-dy test).
GRID_NORM_TOL = {na.FP16: 5e-2, na.BF16: 1.5e-1}
# FP16/BF16 contexts accumulate the table gradient in FP16 (packed atomics): the tests
# scale the loss as training does, so that it is neither subnormal nor overflowing.
SCALE = {na.FP32: 1.0, na.FP16: 4096.0, na.BF16: 4096.0}
PRECISIONS = [na.FP32, na.FP16, na.BF16]


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    assert torch.cuda.is_available()
    return torch


def _setup(torch, prec, n=N, seed=1, **kw):
    cfg = na.default_config(precision=prec, n_pixels=n, max_queue=n, max_train=n, **kw)
    ctx = na.Context(cfg)
    rng = np.random.default_rng(seed)
    for net in (na.STATNET, na.RRSNET):
        p = ctx.params_get(net)
        g, _ = ctx.param_count(net)
        p[:g] = rng.uniform(-0.3, 0.3, g).astype(np.float32)
        ctx.params_set(net, p)
    vb = na.synth_vertices(n, 100 + seed, n)
    film_np = rng.lognormal(-1, 1, (3, n)).astype(np.float32)
    return ctx, vb, film_np, torch.from_numpy(film_np).cuda()


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


def _rrs_side_np(n, seed):
    rng = np.random.default_rng(seed)
    qt = rng.uniform(0.05, 5.0, n).astype(np.float32)
    qt[rng.uniform(size=n) < 0.1] = -1.0
    return dict(error=rng.uniform(0, 2, n).astype(np.float32),
                sample_weight=(1.0 / rng.integers(1, 4, n)).astype(np.float32), q_target=qt,
                error_sum=float(n) * 0.7)


@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("net", [na.STATNET, na.RRSNET])
def test_backward_external_matches_fp64(torch_cuda, ora, prec, net):
    torch = torch_cuda
    ctx, vb, film_np, film = _setup(torch, prec)
    n_out = 6 if net == na.STATNET else 1
    rng = np.random.default_rng(3)
    stats_np = np.abs(rng.normal(0.5, 0.5, (6, N))).astype(np.float32) if net == na.RRSNET else None
    stats = None if stats_np is None else torch.from_numpy(stats_np).cuda()
    hv = ora.HostVertices(vb.to_numpy())
    if prec == na.FP32:
        # zero-mean noise: the weight gradients are pure cancellation sums
        dy = rng.normal(0, 1e-3, (n_out, N)).astype(np.float32)
    else:
        # dL/dy of 1e-3 * y^2 / 2 plus noise: with zero-mean noise alone, ReLU-mask flips
        # of the low-precision forward (O(1) errors on ~1% of the terms) dominate the
        # cancellation sums; a coherent dL/dy measures the rounding of the backward itself
        y64 = ora.net_forward(ctx.cfg, net, hv, film_np, None if stats_np is None else stats_np.astype(np.float64),
                              ctx.params_get(net).astype(np.float64)).T
        dy = (1e-3 * y64 + rng.normal(0, 1e-5, (n_out, N))).astype(np.float32)
    ctx.grads_zero(net)
    na.check(ctx.backward(net, vb, na.LOSS_EXTERNAL, film=film, stats=stats, dy=torch.from_numpy(dy).cuda(),
                          loss_scale=SCALE[prec]), "backward")
    g_dev = ctx.grads_get(net) / SCALE[prec]
    p = ctx.params_get(net).astype(np.float64)
    _, g_ref, _ = ora.train_grads(ctx.cfg, net, hv, p, 0, film=film_np,
                                  stats=None if stats_np is None else stats_np.astype(np.float64),
                                  dy_ext=dy.astype(np.float64))
    g, _m = ctx.param_count(net)
    e_grid, e_mlp = _rel(g_dev[:g], g_ref[:g]), _rel(g_dev[g:], g_ref[g:])
    n_grid = float(np.linalg.norm(g_dev[:g] - g_ref[:g]) / np.linalg.norm(g_ref[:g]))
    print(f"prec {prec} net {net}: grid grad err max {e_grid:.2e} norm {n_grid:.2e}, mlp grad err {e_mlp:.2e}")
    assert e_mlp < GRAD_TOL[prec]
    if prec == na.FP32:
        assert e_grid < GRAD_TOL[prec]
    else:
        # A fine-level table entry is touched by one or two vertices, so its gradient is
        # one term w * dL/denc: a ReLU-mask flip of the low-precision forward in that
        # vertex moves it by O(1). Entry-wise maxima are therefore not a precision
        # measure there; the norm-wise error over the table is.
        assert n_grid < GRID_NORM_TOL[prec]


# Entry-wise table-gradient error of FP16/BF16 when no ReLU mask can flip, relative to the
# largest entry (ENTRY_TOL) and per entry among the entries above a twentieth of the largest
# (BIG_ENTRY_TOL). A term of an entry's gradient carries the rounding of 16-bit operands in each
# layer of the backward (2^-9 relative for BF16, 2^-12 for FP16, four layers deep) and of the
# FP16 record value.
ENTRY_TOL = {na.FP16: 4e-3, na.BF16: 3e-2}
BIG_ENTRY_TOL = {na.FP16: 8e-3, na.BF16: 6e-2}
MARGIN = 0.05  # smallest hidden pre-activation the weights below leave


def _flip_free_weights(rng, n_hidden, n_out, const):
    """MLP weights (the flat MLP part of a benchmark net's parameters) whose hidden
    pre-activations stay positive: layer 0 puts 2.0 on the constant input channel and at most
    0.01 on the others (80 inputs of magnitude below a few units move it by well under 1.0);
    hidden layers have positive weights in [0.005, 0.03] on positive inputs; the output layer is
    arbitrary."""
    k, h = na.ENC_WIDTH, 64
    w0 = rng.uniform(-0.01, 0.01, (h, k))
    w0[:, const] = 2.0
    mats = [w0] + [rng.uniform(0.005, 0.03, (h, h)) for _ in range(n_hidden - 1)]
    mats.append(rng.uniform(-0.3, 0.3, (n_out, h)))
    return mats


@pytest.mark.parametrize("prec", [na.FP16, na.BF16])
@pytest.mark.parametrize("net", [na.STATNET, na.RRSNET])
def test_table_gradient_entrywise_without_mask_flips(torch_cuda, ora, prec, net):
    """Entry-wise table gradient of the FP16/BF16 backward against the FP64 oracle. The weights
    keep every hidden pre-activation of the batch at least MARGIN above zero (checked on the
    FP64 forward, whose layer layout is checked against the oracle's outputs), so the
    low-precision forward cannot differ from the oracle's in any ReLU mask and each entry's
    error is the rounding of the backward alone, bounded entry by entry."""
    torch = torch_cuda
    ctx, vb, film_np, film = _setup(torch, prec)
    mcfg = ctx.cfg.statnet if net == na.STATNET else ctx.cfg.rrsnet
    n_out = int(mcfg.n_outputs)
    mats = _flip_free_weights(np.random.default_rng(21), int(mcfg.n_hidden_layers), n_out,
                              71 if net == na.STATNET else 75)  # the constant input channel
    p = ctx.params_get(net)
    g, m = ctx.param_count(net)
    flat = np.concatenate([w.ravel() for w in mats]).astype(np.float32)
    assert flat.size == m
    p[g:] = flat
    ctx.params_set(net, p)
    p64 = ctx.params_get(net).astype(np.float64)
    rng = np.random.default_rng(3)
    stats_np = np.abs(rng.normal(0.5, 0.5, (6, N))).astype(np.float32) if net == na.RRSNET else None
    stats64 = None if stats_np is None else stats_np.astype(np.float64)
    stats = None if stats_np is None else torch.from_numpy(stats_np).cuda()
    hv = ora.HostVertices(vb.to_numpy())
    # FP64 forward with the hidden pre-activations, from the oracle's network input
    x = ora.encode(ctx.cfg, net, hv, film_np, stats64, p64)
    w64 = [w.astype(np.float32).astype(np.float64) for w in mats]
    z_min = np.inf
    for w in w64[:-1]:
        z = x @ w.T
        z_min = min(z_min, float(z.min()))
        x = np.maximum(z, 0.0)
    y_np = x @ w64[-1].T
    y64 = ora.net_forward(ctx.cfg, net, hv, film_np, stats64, p64)
    assert np.max(np.abs(y_np - y64)) <= 1e-9 * max(1.0, np.max(np.abs(y64)))
    assert z_min >= MARGIN, z_min
    dy = (1e-3 * y64.T + rng.normal(0, 1e-5, (n_out, N))).astype(np.float32)
    ctx.grads_zero(net)
    na.check(ctx.backward(net, vb, na.LOSS_EXTERNAL, film=film, stats=stats, dy=torch.from_numpy(dy).cuda(),
                          loss_scale=SCALE[prec]), "backward")
    g_dev = ctx.grads_get(net)[:g] / SCALE[prec]
    _, g_ref, _ = ora.train_grads(ctx.cfg, net, hv, p64, 0, film=film_np, stats=stats64,
                                  dy_ext=dy.astype(np.float64))
    g_ref = g_ref[:g]
    top = np.max(np.abs(g_ref))
    err = np.abs(g_dev - g_ref)
    big = np.abs(g_ref) >= top / 20
    e_max = float(err.max() / top)
    e_big = float(np.max(err[big] / np.abs(g_ref[big])))
    print(f"prec {prec} net {net}: min pre-activation {z_min:.3f}; table gradient error {e_max:.2e} of the "
          f"largest entry, {e_big:.2e} per entry over {int(big.sum())} entries above 1/20 of it")
    assert e_max < ENTRY_TOL[prec]
    assert e_big < BIG_ENTRY_TOL[prec]


@pytest.mark.parametrize("case", ["table_2p20", "z_plane"])
def test_low_precision_table_gradient_fallbacks(torch_cuda, ora, case):
    """BF16 StatNet backward through the two table-gradient paths that the default
    configuration does not take, against the FP64 oracle:
    - table_2p20: a 2^20-entry hash table has 128 segments per level, more than the binned
      scatter handles, so the gradient goes through direct packed-FP16 atomics;
    - z_plane: every vertex on one z-plane packs the records of the dense levels 2 and 3 into
      one or two segments each, beyond their record slots, so the records that do not fit are
      added with packed-FP16 atomics (a few per table entry, hence FP16-accurate)."""
    torch = torch_cuda
    prec, net = na.BF16, na.STATNET
    kw = dict(grid=dict(log2_table_size=20)) if case == "table_2p20" else {}
    ctx, vb, film_np, film = _setup(torch, prec, **kw)
    if case == "z_plane":
        vb.pos[2, :N] = 0.37
    hv = ora.HostVertices(vb.to_numpy())
    if case == "z_plane":
        # precondition: some segment of a dense level receives more records than its slots.
        # Records: one per x-neighbour corner pair whose entries differ by 1, 3 or 7 (xor),
        # else one per corner; slots of a dense segment: min(8n, 3 x its share of 4.5n + 4096).
        lay = ora.grid_layout(ctx.cfg.grid)
        idx = ora.grid_indices(ctx.cfg, hv).astype(np.int64)  # [vertex][level][corner x | y<<1 | z<<2]
        over = False
        for l in range(ctx.cfg.grid.n_levels):
            if not lay["dense"][l]:
                continue
            size = int(lay["size"][l])
            e0, e1 = idx[:, l, 0::2].ravel(), idx[:, l, 1::2].ravel()
            single = ~np.isin(e0 ^ e1, [1, 3, 7])
            segs = np.concatenate([e0 >> 13, e1[single] >> 13])
            counts = np.bincount(segs, minlength=(size + 8191) // 8192)
            for b, c in enumerate(counts):
                ne = min(8192, size - b * 8192)
                cap = min(8 * N, 3 * (9 * N // 2 * ne // size) + 4096)
                over |= c > (cap + 63) // 64 * 64
        assert over
    rng = np.random.default_rng(3)
    y64 = ora.net_forward(ctx.cfg, net, hv, film_np, None, ctx.params_get(net).astype(np.float64)).T
    dy = (1e-3 * y64 + rng.normal(0, 1e-5, (6, N))).astype(np.float32)
    ctx.grads_zero(net)
    na.check(ctx.backward(net, vb, na.LOSS_EXTERNAL, film=film, dy=torch.from_numpy(dy).cuda(),
                          loss_scale=SCALE[prec]), "backward")
    g_dev = ctx.grads_get(net) / SCALE[prec]
    p = ctx.params_get(net).astype(np.float64)
    _, g_ref, _ = ora.train_grads(ctx.cfg, net, hv, p, 0, film=film_np, dy_ext=dy.astype(np.float64))
    g, _m = ctx.param_count(net)
    n_grid = float(np.linalg.norm(g_dev[:g] - g_ref[:g]) / np.linalg.norm(g_ref[:g]))
    e_mlp = _rel(g_dev[g:], g_ref[g:])
    print(f"{case}: grid grad norm err {n_grid:.2e}, mlp grad err {e_mlp:.2e}")
    assert e_mlp < GRAD_TOL[prec]
    assert n_grid < GRID_NORM_TOL[prec]
    # the optimizer step after these gradients is taken; after a backward whose scale overflows
    # FP16 (records and direct atomics alike) it is skipped
    ctx.adam_step(net)
    assert ctx.train_state(net)["step"] == 1 and ctx.train_state(net)["skipped"] == 0
    ctx.grads_zero(net)
    na.check(ctx.backward(net, vb, na.LOSS_EXTERNAL, film=film, dy=torch.from_numpy(dy).cuda(), loss_scale=2.0**40),
             "backward")
    ctx.adam_step(net)
    st = ctx.train_state(net)
    assert st["step"] == 1 and st["skipped"] == 1


@pytest.mark.parametrize("n", [777, 300])
def test_low_precision_backward_ragged_batch(torch_cuda, ora, n):
    """BF16 StatNet backward at batch sizes that are no multiple of the 256-vertex runs of the
    record counts or of the 512-vertex workgroups of the record kernel (partial runs, a partial
    last workgroup, scan tails), against the FP64 oracle."""
    torch = torch_cuda
    prec, net = na.BF16, na.STATNET
    ctx, vb, film_np, film = _setup(torch, prec, n=n, seed=7)
    hv = ora.HostVertices(vb.to_numpy())
    rng = np.random.default_rng(11)
    y64 = ora.net_forward(ctx.cfg, net, hv, film_np, None, ctx.params_get(net).astype(np.float64)).T
    dy = (1e-3 * y64 + rng.normal(0, 1e-5, (6, n))).astype(np.float32)
    ctx.grads_zero(net)
    na.check(ctx.backward(net, vb, na.LOSS_EXTERNAL, film=film, dy=torch.from_numpy(dy).cuda(),
                          loss_scale=SCALE[prec]), "backward")
    g_dev = ctx.grads_get(net) / SCALE[prec]
    p = ctx.params_get(net).astype(np.float64)
    _, g_ref, _ = ora.train_grads(ctx.cfg, net, hv, p, 0, film=film_np, dy_ext=dy.astype(np.float64))
    g, _m = ctx.param_count(net)
    n_grid = float(np.linalg.norm(g_dev[:g] - g_ref[:g]) / np.linalg.norm(g_ref[:g]))
    e_mlp = _rel(g_dev[g:], g_ref[g:])
    print(f"n={n}: grid grad norm err {n_grid:.2e}, mlp grad err {e_mlp:.2e}")
    assert e_mlp < GRAD_TOL[prec]
    assert n_grid < GRID_NORM_TOL[prec]


@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("train_sigma", [0, 1])
def test_ll2_and_rrs_loss_backward(torch_cuda, ora, prec, train_sigma):
    """Both losses through nrrs_backward against the FP64 oracle; with train_sigma the StatNet
    outputs 3-5 are sigmas (LL2 scores sigma^2 + L^2 against L^2) and RRSNet's training input
    and loss read them as such (statistics of both signs)."""
    torch = torch_cuda
    ctx, vb, film_np, film = _setup(torch, prec, seed=2, train_sigma=train_sigma)
    hv = ora.HostVertices(vb.to_numpy())
    tL = ora.synth_targets(N, 4)
    loss = torch.zeros(1, device="cuda")
    ctx.grads_zero(na.STATNET)
    na.check(ctx.backward(na.STATNET, vb, na.LOSS_LL2, target_L=torch.from_numpy(tL).cuda(), loss=loss,
                          loss_scale=SCALE[prec]), "ll2")
    p = ctx.params_get(na.STATNET).astype(np.float64)
    l_ref, g_ref, _ = ora.train_grads(ctx.cfg, na.STATNET, hv, p, 1, film=film_np, target_L=tL)
    g_dev = ctx.grads_get(na.STATNET) / SCALE[prec]
    print(f"prec {prec}: LL2 loss {loss.item():.6g} vs {l_ref:.6g}, grad err {_rel(g_dev, g_ref):.2e}")
    assert abs(loss.item() - l_ref) / l_ref < GRAD_TOL[prec]
    assert _rel(g_dev, g_ref) < GRAD_TOL[prec]

    sd = _rrs_side_np(N, 5)
    stats_np = np.abs(np.random.default_rng(6).normal(0.5, 0.5, (6, N))).astype(np.float32)
    if train_sigma:
        stats_np[3:] = np.random.default_rng(7).normal(0.0, 1.5, (3, N))
    keep = {k: torch.from_numpy(v).cuda() for k, v in sd.items() if k != "error_sum"}
    side = ctx.side(error=keep["error"], sample_weight=keep["sample_weight"], q_target=keep["q_target"],
                    error_sum=sd["error_sum"])
    ctx.grads_zero(na.RRSNET)
    na.check(ctx.backward(na.RRSNET, vb, na.LOSS_RRS, film=film, stats=torch.from_numpy(stats_np).cuda(),
                          side=side, loss=loss, loss_scale=SCALE[prec]), "rrs")
    p = ctx.params_get(na.RRSNET).astype(np.float64)
    ref_side = ora.RrsSide(None, sd["error"], sd["sample_weight"], sd["q_target"], sd["error_sum"])
    l_ref, g_ref, _ = ora.train_grads(ctx.cfg, na.RRSNET, hv, p, 2, film=film_np, stats=stats_np.astype(np.float64),
                                      side=ref_side)
    g_dev = ctx.grads_get(na.RRSNET) / SCALE[prec]
    print(f"prec {prec}: RRS loss {loss.item():.6g} vs {l_ref:.6g}, grad err {_rel(g_dev, g_ref):.2e}")
    assert abs(loss.item() - l_ref) / l_ref < GRAD_TOL[prec]
    assert _rel(g_dev, g_ref) < GRAD_TOL[prec]


def test_finite_differences_of_device_forward(torch_cuda, ora):
    """Central differences (eps 1e-3) of the FP32 device forward through the
    frozen-denominator LL2 loss, on grid entries and MLP weights."""
    torch = torch_cuda
    ctx, vb, film_np, film = _setup(torch, na.FP32, seed=3)
    tL = ora.synth_targets(N, 7)
    hv = ora.HostVertices(vb.to_numpy())
    y0 = ctx.statnet_infer(vb).cpu().numpy().astype(np.float64)
    ctx.grads_zero(na.STATNET)
    na.check(ctx.backward(na.STATNET, vb, na.LOSS_LL2, target_L=torch.from_numpy(tL).cuda()), "ll2")
    grad = ctx.grads_get(na.STATNET).astype(np.float64)
    p0 = ctx.params_get(na.STATNET)
    g, m = ctx.param_count(na.STATNET)
    rng = np.random.default_rng(0)
    touched = np.nonzero(np.abs(grad[:g]) > 1e-3 * np.abs(grad[:g]).max())[0]
    idx = np.concatenate([rng.choice(touched, 16, replace=False), g + rng.choice(m, 16, replace=False)])

    def loss_at(p):
        ctx.params_set(na.STATNET, p)
        y = ctx.statnet_infer(vb).cpu().numpy().astype(np.float64)
        return ora.loss(ctx.cfg, 1, hv, y, target_L=tL, y_den=y0)[0]

    errs = []
    for k in idx:
        p1 = p0.copy()
        p1[k] += 1e-3
        p2 = p0.copy()
        p2[k] -= 1e-3
        fd = (loss_at(p1) - loss_at(p2)) / 2e-3
        errs.append(abs(fd - grad[k]) / np.abs(grad[idx]).max())
    ctx.params_set(na.STATNET, p0)
    errs = np.array(errs)
    print(f"FD errors: median {np.median(errs):.2e} max {errs.max():.2e}")
    # FP32 forward noise / (2 eps) and ReLU kinks: a few may be off, the bulk is tight
    assert np.median(errs) < 1e-3 and np.mean(errs < 1e-2) >= 0.9


@pytest.mark.parametrize("prec", [na.FP32, na.BF16])
def test_adam_matches_oracle(torch_cuda, ora, prec):
    torch = torch_cuda
    ctx, vb, film_np, film = _setup(torch, prec, seed=4, adam_l2_reg=1e-4)
    tL = torch.from_numpy(ora.synth_targets(N, 8)).cuda()
    p = ctx.params_get(na.STATNET).astype(np.float64)
    mo, vo = np.zeros_like(p), np.zeros_like(p)
    for step in range(1, 4):
        ctx.grads_zero(na.STATNET)
        na.check(ctx.backward(na.STATNET, vb, na.LOSS_LL2, target_L=tL,
                              loss_scale=ctx.train_state(na.STATNET)["loss_scale"]), "ll2")
        gd = ctx.grads_get(na.STATNET).astype(np.float64) / ctx.train_state(na.STATNET)["loss_scale"]
        ctx.adam_step(na.STATNET)
        ora.adam(p, mo, vo, gd, ctx.cfg.lr_statnet, ctx.cfg.adam_beta1, ctx.cfg.adam_beta2, ctx.cfg.adam_eps,
                 ctx.cfg.adam_l2_reg, step)
        pd = ctx.params_get(na.STATNET).astype(np.float64)
        err = np.max(np.abs(pd - p))
        print(f"prec {prec} step {step}: max |p_dev - p_ref| = {err:.2e}")
        assert err < 1e-5 * max(1.0, step)
        p = pd  # follow the device trajectory (FP32 rounding)
    assert ctx.train_state(na.STATNET)["step"] == 3


def test_loss_scale_overflow_skips_step(torch_cuda, ora):
    """FP16: an initial scale of 2^40 overflows the FP16 dy operands; the step is
    skipped (parameters unchanged, gradients cleared) and the scale halves."""
    torch = torch_cuda
    ctx, vb, _, _ = _setup(torch, na.FP16, seed=5, loss_scale_init=float(2 ** 40))
    tL = torch.from_numpy(ora.synth_targets(N, 9)).cuda()
    p0 = ctx.params_get(na.STATNET)
    na.check(ctx.train_step(vb, 1, target_L=tL), "train_step")
    st = ctx.train_state(na.STATNET)
    assert st["skipped"] == 1 and st["step"] == 0 and st["loss_scale"] == float(2 ** 39)
    assert np.array_equal(ctx.params_get(na.STATNET), p0)
    assert not ctx.grads_get(na.STATNET).any()


@pytest.mark.parametrize("prec", PRECISIONS)
def test_training_reduces_losses(torch_cuda, ora, prec):
    """StatNet alone, then RRSNet alone on the frozen StatNet (whose statistics are its
    inputs, so its loss is stationary only then), then both jointly."""
    torch = torch_cuda
    n = 1 << 14
    ctx, vb, film_np, film = _setup(torch, prec, n=n, seed=6, lr_rrsnet=1e-3)
    tL = torch.from_numpy(ora.synth_targets(n, 10)).cuda()
    sd = _rrs_side_np(n, 11)
    keep = {k: torch.from_numpy(v).cuda() for k, v in sd.items() if k != "error_sum"}
    side = ctx.side(error=keep["error"], sample_weight=keep["sample_weight"], q_target=keep["q_target"],
                    error_sum=sd["error_sum"])
    loss = torch.zeros(2, device="cuda")
    hist = {}
    for mask, steps in ((1, 40), (2, 40), (3, 10)):
        h = []
        for _ in range(steps):
            na.check(ctx.train_step(vb, mask, film=film, target_L=tL, side=side, loss=loss), "train_step")
            h.append(loss.cpu().numpy().copy())
        hist[mask] = np.array(h)
    print(f"prec {prec}: StatNet loss {hist[1][0, 0]:.4g} -> {hist[1][-1, 0]:.4g}, RRSNet {hist[2][0, 1]:.4g} -> "
          f"{hist[2][-1, 1]:.4g}")
    assert all(np.all(np.isfinite(h)) for h in hist.values())
    assert hist[1][-5:, 0].mean() < 0.7 * hist[1][:3, 0].mean()
    assert hist[2][-5:, 1].mean() < 0.9 * hist[2][:3, 1].mean()
