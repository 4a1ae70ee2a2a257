"""The Ema optimizer (ema_decay; Ema{Adam}, 2net-simple.json:16-18, 171-173) on the device:

- the moving average the inference entry points read (nrrs_params_get_inference) against the
  FP64 oracle recursion run over the device's own trained parameters after each step, in
  FP32/FP16/BF16, for the benchmark and the shipped nets;
- inference reads the average: StatNet statistics against the oracle forward with the averaged
  parameters, away from those of the trained parameters; the frame kernel's q equals the chained
  stage's q with the average;
- a skipped step (FP16 loss-scale overflow) leaves the average; nrrs_params_set restarts it."""
import numpy as np
import pytest

import nrrs_amd as na
from test_gpu_joint import _batch, _device_step
from test_gpu_rrs import OUT_TOL

pytestmark = pytest.mark.gpu

PRECISIONS = [na.FP32, na.FP16, na.BF16]
DECAY = 0.9


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    assert torch.cuda.is_available()
    return torch


def _ctx(prec, shipped, n=4096, **kw):
    config = na.shipped_config if shipped else na.default_config
    return na.Context(config(precision=prec, n_pixels=n, max_queue=2 * n, max_train=n, ema_decay=DECAY, **kw))


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("shipped", [0, 1])
def test_ema_tracks_the_trained_parameters(torch_cuda, ora, prec, shipped):
    torch = torch_cuda
    n = 4096
    ctx = _ctx(prec, shipped, n)
    vb, film_np, tL, side = _batch(torch, ora, ctx, n, seed=2)
    nets = (na.STATNET, na.RRSNET)
    ema = [ctx.params_get(net).astype(np.float64) for net in nets]
    for net in nets:  # the average starts at the parameters (nrrs_params_set restarted it)
        assert np.array_equal(ctx.params_get_inference(net), ctx.params_get(net))
    for t in range(1, 5):
        loss = _device_step(torch, ctx, vb, film_np, tL, side)
        assert np.all(np.isfinite(loss))
        for net in nets:
            ora.ema_step(ema[net], ctx.params_get(net), DECAY, t)
    for net in nets:
        e_dev = ctx.params_get_inference(net).astype(np.float64)
        p_dev = ctx.params_get(net).astype(np.float64)
        err = np.max(np.abs(e_dev - ema[net])) / np.max(np.abs(ema[net]))
        gap = np.max(np.abs(p_dev - ema[net])) / np.max(np.abs(ema[net]))
        print(f"prec {prec} shipped {shipped} net {net}: average error {err:.2e} (trained parameters {gap:.2e} away)")
        assert err < 2e-6  # FP32 arithmetic of the average
        assert gap > 100 * err
        assert ctx.train_state(net)["step"] == 4
    # inference reads the average
    hv = ora.HostVertices(vb.to_numpy())
    stats = ctx.statnet_infer(vb).cpu().numpy()
    e_stat = ctx.params_get_inference(na.STATNET).astype(np.float64)
    ref = ora.net_forward(ctx.cfg, na.STATNET, hv, film_np, None, e_stat).T
    ref_trained = ora.net_forward(ctx.cfg, na.STATNET, hv, film_np, None,
                                  ctx.params_get(na.STATNET).astype(np.float64)).T
    err, off = _rel(stats, ref), _rel(stats, ref_trained)
    print(f"statistics: {err:.2e} from the averaged net, {off:.2e} from the trained one")
    assert err < OUT_TOL[prec]
    if prec == na.FP32:
        assert off > 100 * err


def test_ema_skipped_step_and_restart(torch_cuda, ora):
    """FP16 with an initial loss scale of 2^40: the first step overflows and is skipped, and the
    average stays at the parameters; nrrs_params_set then restarts the average at new values."""
    torch = torch_cuda
    n = 4096
    ctx = _ctx(na.FP16, 0, n, loss_scale_init=float(2 ** 40))
    vb, film_np, tL, side = _batch(torch, ora, ctx, n, seed=3)
    p0 = [ctx.params_get(net) for net in (na.STATNET, na.RRSNET)]
    _device_step(torch, ctx, vb, film_np, tL, side)
    for net in (na.STATNET, na.RRSNET):
        st = ctx.train_state(net)
        assert st["skipped"] == 1 and st["step"] == 0
        assert np.array_equal(ctx.params_get_inference(net), p0[net])
    q = p0[na.STATNET].copy()
    q[::7] *= -0.5
    ctx.params_set(na.STATNET, q)
    assert np.array_equal(ctx.params_get_inference(na.STATNET), q)


def test_ema_frame_reads_the_average(torch_cuda, ora):
    """BF16 benchmark nets after two training steps with Ema: the frame kernel's bounce-0 q equals
    the chained stage's q on the same vertices (both read the average), and differs from the
    q of a context without Ema holding the trained parameters."""
    torch = torch_cuda
    n = 1 << 14
    ctx = _ctx(na.BF16, 0, n)
    vb, film_np, tL, side = _batch(torch, ora, ctx, n, seed=4)
    film = torch.from_numpy(film_np).cuda()
    for _ in range(2):
        _device_step(torch, ctx, vb, film_np, tL, side)
    a, b = ctx.frame_buffers(n, 1, capacity=2 * n, trace=True, record_q=True)
    a.film_dev = film.data_ptr()
    a.seed = 99
    na.check(ctx.frame_run(a), "frame_run")
    torch.cuda.synchronize()
    assert int(b["status"].item()) == 0
    qf = b["q"][0, :n].cpu().numpy()
    queue = b["trace"][0, :n]
    fv = na.frame_vertices(99, 0, queue, n)

    def chained_q(c):
        q = torch.empty(n, device="cuda")
        surv = torch.empty(8 * n + 1, dtype=torch.int32, device="cuda")
        ns = torch.zeros(1, dtype=torch.int32, device="cuda")
        args = c.decide_args(fv, film, q, surv, ns, q_mode=na.Q_RRSNET_CHAINED, seed=99, stream_id=1000)
        na.check(c.rrs_infer_decide(args), "rrs_infer_decide")
        torch.cuda.synchronize()
        return q.cpu().numpy()

    assert np.array_equal(chained_q(ctx), qf)
    plain = na.Context(na.default_config(precision=na.BF16, n_pixels=n, max_queue=2 * n, max_train=n))
    for net in (na.STATNET, na.RRSNET):
        plain.params_set(net, ctx.params_get(net))
    assert not np.array_equal(chained_q(plain), qf)
