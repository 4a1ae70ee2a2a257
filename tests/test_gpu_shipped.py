"""The reference's shipped nets on the device (net_set 1, nrrs_shipped_config;
2net-simple.json:29-75, 210-291) against the FP64 oracle:

- the input channels (nrrs_encode): StatNet's OneBlob of the lat-long direction and of the
  roughness, RRSNet's NRRSNetworkInput1 channels (no grid), the AID RRSNet's channels;
- StatNet statistics and the RRS stage (StatNet -> RRSNet -> q, and the AID RRSNet alone) in
  FP32/FP16/BF16, to the tolerances of the benchmark nets, and the decision bit-exact given the
  device q;
- the joint training step in the linear-Adam regime of the joint-step test (LeakyReLU backward,
  the grid-less RRSNet, the AID RRSNet);
- the weights outside the 32-wide nets stay zero through training;
- the entry points the shipped nets do not support refuse them."""
import numpy as np
import pytest

import nrrs_amd as na
from test_gpu_joint import LINEAR_ADAM, _batch, _device_step, _joint_step_vs_fp64
from test_gpu_rrs import OUT_TOL, Q_TOL, Q_TOL_ADRRS, _decide_buffers, make_film, random_params

pytestmark = pytest.mark.gpu

PRECISIONS = [na.FP32, na.FP16, na.BF16]
N = 1 << 14


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    assert torch.cuda.is_available()
    return torch


def _cfg(prec, **kw):
    kw.setdefault("ema_decay", 0.0)
    return na.shipped_config(precision=prec, **kw)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


@pytest.mark.parametrize("aid", [0, 1])
def test_shipped_encode_matches_oracle(torch_cuda, ora, aid):
    torch = torch_cuda
    ctx = na.Context(_cfg(na.FP32, rrs_aid=aid))
    vb = na.synth_vertices(4096, 21, ctx.cfg.n_pixels)
    hv = ora.HostVertices(vb.to_numpy())
    film_np, film = make_film(torch, ctx.cfg.n_pixels, 5)
    stats = torch.rand(6, 4096, device="cuda") * 3.0 - 0.5
    for net in (na.STATNET, na.RRSNET):
        p = random_params(ctx, net, 3 + net)
        st = stats if net == na.RRSNET else None
        enc = ctx.encode(net, vb, film, st).cpu().numpy()
        ref = ora.encode(ctx.cfg, net, hv, film_np, None if st is None else st.cpu().numpy().astype(np.float64),
                         p.astype(np.float64))
        # FP32 trilinear, lat-long and OneBlob arithmetic against FP64
        assert np.allclose(enc, ref, rtol=1e-5, atol=4e-6), net
    assert ctx.param_count(na.RRSNET)[0] == (ctx.param_count(na.STATNET)[0] if aid else 0)


@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("n", [N, 3001])
def test_shipped_rrs_stage_matches_fp64(torch_cuda, ora, prec, n):
    """The chained RRS stage (FP16/BF16: one kernel from StatNet to q; FP32: a StatNet kernel,
    then RRSNet on its statistics) and the ADRRS form against the FP64 oracle, at a full and a
    ragged batch (3001: no multiple of the 32-vertex blocks)."""
    torch = torch_cuda
    N = n
    ctx = na.Context(_cfg(prec, n_pixels=N, max_queue=N, max_train=N))
    ps = random_params(ctx, na.STATNET, 11)
    pr = random_params(ctx, na.RRSNET, 12)
    vb = na.synth_vertices(N, 31, N)
    hv = ora.HostVertices(vb.to_numpy())
    film_np, film = make_film(torch, N, 6)
    stats = ctx.statnet_infer(vb).cpu().numpy()
    ref_stats = ora.net_forward(ctx.cfg, na.STATNET, hv, film_np, None, ps.astype(np.float64)).T
    e_stats = _rel(stats, ref_stats)
    # chained RRS stage: StatNet statistics -> NRRSNetworkInput1 -> RRSNet -> q
    b = _decide_buffers(torch, N, 8 * N)
    args = ctx.decide_args(vb, film, b["q"], b["survivors"], b["n_surv"], b["overflow"],
                           q_mode=na.Q_RRSNET_CHAINED, seed=5, stream_id=1, qn=b["qn"], split=b["split"])
    assert ctx.rrs_infer_decide(args) == na.OK
    torch.cuda.synchronize()
    q = b["q"][:N].cpu().numpy()
    raw64 = ora.net_forward(ctx.cfg, na.RRSNET, hv, film_np, ref_stats, pr.astype(np.float64))[:, 0]
    q64 = ora.q_f64(ctx.cfg, na.Q_RRSNET, hv, raw64, film_np)
    qerr = np.abs(q - q64) / q64
    # ADRRS form from the StatNet statistics
    b2 = _decide_buffers(torch, N, 8 * N)
    args2 = ctx.decide_args(vb, film, b2["q"], b2["survivors"], b2["n_surv"], b2["overflow"],
                            q_mode=na.Q_STATNET_ADRRS, seed=5, stream_id=1)
    assert ctx.rrs_infer_decide(args2) == na.OK
    torch.cuda.synchronize()
    qa = b2["q"][:N].cpu().numpy()
    qa64 = ora.q_f64(ctx.cfg, na.Q_STATNET_ADRRS, hv, ref_stats, film_np)
    print(f"shipped prec {prec}: stats err {e_stats:.2e}, q rel err max {qerr.max():.2e} mean {qerr.mean():.2e}, "
          f"ADRRS q err {np.max(np.abs(qa - qa64)) / qa64.max():.2e}")
    assert e_stats < OUT_TOL[prec]
    assert qerr.max() < Q_TOL[prec]
    assert np.max(np.abs(qa - qa64)) / qa64.max() < Q_TOL_ADRRS[prec]
    ref = ora.rrs_decide(ctx.cfg, q, stid=None, seed=5, stream=1)
    ns = int(b["n_surv"].item())
    assert ns == ref["n_survivors"]
    assert np.array_equal(b["split"][:N].cpu().numpy(), ref["s"])
    assert np.array_equal(b["survivors"][:ns].cpu().numpy().astype(np.uint32), ref["survivors"])


@pytest.mark.parametrize("prec", PRECISIONS)
def test_shipped_aid_rrs_stage_matches_fp64(torch_cuda, ora, prec):
    torch = torch_cuda
    ctx = na.Context(_cfg(prec, rrs_aid=1, n_pixels=N, max_queue=N, max_train=N))
    pr = random_params(ctx, na.RRSNET, 42)
    vb = na.synth_vertices(N, 77, N)
    hv = ora.HostVertices(vb.to_numpy())
    film_np, film = make_film(torch, N, 8)
    b = _decide_buffers(torch, N, 8 * N)
    args = ctx.decide_args(vb, film, b["q"], b["survivors"], b["n_surv"], b["overflow"], q_mode=na.Q_RRSNET,
                           seed=5, stream_id=1)
    assert ctx.rrs_infer_decide(args) == na.OK
    torch.cuda.synchronize()
    q = b["q"][:N].cpu().numpy()
    raw64 = ora.net_forward(ctx.cfg, na.RRSNET, hv, film_np, None, pr.astype(np.float64))[:, 0]
    q64 = ora.q_f64(ctx.cfg, na.Q_RRSNET, hv, raw64, film_np)
    qerr = np.abs(q - q64) / q64
    print(f"shipped AID prec {prec}: q rel err max {qerr.max():.2e} mean {qerr.mean():.2e}")
    assert qerr.max() < Q_TOL[prec]


@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("aid", [0, 1])
def test_shipped_joint_step_matches_fp64(torch_cuda, ora, prec, aid):
    _joint_step_vs_fp64(torch_cuda, ora, prec, 4096, 1, config=_cfg_kw, rrs_aid=aid)


def test_shipped_joint_step_ragged_batch(torch_cuda, ora):
    """BF16 without the paired table, at a batch that is not a multiple of the tile."""
    _joint_step_vs_fp64(torch_cuda, ora, na.BF16, 3001, 0, config=_cfg_kw)


def _cfg_kw(**kw):
    kw.setdefault("ema_decay", 0.0)
    return na.shipped_config(**kw)


def _live(p, g, n_out, net, aid):
    """Whether every weight outside the 32-wide net is zero."""
    pos, fin = g, 80
    cols = {na.STATNET: np.r_[0:16, 32:48]}.get(net, np.r_[0:16, 32:44, 65:69] if aid else np.r_[64:80])
    for layer in range(4):
        o = n_out if layer == 3 else 64
        W = p[pos:pos + o * fin].reshape(o, fin)
        mk = np.zeros((o, fin), bool)
        rows = slice(0, o) if layer == 3 else slice(0, 32)
        if layer == 0:
            mk[rows, cols] = True
        else:
            mk[rows, :32] = True
        if W[~mk].any():
            return False
        pos += o * fin
        fin = o
    return True


@pytest.mark.parametrize("prec", PRECISIONS)
def test_shipped_padding_stays_zero(torch_cuda, ora, prec):
    """The weights outside the 32-wide nets: zero after nrrs_params_set of random values, and
    still zero after training steps (they receive exact zero gradients), while the live weights
    move."""
    torch = torch_cuda
    n = 4096
    ctx = na.Context(_cfg(prec, n_pixels=n, max_queue=n, max_train=n))
    vb, film_np, tL, side = _batch(torch, ora, ctx, n, seed=4)
    rng = np.random.default_rng(0)
    for net in (na.STATNET, na.RRSNET):
        p = ctx.params_get(net)
        g, _ = ctx.param_count(net)
        q = p.copy()
        q[g:] = rng.normal(0, 0.1, q.size - g).astype(np.float32)
        ctx.params_set(net, q)
        assert _live(ctx.params_get(net), g, 6 if net == na.STATNET else 1, net, 0)
    p0 = [ctx.params_get(net) for net in (na.STATNET, na.RRSNET)]
    for _ in range(3):
        loss = _device_step(torch, ctx, vb, film_np, tL, side)
        assert np.all(np.isfinite(loss))
    for net in (na.STATNET, na.RRSNET):
        p = ctx.params_get(net)
        g, _ = ctx.param_count(net)
        assert _live(p, g, 6 if net == na.STATNET else 1, net, 0)
        assert np.any(p[g:] != p0[net][g:])
        assert ctx.train_state(net)["step"] == 3


def test_shipped_refusals(torch_cuda):
    torch = torch_cuda
    n = 4096
    ctx = na.Context(_cfg(na.BF16, n_pixels=n, max_queue=2 * n, max_train=n))
    vb = na.synth_vertices(n, 1, n)
    film = torch.ones(3, n, device="cuda")
    b = _decide_buffers(torch, n, 8 * n)
    # RRSNet on NRRSNetworkInput1 needs StatNet statistics
    args = ctx.decide_args(vb, film, b["q"], b["survivors"], b["n_surv"], b["overflow"], q_mode=na.Q_RRSNET)
    assert ctx.rrs_infer_decide(args) == na.ERR_UNSUPPORTED
    with pytest.raises(na.NrrsError) as ei:
        ctx.rrsnet_infer(vb, film)
    assert ei.value.code == na.ERR_ARG
    a, _ = ctx.frame_buffers(n, 2, capacity=2 * n)
    a.film_dev = film.data_ptr()
    assert ctx.frame_run(a) == na.ERR_UNSUPPORTED
    # two-headed StatNets and other depths are not shipped shapes
    for kw in (dict(ll2_two_head=1, statnet=dict(n_hidden_layers=3, n_outputs=6)),
               dict(statnet=dict(n_hidden_layers=4, n_outputs=6))):
        with pytest.raises(na.NrrsError) as ei:
            na.Context(_cfg(na.BF16, **kw))
        assert ei.value.code == na.ERR_UNSUPPORTED
