"""The AID RRSNet input on the device (rrs_aid; net1_aid_net2, network.cpp:31, 209-258,
303-343; NRRSNetworkInput1AID, train.h:58-68) against the FP64 oracle:

- the input channels (nrrs_encode) and RRSNet's output, whatever statistics the caller passes;
- the inference stage NRRS_Q_RRSNET (RRSNet alone, network.cpp:303-308) in FP32/FP16/BF16, to
  the one-net tolerances, and its decision bit-exact given the device q;
- the joint training step: StatNet as in the one-net step, RRSNet on the AID input with the
  NRRS_RRS loss on StatNet's statistics of the batch, after one step, in the linear-Adam regime
  of the joint-step test;
- the chained inference mode and the frame kernel refuse AID contexts."""
import numpy as np
import pytest

import nrrs_amd as na
from test_gpu_joint import _joint_step_vs_fp64
from test_gpu_rrs import OUT_TOL, Q_TOL, _decide_buffers, make_film, random_params

pytestmark = pytest.mark.gpu

PRECISIONS = [na.FP32, na.FP16, na.BF16]
N = 1 << 14


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    assert torch.cuda.is_available()
    return torch


def _ctx(prec, **kw):
    return na.Context(na.default_config(precision=prec, rrs_aid=1, **kw))


def test_aid_encode_matches_oracle(torch_cuda, ora):
    torch = torch_cuda
    ctx = _ctx(na.FP32)
    p = random_params(ctx, na.RRSNET, 3)
    vb = na.synth_vertices(4096, 21, ctx.cfg.n_pixels)
    film_np, film = make_film(torch, ctx.cfg.n_pixels, 5)
    stats = torch.rand(6, 4096, device="cuda") * 3.0
    enc = ctx.encode(na.RRSNET, vb, film, stats).cpu().numpy()
    ref = ora.encode(ctx.cfg, na.RRSNET, ora.HostVertices(vb.to_numpy()), film_np, None, p.astype(np.float64))
    assert np.allclose(enc, ref, rtol=1e-5, atol=2e-6)
    assert not enc[:, 48:64].any() and not enc[:, 69:75].any()


@pytest.mark.parametrize("prec", PRECISIONS)
def test_aid_rrs_stage_matches_fp64(torch_cuda, ora, prec):
    torch = torch_cuda
    ctx = _ctx(prec)
    pr = random_params(ctx, na.RRSNET, 42)
    vb = na.synth_vertices(N, 77, ctx.cfg.n_pixels)
    film_np, film = make_film(torch, ctx.cfg.n_pixels, 8)
    hv = ora.HostVertices(vb.to_numpy())
    raw64 = ora.net_forward(ctx.cfg, na.RRSNET, hv, film_np, None, pr.astype(np.float64))[:, 0]
    # RRSNet's raw output with and without caller statistics: the AID input ignores them
    raw = ctx.rrsnet_infer(vb, film).cpu().numpy()
    raw_s = ctx.rrsnet_infer(vb, film, torch.rand(6, N, device="cuda")).cpu().numpy()
    assert np.array_equal(raw, raw_s)
    e_raw = np.max(np.abs(raw - raw64)) / np.max(np.abs(raw64))
    b = _decide_buffers(torch, N, 8 * N)
    args = ctx.decide_args(vb, film, b["q"], b["survivors"], b["n_surv"], b["overflow"], q_mode=na.Q_RRSNET, seed=5,
                           stream_id=1, qn=b["qn"], split=b["split"])
    assert ctx.rrs_infer_decide(args) == na.OK
    torch.cuda.synchronize()
    q = b["q"][:N].cpu().numpy()
    q64 = ora.q_f64(ctx.cfg, na.Q_RRSNET, hv, raw64, film_np)
    qerr = np.abs(q - q64) / q64
    print(f"AID precision {prec}: raw err {e_raw:.2e}, q rel err max {qerr.max():.2e} mean {qerr.mean():.2e}")
    assert e_raw < OUT_TOL[prec]
    assert qerr.max() < Q_TOL[prec]
    ref = ora.rrs_decide(ctx.cfg, q, stid=None, seed=5, stream=1)
    ns = int(b["n_surv"].item())
    assert ns == ref["n_survivors"]
    assert np.array_equal(b["split"][:N].cpu().numpy(), ref["s"])
    assert np.array_equal(b["survivors"][:ns].cpu().numpy().astype(np.uint32), ref["survivors"])


@pytest.mark.parametrize("prec", PRECISIONS)
def test_aid_joint_step_matches_fp64(torch_cuda, ora, prec):
    _joint_step_vs_fp64(torch_cuda, ora, prec, 4096, 1, rrs_aid=1)


def test_aid_joint_step_per_net_path(torch_cuda, ora):
    """BF16 without the paired table, at a batch that is not a multiple of the tile."""
    _joint_step_vs_fp64(torch_cuda, ora, na.BF16, 3001, 0, rrs_aid=1)


def test_aid_refuses_chained_and_frame(torch_cuda):
    torch = torch_cuda
    ctx = _ctx(na.BF16, n_pixels=4096, max_queue=8192, max_train=4096)
    vb = na.synth_vertices(4096, 1, 4096)
    film = torch.ones(3, 4096, device="cuda")
    b = _decide_buffers(torch, 4096, 8 * 4096)
    args = ctx.decide_args(vb, film, b["q"], b["survivors"], b["n_surv"], b["overflow"], q_mode=na.Q_RRSNET_CHAINED)
    assert ctx.rrs_infer_decide(args) == na.ERR_UNSUPPORTED
    a, _ = ctx.frame_buffers(4096, 2, capacity=8192)
    a.film_dev = film.data_ptr()
    assert ctx.frame_run(a) == na.ERR_UNSUPPORTED
