"""Per-pixel error pipeline on the GPU against the oracle: the running radiance blend, the
relative error rows and their device-side sum, and the RRS loss fed from that device sum."""
import numpy as np
import pytest

import nrrs_amd as na

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    assert torch.cuda.is_available()
    return torch


@pytest.mark.parametrize("n", [1, 1000, 1 << 18])
def test_error_pipeline_matches_oracle(torch_cuda, ora, n):
    torch = torch_cuda
    ctx = na.Context(n_pixels=max(n, 16))
    rng = np.random.default_rng(n)
    ref = rng.lognormal(-1, 1, (3, n)).astype(np.float32)
    ref[:, rng.uniform(size=n) < 0.05] = 0.0          # black pixels
    w0 = (ref * rng.lognormal(0, 0.5, (3, n))).astype(np.float32)
    Lf = rng.lognormal(-1, 1.5, (3, n)).astype(np.float32)
    spp = rng.integers(1, 16, n).astype(np.uint32)
    sppf = rng.integers(1, 64, n).astype(np.float32)
    wL = torch.from_numpy(w0).cuda()
    ctx.error_blend(torch.from_numpy(Lf).cuda(), torch.from_numpy(spp.astype(np.int32)).cuda(), 0.3, wL)
    w_ref = w0.astype(np.float64)
    ora.error_blend(Lf, spp, 0.3, w_ref)
    w_gpu = wL.cpu().numpy()
    # FP32 blend of FP32 inputs: a few ulp
    assert np.allclose(w_gpu, w_ref, rtol=2e-6, atol=1e-30)
    for clamp in (True, False):
        err, total = ctx.error_compute(torch.from_numpy(ref).cuda(), wL, torch.from_numpy(sppf).cuda(), clamp=clamp)
        e_ref, t_ref = ora.error_compute(ref, w_gpu, sppf, clamp)
        e = err.cpu().numpy()
        # FP32 arithmetic on identical inputs against FP64: relative 2e-6 per pixel
        assert np.allclose(e, e_ref, rtol=2e-6, atol=1e-12)
        assert abs(float(total.item()) - t_ref) <= 1e-6 * t_ref
    err, total = ctx.error_compute(torch.from_numpy(ref).cuda(), wL, None, clamp=True)
    assert np.array_equal(err[0].cpu().numpy(), err[1].cpu().numpy())


def test_rrs_loss_reads_device_error_sum(torch_cuda):
    """The RRS loss with error_sum_dev equals the one with the same value given on the host."""
    torch = torch_cuda
    n = 4096
    ctx = na.Context(precision=na.FP32, n_pixels=n, max_queue=n, max_train=n)
    vb = na.synth_vertices(n, 3, n)
    rng = np.random.default_rng(4)
    film = torch.from_numpy(rng.lognormal(-1, 1, (3, n)).astype(np.float32)).cuda()
    stats = torch.from_numpy(np.abs(rng.normal(0.5, 0.5, (6, n))).astype(np.float32)).cuda()
    err = torch.from_numpy(rng.uniform(0, 2, n).astype(np.float32)).cuda()
    sw = torch.ones(n, device="cuda")
    qt = torch.from_numpy(rng.uniform(0.05, 5, n).astype(np.float32)).cuda()
    esum = torch.tensor([1234.5], device="cuda")
    losses = []
    for dev in (False, True):
        side = ctx.side(error=err, sample_weight=sw, q_target=qt, error_sum=0.0 if dev else 1234.5,
                        error_sum_dev=esum if dev else None)
        loss = torch.zeros(1, device="cuda")
        na.check(ctx.backward(na.RRSNET, vb, na.LOSS_RRS, film=film, stats=stats, side=side, loss=loss), "bwd")
        losses.append(float(loss.item()))
        ctx.grads_zero(na.RRSNET)
    # same ebar; the batch loss is summed across workgroups with atomics (order-dependent bits)
    assert losses[0] != 0.0 and abs(losses[0] - losses[1]) <= 1e-6 * abs(losses[1])
