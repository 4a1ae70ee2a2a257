"""The persistent multi-bounce frame kernel against the per-bounce entry points
and the oracle: per-bounce q equal to the stand-alone chained inference, every
queue bit-exact to the decision + stable compaction of the oracle, identical
results for every tile shape, and hipGraph replays equal to direct launches."""
import numpy as np
import pytest

import nrrs_amd as na

pytestmark = pytest.mark.gpu

P = 1 << 14
BOUNCES = 5


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    assert torch.cuda.is_available()
    return torch


def _ctx(torch, prec=na.BF16):
    ctx = na.Context(precision=prec, n_pixels=P, max_queue=2 * P)
    rng = np.random.default_rng(1)
    for net in (na.STATNET, na.RRSNET):
        p = ctx.params_get(net)
        g, _ = ctx.param_count(net)
        p[:g] = rng.uniform(-0.3, 0.3, g).astype(np.float32)
        ctx.params_set(net, p)
    film_np = rng.lognormal(-1, 1, (3, P)).astype(np.float32)
    return ctx, film_np, torch.from_numpy(film_np).cuda()


def _run(torch, ctx, film, seed, shape=None):
    if shape:
        ctx.frame_set_shape(*shape)
    a, b = ctx.frame_buffers(P, BOUNCES, capacity=2 * P, trace=True, record_q=True)
    a.film_dev = film.data_ptr()
    a.seed = seed
    na.check(ctx.frame_run(a), "frame_run")
    torch.cuda.synchronize()
    assert int(b["status"].item()) == 0
    return {k: v.cpu().numpy() for k, v in b.items()}


@pytest.mark.parametrize("prec", [na.BF16, na.FP16])
def test_frame_matches_per_bounce_stage_and_oracle(torch_cuda, ora, prec):
    torch = torch_cuda
    ctx, film_np, film = _ctx(torch, prec)
    seed = 0x1234567
    r = _run(torch, ctx, film, seed)
    sizes = r["sizes"].astype(np.int64)
    print("queue sizes per bounce:", sizes.tolist())
    assert sizes[0] == P
    for b in range(BOUNCES):
        n = int(sizes[b])
        queue = r["trace"][b, :n].astype(np.uint32)
        qf = r["q"][b, :n]
        # the stand-alone chained stage on the same vertices gives the same q
        vb = na.frame_vertices(seed, b, torch.from_numpy(queue.astype(np.int32)).cuda(), n)
        q = torch.empty(n, device="cuda")
        surv = torch.empty(8 * n + 1, dtype=torch.int32, device="cuda")
        ns = torch.zeros(1, dtype=torch.int32, device="cuda")
        args = ctx.decide_args(vb, film, q, surv, ns, q_mode=na.Q_RRSNET_CHAINED, seed=seed, stream_id=1000 + b)
        na.check(ctx.rrs_infer_decide(args), "rrs_infer_decide")
        torch.cuda.synchronize()
        qs = q.cpu().numpy()
        # bit-identical: both kernels run the same encoding, MLP and epilogue code, and device
        # code fuses multiply-adds only within one expression, so its rounding does not depend
        # on the basic blocks each kernel gives that code
        print(f"bounce {b}: {int(np.sum(qs != qf))} of {n} q differ, max rel {np.max(np.abs(qs - qf) / qf):.2e}")
        assert np.array_equal(qs, qf)
        # the next queue is the oracle's decision + stable compaction of this bounce's q
        ref = ora.rrs_decide(ctx.cfg, qf, stid=queue, seed=seed, stream=1000 + b, capacity=2 * P)
        assert int(sizes[b + 1]) == ref["n_survivors"]
        assert np.array_equal(r["trace"][b + 1, : sizes[b + 1]].astype(np.uint32), ref["survivors"])


def test_frame_results_independent_of_tile_shape(torch_cuda):
    torch = torch_cuda
    ctx, _, film = _ctx(torch)
    ref = _run(torch, ctx, film, 99, shape=(4, 128))
    for shape in ((1, 64), (1, 256), (2, 64), (2, 128), (4, 256), (8, 256)):
        r = _run(torch, ctx, film, 99, shape=shape)
        assert np.array_equal(r["sizes"], ref["sizes"]), shape
        for b in range(BOUNCES + 1):
            n = int(ref["sizes"][b])
            assert np.array_equal(r["trace"][b, :n], ref["trace"][b, :n]), (shape, b)


def test_frame_autotune_and_graph_replay(torch_cuda):
    torch = torch_cuda
    ctx, _, film = _ctx(torch)
    a, b = ctx.frame_buffers(P, BOUNCES, capacity=2 * P)
    a.film_dev = film.data_ptr()
    a.seed = 7
    shape, ms = ctx.frame_autotune(a)
    print("autotuned shape", shape, f"{ms:.3f} ms")
    assert shape[1] % (32 * shape[0]) == 0
    # graph with a device seed advanced by every replay
    seed_dev = torch.tensor([1000], dtype=torch.int64, device="cuda")
    a.seed_dev = seed_dev.data_ptr()
    a.advance_seed = 1
    stream = torch.cuda.Stream()
    ctx.frame_graph_capture(a, stream)
    got = []
    for _ in range(3):
        b["status"].zero_()
        ctx.frame_graph_launch(stream)
        stream.synchronize()
        assert int(b["status"].item()) == 0
        got.append(b["sizes"].cpu().numpy().copy())
        torch.cuda.synchronize()
    # direct launches with the seeds the replays used
    s = 1000
    for k in range(3):
        r = _run(torch, ctx, film, s)
        assert np.array_equal(r["sizes"], got[k]), k
        s = (s + 0xD1B54A32D192ED03) & (2 ** 64 - 1)


def test_frame_graph_survives_refused_capacity(torch_cuda):
    """The frame scratch is allocated once, for max_queue vertices, by nrrs_create: a frame asking
    for more is refused (NRRS_ERR_SHAPE) and a graph captured before still replays as captured."""
    torch = torch_cuda
    ctx, _, film = _ctx(torch)
    a, b = ctx.frame_buffers(P, BOUNCES, capacity=2 * P)
    a.film_dev = film.data_ptr()
    a.seed = 11
    stream = torch.cuda.Stream()
    ctx.frame_graph_capture(a, stream)
    ctx.frame_graph_launch(stream)
    stream.synchronize()
    first = b["sizes"].cpu().numpy().copy()
    big, bb = ctx.frame_buffers(P, BOUNCES, capacity=4 * P)
    big.film_dev = film.data_ptr()
    assert ctx.frame_run(big) == na.ERR_SHAPE
    b["status"].zero_()
    b["sizes"].zero_()
    torch.cuda.synchronize()
    ctx.frame_graph_launch(stream)
    stream.synchronize()
    assert int(b["status"].item()) == 0
    assert np.array_equal(b["sizes"].cpu().numpy(), first)
