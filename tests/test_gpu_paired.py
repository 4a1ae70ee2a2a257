"""The paired table (both nets' FP16/BF16 grid entries interleaved per entry, so that one gather
serves both nets) against the per-net tables: the chained RRS stage, the joint training step
and the persistent frame kernel give bitwise the same results with paired_tables = 1 and 0,
including after the Adam steps that refresh each net's half of the paired table."""
import numpy as np
import pytest

import nrrs_amd as na

pytestmark = pytest.mark.gpu

PRECISIONS = [na.BF16, na.FP16]


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    assert torch.cuda.is_available()
    return torch


def _contexts(prec, n, **kw):
    """Contexts with paired_tables 1 and 0 holding the same parameters (grid U[-0.3, 0.3])."""
    ctxs = [na.Context(na.default_config(precision=prec, n_pixels=n, max_queue=2 * n, max_train=n,
                                         paired_tables=paired, **kw)) for paired in (1, 0)]
    rng = np.random.default_rng(5)
    for net in (na.STATNET, na.RRSNET):
        p = ctxs[0].params_get(net)
        g, _ = ctxs[0].param_count(net)
        p[:g] = rng.uniform(-0.3, 0.3, g).astype(np.float32)
        for c in ctxs:
            c.params_set(net, p)
    return ctxs


@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("n", [4096, 3001])
def test_chained_stage_paired_equals_per_net(torch_cuda, prec, n):
    torch = torch_cuda
    ctxs = _contexts(prec, n)
    vb = na.synth_vertices(n, 21, n)
    film = torch.from_numpy(np.random.default_rng(2).lognormal(-1, 1, (3, n)).astype(np.float32)).cuda()
    out = []
    for ctx in ctxs:
        b = dict(q=torch.empty(n, device="cuda"), qn=torch.empty(n, device="cuda"),
                 split=torch.empty(n, dtype=torch.uint8, device="cuda"),
                 surv=torch.empty(8 * n, dtype=torch.int32, device="cuda"),
                 ns=torch.zeros(1, dtype=torch.int32, device="cuda"))
        args = ctx.decide_args(vb, film, b["q"], b["surv"], b["ns"], q_mode=na.Q_RRSNET_CHAINED, seed=9,
                               stream_id=4, qn=b["qn"], split=b["split"])
        na.check(ctx.rrs_infer_decide(args), "rrs_infer_decide")
        torch.cuda.synchronize()
        out.append({k: v.cpu().numpy() for k, v in b.items()})
    p, s = out
    print(f"prec {prec} n {n}: {int(p['ns'][0])} survivors, q in [{p['q'].min():.3g}, {p['q'].max():.3g}]")
    for k in ("q", "qn", "split", "ns"):
        assert np.array_equal(p[k], s[k]), k
    assert np.array_equal(p["surv"][: int(p["ns"][0])], s["surv"][: int(s["ns"][0])])


@pytest.mark.parametrize("prec", PRECISIONS)
def test_joint_training_step_paired_equals_per_net(torch_cuda, ora, prec):
    torch = torch_cuda
    n = 3001
    ctxs = _contexts(prec, n)
    vb = na.synth_vertices(n, 22, n)
    rng = np.random.default_rng(6)
    film = torch.from_numpy(rng.lognormal(-1, 1, (3, n)).astype(np.float32)).cuda()
    tL = torch.from_numpy(ora.synth_targets(n, 12)).cuda()
    qt = rng.uniform(0.05, 5.0, n).astype(np.float32)
    keep = dict(error=torch.from_numpy(rng.uniform(0, 2, n).astype(np.float32)).cuda(),
                sample_weight=torch.from_numpy((1.0 / rng.integers(1, 4, n)).astype(np.float32)).cuda(),
                q_target=torch.from_numpy(qt).cuda())
    res = []
    for ctx in ctxs:
        side = ctx.side(error=keep["error"], sample_weight=keep["sample_weight"], q_target=keep["q_target"],
                        error_sum=0.7 * n)
        loss = torch.zeros(2, device="cuda")
        losses = []
        # StatNet alone (its update then refreshes the paired table), then both nets jointly
        for mask in (1, 1, 3, 3, 3):
            na.check(ctx.train_step(vb, mask, film=film, target_L=tL, side=side, loss=loss), "train_step")
            losses.append(loss.cpu().numpy().copy())
        # the chained stage on the trained parameters reads the refreshed paired table
        q = torch.empty(n, device="cuda")
        surv = torch.empty(8 * n, dtype=torch.int32, device="cuda")
        ns = torch.zeros(1, dtype=torch.int32, device="cuda")
        args = ctx.decide_args(vb, film, q, surv, ns, q_mode=na.Q_RRSNET_CHAINED, seed=3, stream_id=2)
        na.check(ctx.rrs_infer_decide(args), "rrs_infer_decide")
        torch.cuda.synchronize()
        res.append(dict(losses=np.array(losses), stat=ctx.params_get(na.STATNET), rrs=ctx.params_get(na.RRSNET),
                        q=q.cpu().numpy()))
    a, b = res
    print(f"prec {prec}: StatNet losses {a['losses'][:, 0].tolist()} / {b['losses'][:, 0].tolist()}, "
          f"RRSNet {a['losses'][:, 1].tolist()} / {b['losses'][:, 1].tolist()}")
    for k in ("stat", "rrs"):
        d = np.abs(a[k] - b[k])
        print(f"{k}: {int(np.sum(d > 0))} of {d.size} parameters differ, max {d.max():.3g}")
    # the parameters (gradients summed exactly in fixed point, dW reduced in a fixed order)
    # are bitwise equal; the loss values are summed with float atomics in no fixed order
    assert np.array_equal(a["stat"], b["stat"]) and np.array_equal(a["rrs"], b["rrs"])
    np.testing.assert_allclose(a["losses"], b["losses"], rtol=1e-5)
    assert np.array_equal(a["q"], b["q"])


def test_frame_paired_equals_per_net(torch_cuda):
    torch = torch_cuda
    P = 1 << 13
    ctxs = _contexts(na.BF16, P)
    film = torch.from_numpy(np.random.default_rng(8).lognormal(-1, 1, (3, P)).astype(np.float32)).cuda()
    out = []
    for ctx in ctxs:
        a, b = ctx.frame_buffers(P, 4, capacity=2 * P, trace=True, record_q=True)
        a.film_dev = film.data_ptr()
        a.seed = 31
        na.check(ctx.frame_run(a), "frame_run")
        torch.cuda.synchronize()
        assert int(b["status"].item()) == 0
        out.append({k: v.cpu().numpy() for k, v in b.items()})
    p, s = out
    print("queue sizes per bounce:", p["sizes"].tolist())
    assert np.array_equal(p["sizes"], s["sizes"])
    for bn in range(4):
        n = int(p["sizes"][bn])
        assert np.array_equal(p["trace"][bn, :n], s["trace"][bn, :n]), bn
        assert np.array_equal(p["q"][bn, :n], s["q"][bn, :n]), bn


def test_joint_step_with_full_segments(torch_cuda, ora):
    """Every vertex on one z-plane overfills the record slots of dense-level segments, so the
    joint record pass adds the records that do not fit to both nets' table gradients with
    packed FP16 atomics, whose order is not fixed. One joint step then matches the per-net path
    except where a gradient's sign is decided by that order: Adam's first step moves each
    parameter by lr * sign(gradient)."""
    torch = torch_cuda
    n = 4096
    ctxs = _contexts(na.BF16, n)
    vb = na.synth_vertices(n, 23, n)
    vb.pos[2, :n] = 0.37
    # precondition: some dense-level segment receives more records than its slots (records: one
    # per x-neighbour corner pair whose entries differ by 1, 3 or 7, else one per corner; slots
    # of a dense segment: min(8n, 3 x its share of 4.5n + 4096), rounded up to 64)
    hv = ora.HostVertices(vb.to_numpy())
    lay = ora.grid_layout(ctxs[0].cfg.grid)
    idx = ora.grid_indices(ctxs[0].cfg, hv).astype(np.int64)
    over = False
    for l in range(ctxs[0].cfg.grid.n_levels):
        if not lay["dense"][l]:
            continue
        size = int(lay["size"][l])
        e0, e1 = idx[:, l, 0::2].ravel(), idx[:, l, 1::2].ravel()
        single = ~np.isin(e0 ^ e1, [1, 3, 7])
        counts = np.bincount(np.concatenate([e0 >> 13, e1[single] >> 13]), minlength=(size + 8191) // 8192)
        for b, c in enumerate(counts):
            ne = min(8192, size - b * 8192)
            cap = min(8 * n, 3 * (9 * n // 2 * ne // size) + 4096)
            over |= c > (cap + 63) // 64 * 64
    assert over
    rng = np.random.default_rng(7)
    film = torch.from_numpy(rng.lognormal(-1, 1, (3, n)).astype(np.float32)).cuda()
    tL = torch.from_numpy(ora.synth_targets(n, 13)).cuda()
    keep = dict(error=torch.from_numpy(rng.uniform(0, 2, n).astype(np.float32)).cuda(),
                sample_weight=torch.ones(n, device="cuda"),
                q_target=torch.from_numpy(rng.uniform(0.05, 5.0, n).astype(np.float32)).cuda())
    res = []
    for ctx in ctxs:
        side = ctx.side(error=keep["error"], sample_weight=keep["sample_weight"], q_target=keep["q_target"],
                        error_sum=0.7 * n)
        p0 = {net: ctx.params_get(net) for net in (na.STATNET, na.RRSNET)}
        na.check(ctx.train_step(vb, 3, film=film, target_L=tL, side=side), "train_step")
        res.append({net: ctx.params_get(net) - p0[net] for net in (na.STATNET, na.RRSNET)})
    for net in (na.STATNET, na.RRSNET):
        a, b = res[0][net], res[1][net]
        moved = np.abs(b) > 0
        differ = a != b
        print(f"net {net}: {int(moved.sum())} parameters moved, {int(differ.sum())} differ")
        assert moved.sum() > 1000
        assert differ.sum() <= 1e-3 * moved.sum() + 8


def test_paired_table_follows_single_net_updates(torch_cuda, ora):
    """A StatNet update through nrrs_backward + nrrs_adam_step (not the joint step) refreshes the
    paired table: the chained stage afterwards is bitwise the same as with per-net tables."""
    torch = torch_cuda
    n = 3001
    ctxs = _contexts(na.BF16, n)
    vb = na.synth_vertices(n, 24, n)
    film = torch.from_numpy(np.random.default_rng(9).lognormal(-1, 1, (3, n)).astype(np.float32)).cuda()
    tL = torch.from_numpy(ora.synth_targets(n, 14)).cuda()
    qs = []
    for ctx in ctxs:
        for _ in range(2):
            na.check(ctx.backward(na.STATNET, vb, na.LOSS_LL2, target_L=tL, loss_scale=4096.0), "backward")
            ctx.adam_step(na.STATNET)
        q = torch.empty(n, device="cuda")
        surv = torch.empty(8 * n, dtype=torch.int32, device="cuda")
        ns = torch.zeros(1, dtype=torch.int32, device="cuda")
        args = ctx.decide_args(vb, film, q, surv, ns, q_mode=na.Q_RRSNET_CHAINED, seed=4, stream_id=6)
        na.check(ctx.rrs_infer_decide(args), "rrs_infer_decide")
        torch.cuda.synchronize()
        qs.append(q.cpu().numpy())
    assert ctxs[0].train_state(na.STATNET)["step"] == 2
    assert np.array_equal(qs[0], qs[1])
