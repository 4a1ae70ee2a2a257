"""The allocation-free decision (nrrs_rrs_decide_static, fixed device scratch, no context)
against the context's nrrs_rrs_decide on the same q: normalised q', split counts, survivor queue,
its length and overflow bit-exact, on two calls in a row (the second smaller, so that stale
scratch from the first would show), with and without caller-kept split counts."""
import numpy as np
import pytest

import nrrs_amd as na

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    assert torch.cuda.is_available()
    return torch


def _case(torch, n, n_spec, seed, keep_split):
    rng = np.random.default_rng(seed)
    q = rng.uniform(0.05, 20.0, n).astype(np.float32)
    q[: n // 7] = rng.uniform(0.05, 1.0, n // 7).astype(np.float32)
    stid = rng.permutation(4 * n)[:n].astype(np.int32)
    spec = (np.arange(n_spec) + 4 * n + 3).astype(np.int32)
    cap = 8 * n + n_spec
    vb = na.VertexBatch(n)
    vb.stid = torch.from_numpy(stid).cuda()
    t = dict(q=torch.from_numpy(q).cuda(), spec=torch.from_numpy(np.concatenate([spec, [0]])).cuda(),
             n_spec=torch.tensor([n_spec], dtype=torch.int32, device="cuda"))
    return vb, t, cap, keep_split


def _run(torch, ctx, cfg, vb, t, cap, keep_split, static):
    n = vb.n
    out = dict(qn=torch.full((n,), -1.0, device="cuda"), split=torch.full((n,), 255, dtype=torch.uint8, device="cuda"),
               surv=torch.full((cap,), -7, dtype=torch.int32, device="cuda"),
               ns=torch.zeros(1, dtype=torch.int32, device="cuda"), ovf=torch.zeros(1, dtype=torch.int32, device="cuda"),
               qn_stid=torch.zeros(4 * n, device="cuda"))
    args = ctx.decide_args(vb, None, t["q"], out["surv"], out["ns"], out["ovf"], spec=t["spec"], n_spec=t["n_spec"],
                           seed=0x5EED, stream_id=3, qn=out["qn"], split=out["split"] if keep_split else None,
                           qn_stid=out["qn_stid"])
    st = na.rrs_decide_static(cfg, args) if static else ctx.rrs_decide(args)
    assert st == na.OK
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in out.items()}


@pytest.mark.parametrize("keep_split", [True, False])
def test_static_decide_matches_context_decide(torch_cuda, keep_split):
    torch = torch_cuda
    cfg = na.default_config(n_pixels=1 << 17, max_queue=1 << 17)
    ctx = na.Context(cfg)
    for n, n_spec, seed in ((100000, 37, 1), (3001, 5, 2)):  # the second call is smaller
        vb, t, cap, _ = _case(torch, n, n_spec, seed, keep_split)
        ref = _run(torch, ctx, cfg, vb, t, cap, keep_split, static=False)
        got = _run(torch, ctx, cfg, vb, t, cap, keep_split, static=True)
        k = int(ref["ns"][0])
        print(f"n {n}: {k} survivors, overflow {int(ref['ovf'][0])}")
        assert k == int(got["ns"][0]) and int(ref["ovf"][0]) == int(got["ovf"][0])
        assert np.array_equal(ref["surv"][:k], got["surv"][:k])
        for key in ("qn", "qn_stid") + (("split",) if keep_split else ()):
            assert np.array_equal(ref[key], got[key]), key


def test_static_decide_refuses_oversize(torch_cuda):
    torch = torch_cuda
    cfg = na.default_config()
    n = na.STATIC_DECIDE_MAX + 1
    vb = na.VertexBatch(1)
    vb.n = n
    q = torch.zeros(1, device="cuda")
    surv = torch.zeros(1, dtype=torch.int32, device="cuda")
    ns = torch.zeros(1, dtype=torch.int32, device="cuda")
    a = na.DecideArgs()
    a.n = n
    a.v = vb.c_struct()
    a.q_dev = q.data_ptr()
    a.survivors_dev = surv.data_ptr()
    a.capacity = 1
    a.n_survivors_dev = ns.data_ptr()
    assert na.rrs_decide_static(cfg, a) == na.ERR_SHAPE
