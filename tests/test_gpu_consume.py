"""The survivor consumer on the GPU: the RRS stage of one bounce over a scatter queue with
specular items, q' stored per scatter item, then the next bounce's rows gathered from the
survivor queue with thp / q' (integrator.cpp:772-794), bit-exact against the oracle."""
import numpy as np
import pytest

import nrrs_amd as na

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    assert torch.cuda.is_available()
    return torch


@pytest.mark.parametrize("n,n_spec,cap_frac", [(4096, 0, 8.0), (4096, 300, 8.0), (1000, 50, 0.5), (0, 7, 1.0)])
def test_consume_matches_oracle(torch_cuda, ora, n, n_spec, cap_frac):
    torch = torch_cuda
    P = 4096
    ctx = na.Context(precision=na.BF16, n_pixels=P, max_queue=max(n, 1))
    rng = np.random.default_rng(n + n_spec)
    n_sc = n + n_spec                                   # scatter queue: inference items and specular items
    sc = na.synth_vertices(max(n_sc, 1), 31, P)
    sc.n = n_sc
    perm = rng.permutation(n_sc).astype(np.int32)
    inf_stid, spec_stid = perm[:n], perm[n:]
    # inference queue = the scatter items at inf_stid, carrying their stid
    d = sc.to_numpy()
    dq = {k: (v[..., inf_stid] if v is not None else None) for k, v in d.items()}
    dq["stid"] = inf_stid.astype(np.uint32)
    vq = na.VertexBatch.from_numpy(dq) if n else na.VertexBatch(0)
    vq.n = n
    film_np = rng.lognormal(-1, 1, (3, P)).astype(np.float32)
    film = torch.from_numpy(film_np).cuda()
    cap = max(int(cap_frac * max(n_sc, 1)), 1)
    q = torch.empty(max(n, 1), device="cuda")
    surv = torch.full((cap,), -1, dtype=torch.int32, device="cuda")
    n_surv = torch.zeros(1, dtype=torch.int32, device="cuda")
    rrs = torch.ones(max(n_sc, 1), device="cuda")    # specular items keep q' = 1
    spec = torch.from_numpy(spec_stid.copy()).cuda() if n_spec else None
    nsp = torch.tensor([n_spec], dtype=torch.int32, device="cuda") if n_spec else None
    args = ctx.decide_args(vq, film, q, surv, n_surv, spec=spec, n_spec=nsp, seed=9, stream_id=2, qn_stid=rrs)
    assert ctx.rrs_infer_decide(args) == na.OK
    node_in = torch.from_numpy(rng.integers(-1, 1000, max(n_sc, 1)).astype(np.int32)).cuda()
    out, node_out, stid_out = ctx.consume_survivors(surv, n_surv, sc, rrs, cap, node_in=node_in)
    torch.cuda.synchronize()
    ns = min(int(n_surv.item()), cap)
    s_np = surv[:ns].cpu().numpy().astype(np.uint32)
    # q' landed at the inference items' stids only
    rrs_np = rrs[:n_sc].cpu().numpy()
    assert np.all(rrs_np[spec_stid] == 1.0)
    ref_o, ref_node = ora.consume_survivors(s_np, ora.HostVertices(d if n_sc else sc.to_numpy()), rrs_np,
                                            node_in[:n_sc].cpu().numpy())
    g = out.to_numpy()
    for k in na.VertexBatch.FIELDS3 + ("rough", "pixel"):
        assert np.array_equal(g[k][..., :ns], ref_o[k][..., :ns]), k
    assert np.array_equal(node_out[:ns].cpu().numpy(), ref_node[:ns])
    assert np.array_equal(stid_out[:ns].cpu().numpy().astype(np.uint32), s_np)
    # each inference item appears s times; the thp sum over its copies is s / q' times its thp
    if n:
        counts = np.bincount(s_np, minlength=n_sc)
        assert np.all(counts[spec_stid] == 1)
        print(f"survivors {ns}: mean copies / q' = {np.mean(counts[inf_stid] / rrs_np[inf_stid]):.4f}")
