"""Path nodes and training samples on the GPU against the oracle: a forest of paths recorded
over four bounces (splits give siblings), radiance added along the paths, training rows
built with each option; then the rows feed a training step with the device error sum."""
import numpy as np
import pytest

import nrrs_amd as na

pytestmark = pytest.mark.gpu

P = 4096


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    assert torch.cuda.is_available()
    return torch


def _forest(torch, ora, ctx, nb, seed, bounces=4, n_dev_cut=None):
    """Records the same forest on the device and in the oracle; returns the node ids."""
    rng = np.random.default_rng(seed)
    prev = None
    all_nodes = []
    for b in range(bounces):
        m = P if b == 0 else int(0.8 * len(prev))
        vb = na.synth_vertices(m, seed * 100 + b, P)
        hv = ora.HostVertices(vb.to_numpy())
        parent = None if b == 0 else rng.choice(prev, m).astype(np.int32)   # repeats: split paths
        q = rng.uniform(0.05, 4.0, m).astype(np.float32)
        kw = {}
        if n_dev_cut is not None and b == 1:
            kw["n_dev"] = torch.tensor([m - n_dev_cut], dtype=torch.int32, device="cuda")
        g = ctx.nodes_record(vb, None if parent is None else torch.from_numpy(parent).cuda(),
                             torch.from_numpy(q).cuda(), depth=b, **kw).cpu().numpy()
        if "n_dev" in kw:
            m = m - n_dev_cut
            hv = ora.HostVertices({k: (v[..., :m] if v is not None else None) for k, v in vb.to_numpy().items()})
            parent = None if parent is None else parent[:m]
            q = q[:m]
            g = g[:m]
        o = nb.record(hv, parent, q, depth=b)
        assert np.array_equal(g, o), f"bounce {b}: node ids"
        # radiance found after this bounce at a random half of the new vertices
        sel = rng.uniform(size=m) < 0.5
        L = rng.lognormal(-1, 1, (3, int(sel.sum()))).astype(np.float32)
        ctx.nodes_add_radiance(torch.from_numpy(g[sel].copy()).cuda(), torch.from_numpy(L).cuda())
        nb.add_radiance(g[sel], L)
        prev = g[g >= 0]
        all_nodes.append(g)
    return all_nodes


@pytest.mark.parametrize("max_nodes", [1 << 16, 9000])
def test_nodes_and_training_rows_match_oracle(torch_cuda, ora, max_nodes):
    torch = torch_cuda
    ctx = na.Context(n_pixels=P, max_queue=P, max_train=P, max_nodes=max_nodes)
    ctx.nodes_reset()
    nb = ora.Nodes(max_nodes)
    _forest(torch, ora, ctx, nb, seed=3, n_dev_cut=17)
    size, inval = ctx.nodes_size()
    assert (size, inval) == (nb.size, nb.invalid)
    print(f"nodes {size}, invalid {inval}")
    rng = np.random.default_rng(5)
    film_np = rng.lognormal(-1, 1, (3, P)).astype(np.float32)
    film = torch.from_numpy(film_np).cuda()
    err_np = rng.uniform(0, 3, (2, P // 4)).astype(np.float32)   # error image at half resolution (64 x 64 -> 32 x 32)
    err = torch.from_numpy(err_np).cuda()
    for opts in (dict(), dict(ignore_zero_L=True), dict(copy_rrs_node=True, mark_duplicates=False)):
        cap = 12000
        vb, o = ctx.generate_training_data(film, cap, error=err, error_scale=2, frame_width=64, **opts)
        torch.cuda.synchronize()
        r = nb.training_data(film_np, cap, err=err_np, scale=2, width=64,
                             ignore_zero_L=opts.get("ignore_zero_L", False), copy_rrs=opts.get("copy_rrs_node", False),
                             mark_dup=opts.get("mark_duplicates", True))
        n = int(o["n_out"].item())
        assert n == r["n"], opts
        g = vb.to_numpy()
        for k in na.VertexBatch.FIELDS3 + ("rough", "pixel"):
            assert np.array_equal(g[k][..., :n], r["v"][k][..., :n]), (k, opts)
        for k in ("q_target", "error", "error_avg"):
            assert np.array_equal(o[k][:n].cpu().numpy(), r[k][:n]), (k, opts)
        assert np.array_equal(o["pixel_samples"].cpu().numpy().astype(np.uint32), r["pixel_samples"])
        # FP32 sums of the same terms in another order / FP32 division: relative 1e-5
        tl = o["target_L"][:, :n].cpu().numpy()
        assert np.allclose(tl, r["target_L"][:, :n], rtol=1e-5, atol=1e-30), opts
        assert np.allclose(o["ref_mean"][:n].cpu().numpy(), r["ref_mean"][:n], rtol=1e-6)
        assert np.allclose(o["sample_weight"][:n].cpu().numpy(), r["sample_weight"][:n], rtol=1e-7)
        print(opts, "rows", n, "duplicates", int((r["q_target"][:n] < 0).sum()))


def test_rows_feed_a_training_step(torch_cuda, ora):
    """Nodes -> rows -> error pipeline -> joint training step, all on the device."""
    torch = torch_cuda
    ctx = na.Context(precision=na.BF16, n_pixels=P, max_queue=P, max_train=1 << 14, max_nodes=1 << 15)
    ctx.nodes_reset()
    nb = ora.Nodes(1 << 15)
    _forest(torch, ora, ctx, nb, seed=8)
    rng = np.random.default_rng(6)
    film = torch.from_numpy(rng.lognormal(-1, 1, (3, P)).astype(np.float32)).cuda()
    wl = film * torch.from_numpy(rng.lognormal(0, 0.3, (3, P)).astype(np.float32)).cuda()
    spp = torch.from_numpy(rng.integers(1, 8, P).astype(np.float32)).cuda()
    err, esum = ctx.error_compute(film, wl, spp)
    vb, o = ctx.generate_training_data(film, 1 << 14, error=err, frame_width=64)
    n = int(o["n_out"].item())
    vb.n = n
    side = ctx.side(ref_mean=o["ref_mean"], error=o["error"], sample_weight=o["sample_weight"],
                    q_target=o["q_target"], error_sum_dev=esum)
    loss = torch.zeros(2, device="cuda")
    for _ in range(3):
        na.check(ctx.train_step(vb, 3, film=film, target_L=o["target_L"], side=side, loss=loss), "train_step")
    torch.cuda.synchronize()
    lv = loss.cpu().numpy()
    print("rows", n, "losses", lv)
    assert n > 0 and np.all(np.isfinite(lv)) and lv[0] > 0
