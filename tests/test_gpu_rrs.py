"""Parity of the HIP inference + RRS decision path with the CPU oracle.

Integer results (hash-grid indices, split counts, survivor queue, its order and
length) must be bit-exact. Floating-point outputs of the FP32 path must match
the FP64 oracle within the tolerances stated in each test.
"""
import numpy as np
import pytest

import nrrs_amd as na

pytestmark = pytest.mark.gpu

N_SMALL = 1 << 14  # the 128x128-pixel workload, one vertex per pixel


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a device"
    return torch


def random_params(ctx, net, seed, grid_amp=0.5):
    """Master parameters with a grid large enough to matter and Xavier MLP."""
    p = ctx.params_get(net)
    g, m = ctx.param_count(net)
    rng = np.random.default_rng(seed)
    p[:g] = rng.uniform(-grid_amp, grid_amp, g).astype(np.float32)
    ctx.params_set(net, p)
    return p


def make_film(torch, n_pixels, seed):
    rng = np.random.default_rng(seed)
    film = rng.lognormal(-1.0, 1.0, (3, n_pixels)).astype(np.float32)
    return film, torch.from_numpy(film).cuda()


def test_synth_matches_oracle_recipe(torch_cuda, ora):
    vb = na.synth_vertices(4096, 1234, 4096)
    torch_cuda.cuda.synchronize()
    g = vb.to_numpy()
    h = ora.synth_vertices(4096, 1234, 4096)
    assert np.array_equal(g["pixel"], h["pixel"])
    for k in na.VertexBatch.FIELDS3 + ("rough",):
        assert np.allclose(g[k], h[k], rtol=2e-6, atol=2e-6), k
    assert np.allclose(np.linalg.norm(g["dir"], axis=0), 1.0, atol=1e-5)
    assert g["rough"].min() >= 0.05 and g["rough"].max() <= 1.0


def test_grid_indices_bit_exact(torch_cuda, ora):
    ctx = na.Context()
    vb = na.synth_vertices(N_SMALL, 7, N_SMALL)
    d = vb.to_numpy()
    # edge positions: AABB corners, outside the box, NaN
    edge = np.array([[0, 0, 0], [1, 1, 1], [-5, 7, 0.5], [0.999999, 1e-7, 0.5], [np.nan, 0.2, 0.3]],
                    np.float32).T
    d["pos"][:, : edge.shape[1]] = edge
    vb = na.VertexBatch.from_numpy(d)
    idx = ctx.grid_indices(vb).cpu().numpy().astype(np.uint32)
    ref = ora.grid_indices(ctx.cfg, ora.HostVertices(d))
    assert np.array_equal(idx, ref)


@pytest.mark.parametrize("net", [na.STATNET, na.RRSNET])
def test_encode_matches_oracle(torch_cuda, ora, net):
    ctx = na.Context()
    p = random_params(ctx, net, 3)
    vb = na.synth_vertices(4096, 21, 4096)
    film_np, film = make_film(torch_cuda, ctx.cfg.n_pixels, 5)
    enc = ctx.encode(net, vb, film).cpu().numpy()
    ref = ora.encode(ctx.cfg, net, ora.HostVertices(vb.to_numpy()), film_np, None, p.astype(np.float64))
    # FP32 trilinear/SH arithmetic against FP64: absolute 2e-6 on O(1) features
    assert np.allclose(enc, ref, rtol=1e-5, atol=2e-6)


def _rel_err(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30)


# Output error relative to the output range, per precision of the MLP contractions:
# FP32 vector path, FP16 and BF16 MFMA with FP32 accumulation (8 / 11-bit mantissas).
OUT_TOL = {na.FP32: 1e-5, na.FP16: 5e-3, na.BF16: 4e-2}
# Relative error of q = activationSigma(raw) against FP64 (the north-star bound is 1e-2), and
# of the StatNet/ADRRS q relative to its range.
Q_TOL = {na.FP32: 1e-4, na.FP16: 5e-3, na.BF16: 1e-2}
Q_TOL_ADRRS = {na.FP32: 1e-4, na.FP16: 5e-3, na.BF16: 3e-2}
PRECISIONS = [na.FP32, na.FP16, na.BF16]


@pytest.mark.parametrize("prec", PRECISIONS)
def test_statnet_rrsnet_forward_match_fp64(torch_cuda, ora, prec):
    ctx = na.Context(precision=prec)
    ps = random_params(ctx, na.STATNET, 11)
    pr = random_params(ctx, na.RRSNET, 12)
    vb = na.synth_vertices(N_SMALL, 31, N_SMALL)
    film_np, film = make_film(torch_cuda, ctx.cfg.n_pixels, 6)
    stats = ctx.statnet_infer(vb)
    raw = ctx.rrsnet_infer(vb, film)
    raw_chained = ctx.rrsnet_infer(vb, film, stats)
    hv = ora.HostVertices(vb.to_numpy())
    ref_stats = ora.net_forward(ctx.cfg, na.STATNET, hv, film_np, None, ps.astype(np.float64)).T
    ref_raw = ora.net_forward(ctx.cfg, na.RRSNET, hv, film_np, None, pr.astype(np.float64))[:, 0]
    # chained: the oracle consumes the GPU stats so only the RRSNet error is measured
    gstats = stats.cpu().numpy()
    ref_ch = ora.net_forward(ctx.cfg, na.RRSNET, hv, film_np, gstats.astype(np.float64),
                             pr.astype(np.float64))[:, 0]
    errs = (_rel_err(gstats, ref_stats), _rel_err(raw.cpu().numpy(), ref_raw),
            _rel_err(raw_chained.cpu().numpy(), ref_ch))
    print(f"precision {prec}: statnet {errs[0]:.2e} rrsnet {errs[1]:.2e} chained {errs[2]:.2e}")
    assert max(errs) < OUT_TOL[prec]


def _decide_buffers(torch, n, capacity):
    return dict(
        q=torch.empty(max(n, 1), dtype=torch.float32, device="cuda"),
        qn=torch.empty(max(n, 1), dtype=torch.float32, device="cuda"),
        split=torch.empty(max(n, 1), dtype=torch.uint8, device="cuda"),
        survivors=torch.full((capacity,), 0xFFFFFFF, dtype=torch.int32, device="cuda"),
        n_surv=torch.zeros(1, dtype=torch.int32, device="cuda"),
        overflow=torch.zeros(1, dtype=torch.int32, device="cuda"),
    )


@pytest.mark.parametrize("n,n_spec,capacity_frac", [
    (N_SMALL, 0, 8.0), (N_SMALL, 100, 8.0), (1, 0, 8.0), (257, 3, 8.0), (1000, 0, 0.5), (0, 5, 1.0),
    (100000, 0, 8.0),
])
def test_decide_bit_exact(torch_cuda, ora, n, n_spec, capacity_frac):
    torch = torch_cuda
    cfg = na.default_config(n_pixels=max(n, 16), max_queue=max(n, 1))
    ctx = na.Context(cfg)
    rng = np.random.default_rng(n + n_spec)
    q_np = rng.uniform(0.05, 20.0, n).astype(np.float32)
    q_np[: n // 10] = rng.uniform(0.05, 1.0, n // 10).astype(np.float32)
    stid_np = rng.permutation(max(4 * n, 8))[:n].astype(np.uint32)
    spec_np = (np.arange(n_spec) + 10 * n + 5).astype(np.uint32)
    capacity = max(int(capacity_frac * (n + n_spec)), 1)
    b = _decide_buffers(torch, n, capacity)
    b["q"][:n] = torch.from_numpy(q_np).cuda()
    d = {k: np.zeros((3, n), np.float32) for k in na.VertexBatch.FIELDS3}
    d.update(rough=np.zeros(n, np.float32), pixel=np.zeros(n, np.uint32), stid=stid_np)
    vb = na.VertexBatch.from_numpy(d) if n else na.VertexBatch(0)
    vb.n = n
    if n:
        vb.stid = torch.from_numpy(stid_np.astype(np.int32)).cuda()
    spec = torch.from_numpy(np.concatenate([spec_np, [0]]).astype(np.int32)).cuda()
    nspec = torch.tensor([n_spec], dtype=torch.int32, device="cuda")
    args = ctx.decide_args(vb, None, b["q"], b["survivors"], b["n_surv"], b["overflow"], spec=spec,
                           n_spec=nspec, seed=0xABCDEF, stream_id=4, qn=b["qn"], split=b["split"])
    assert ctx.rrs_decide(args) == na.OK
    torch.cuda.synchronize()
    ref = ora.rrs_decide(cfg, q_np, stid=stid_np, spec=spec_np, seed=0xABCDEF, stream=4, capacity=capacity)
    ns = int(b["n_surv"].item())
    assert ns == ref["n_survivors"]
    assert int(b["overflow"].item()) == ref["overflow"]
    assert np.array_equal(b["survivors"][:ns].cpu().numpy().astype(np.uint32), ref["survivors"])
    if n:
        assert np.array_equal(b["split"][:n].cpu().numpy(), ref["s"])
        assert np.array_equal(b["qn"][:n].cpu().numpy(), ref["qn"])


@pytest.mark.parametrize("prec", PRECISIONS)
@pytest.mark.parametrize("mode", [na.Q_RRSNET, na.Q_RRSNET_CHAINED, na.Q_STATNET_ADRRS])
def test_infer_decide_pipeline(torch_cuda, ora, mode, prec):
    """Full stage: q within tolerance of FP64, decision bit-exact given the device q."""
    torch = torch_cuda
    n = N_SMALL
    ctx = na.Context(precision=prec)
    ps = random_params(ctx, na.STATNET, 41)
    pr = random_params(ctx, na.RRSNET, 42)
    vb = na.synth_vertices(n, 77, ctx.cfg.n_pixels)
    film_np, film = make_film(torch, ctx.cfg.n_pixels, 8)
    b = _decide_buffers(torch, n, 8 * n)
    args = ctx.decide_args(vb, film, b["q"], b["survivors"], b["n_surv"], b["overflow"], q_mode=mode,
                           seed=5, stream_id=1, qn=b["qn"], split=b["split"])
    assert ctx.rrs_infer_decide(args) == na.OK
    torch.cuda.synchronize()
    q = b["q"][:n].cpu().numpy()
    hv = ora.HostVertices(vb.to_numpy())
    # FP64 q of the same stage
    if mode == na.Q_STATNET_ADRRS:
        src = ora.net_forward(ctx.cfg, na.STATNET, hv, film_np, None, ps.astype(np.float64)).T
    else:
        stats = None
        if mode == na.Q_RRSNET_CHAINED:
            stats = ora.net_forward(ctx.cfg, na.STATNET, hv, film_np, None, ps.astype(np.float64)).T
        src = ora.net_forward(ctx.cfg, na.RRSNET, hv, film_np, stats, pr.astype(np.float64))[:, 0]
    q64 = ora.q_f64(ctx.cfg, mode, hv, src, film_np)
    qerr = np.abs(q - q64) / q64
    print(f"precision {prec} mode {mode}: q rel err max {qerr.max():.2e} mean {qerr.mean():.2e}")
    if mode == na.Q_STATNET_ADRRS:
        # q = mean(clamp(L) thp / I): relative error of L near 0 is unbounded in low
        # precision, so the bound is on the error against the q range, and the mean
        assert np.max(np.abs(q - q64)) / q64.max() < Q_TOL_ADRRS[prec] and qerr.mean() < Q_TOL_ADRRS[prec]
    else:
        assert qerr.max() < Q_TOL[prec]
    # the decision from the device's own FP32 q is bit-exact
    ref = ora.rrs_decide(ctx.cfg, q, stid=None, seed=5, stream=1)
    ns = int(b["n_surv"].item())
    assert ns == ref["n_survivors"] and int(b["overflow"].item()) == 0
    assert np.array_equal(b["split"][:n].cpu().numpy(), ref["s"])
    assert np.array_equal(b["qn"][:n].cpu().numpy(), ref["qn"])
    assert np.array_equal(b["survivors"][:ns].cpu().numpy().astype(np.uint32), ref["survivors"])
    # decisions that would flip under the FP64 q (reported, tiny)
    ref64 = ora.rrs_decide(ctx.cfg, q64.astype(np.float32), seed=5, stream=1)
    flips = int(np.sum(ref64["s"] != ref["s"]))
    print(f"mode {mode}: {flips} of {n} split counts differ under the FP64 q")
    assert flips <= n * Q_TOL_ADRRS[prec] + 2


@pytest.mark.parametrize("prec", [na.FP32, na.BF16])
@pytest.mark.parametrize("knobs", [
    dict(rrs_weight_window=1),                    # RRSWeightWindow (commoncudautils.h:14-28)
    dict(rrs_use_ceil=0),                         # normaliseFactor * rate (integrator.cpp:1376-1377)
    dict(rrs_no_upscale=0),                       # q' = q * nf also when nf > 1 (integrator.cpp:1391-1397)
    dict(rrs_weight_window=1, rrs_use_ceil=0, rrs_no_upscale=0, rrs_normalize_rate=0.7),
])
def test_decision_branches(torch_cuda, ora, prec, knobs):
    """The knobs of the RRS decision that the default configuration leaves off, through the full
    stage: q against FP64, then the decision bit-exact against the oracle on the device's own q.
    The film has four pixels per vertex, so the normalisation factor exceeds 1 and the
    no-upscale switch decides whether q' grows."""
    torch = torch_cuda
    n, P = N_SMALL, 4 * N_SMALL
    ctx = na.Context(na.default_config(precision=prec, n_pixels=P, max_queue=n, **knobs))
    ps = random_params(ctx, na.STATNET, 51)
    pr = random_params(ctx, na.RRSNET, 52)
    vb = na.synth_vertices(n, 78, P)
    film_np, film = make_film(torch, P, 9)
    b = _decide_buffers(torch, n, 8 * n)
    args = ctx.decide_args(vb, film, b["q"], b["survivors"], b["n_surv"], b["overflow"], q_mode=na.Q_RRSNET_CHAINED,
                           seed=6, stream_id=2, qn=b["qn"], split=b["split"])
    assert ctx.rrs_infer_decide(args) == na.OK
    torch.cuda.synchronize()
    q = b["q"][:n].cpu().numpy()
    hv = ora.HostVertices(vb.to_numpy())
    stats = ora.net_forward(ctx.cfg, na.STATNET, hv, film_np, None, ps.astype(np.float64)).T
    raw = ora.net_forward(ctx.cfg, na.RRSNET, hv, film_np, stats, pr.astype(np.float64))[:, 0]
    q64 = ora.q_f64(ctx.cfg, na.Q_RRSNET_CHAINED, hv, raw, film_np)
    qerr = np.abs(q - q64) / q64
    ref = ora.rrs_decide(ctx.cfg, q, seed=6, stream=2)
    rate = 1.0 if ctx.cfg.rrs_use_ceil else ctx.cfg.rrs_normalize_rate
    nf = P / ref["sum"] * rate
    print(f"knobs {knobs} prec {prec}: q rel err max {qerr.max():.2e}; nf {nf:.3f}; "
          f"{ref['n_survivors']} survivors; q' max {ref['qn'].max():.3f}")
    assert qerr.max() < Q_TOL[prec]
    if ctx.cfg.rrs_weight_window:
        # inside the window [2 / (1 + size), 2 size / (1 + size)] q becomes exactly 1
        assert np.any(q == 1.0)
    assert nf > 1.0
    if not ctx.cfg.rrs_no_upscale:
        assert np.any(ref["qn"] > q)  # upscaled
    ns = int(b["n_surv"].item())
    assert ns == ref["n_survivors"] and int(b["overflow"].item()) == 0
    assert np.array_equal(b["split"][:n].cpu().numpy(), ref["s"])
    assert np.array_equal(b["qn"][:n].cpu().numpy(), ref["qn"])
    assert np.array_equal(b["survivors"][:ns].cpu().numpy().astype(np.uint32), ref["survivors"])


def test_fp32_level_major_equals_per_vertex_gathers(torch_cuda, ora):
    """FP32 contexts gather the grid level by level into a level-major encoding by default
    (fp32_level_major = 1); the per-vertex gathers inside the network kernel (0) give the same
    bits: StatNet statistics and the chained stage's q, q', splits and survivors."""
    torch = torch_cuda
    n = 3001
    out = []
    for lm in (1, 0):
        ctx = na.Context(na.default_config(precision=na.FP32, n_pixels=n, max_queue=n, fp32_level_major=lm))
        random_params(ctx, na.STATNET, 61)
        random_params(ctx, na.RRSNET, 62)
        vb = na.synth_vertices(n, 79, n)
        film_np, film = make_film(torch, n, 10)
        b = _decide_buffers(torch, n, 8 * n)
        args = ctx.decide_args(vb, film, b["q"], b["survivors"], b["n_surv"], b["overflow"],
                               q_mode=na.Q_RRSNET_CHAINED, seed=7, stream_id=3, qn=b["qn"], split=b["split"])
        assert ctx.rrs_infer_decide(args) == na.OK
        st = ctx.statnet_infer(vb)
        torch.cuda.synchronize()
        r = {k: v.cpu().numpy() for k, v in b.items()}
        r["stats"] = st.cpu().numpy()
        out.append(r)
    for k in ("q", "qn", "split", "n_surv", "stats"):
        assert np.array_equal(out[0][k], out[1][k]), k
    ns = int(out[0]["n_surv"][0])
    assert np.array_equal(out[0]["survivors"][:ns], out[1]["survivors"][:ns])
