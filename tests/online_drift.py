"""A reduced online loop trained twice on one seed stream: on the device through the C ABI and
by the FP64 oracle, from the same initial parameters (the reference's loop:
integrator.cpp:2337-2657, one NRRSNetwork::train per frame, network.cpp:196-295).

Per frame f, both sides take the same inputs: `n` training vertices from the synthetic recipe
with seed SEED + f, log-normal(-1, 1.5) radiance targets with a 5% tail scaled x100, and
per-sample pixel errors, sample weights and q targets (10% of them -1, duplicate nodes) drawn
from a generator seeded with f. The device runs one persistent multi-bounce frame launch (FP16/BF16 contexts; the frame
kernel has no FP32 form) and then `nrrs_train_step`: StatNet alone for the first `warm` frames,
StatNet + RRSNet afterwards. The oracle runs the same composition in FP64: StatNet statistics
from its forward with the parameters before the step, both nets' gradients (NRRS_LL2, then
NRRS_RRS on those statistics), then Adam with its own step count per net.

After every frame q of a fixed probe set (the chained StatNet -> RRSNet stage, q before
normalisation) is evaluated by each side on its own parameters; the drift at frame f is the
relative difference |q_dev - q_fp64| / q_fp64 over the probe set. Frame -1 is the untrained
start, where both sides hold the same parameters: its drift is the inference error alone.

Test infrastructure: loaded by tests/test_gpu_online.py and by the online report of bench.py."""
import numpy as np

SEED = 2024
PROBE_SEED = 424242


def _film(n_pixels):
    return np.random.default_rng(SEED).lognormal(-1.0, 1.0, (3, n_pixels)).astype(np.float32)


def _frame_inputs(f, n):
    """Radiance targets [3][n] and the RRS side arrays of frame f."""
    rng = np.random.default_rng(SEED + 31 * f)
    L = rng.lognormal(-1.0, 1.5, (3, n))
    L[:, rng.uniform(size=n) < 0.05] *= 100.0
    qt = rng.uniform(0.05, 5.0, n).astype(np.float32)
    qt[rng.uniform(size=n) < 0.1] = -1.0
    side = dict(error=rng.uniform(0, 2, n).astype(np.float32),
                sample_weight=(1.0 / rng.integers(1, 4, n)).astype(np.float32), q_target=qt)
    return L.astype(np.float32), side


def config(na, prec, n, n_pixels):
    return na.default_config(precision=prec, n_pixels=n_pixels, max_queue=2 * n_pixels, max_train=n,
                             lr_rrsnet=1e-3)


def initial_params(na, cfg):
    """The parameters a fresh context of this configuration starts from (its seeded init)."""
    ctx = na.Context(cfg)
    p = [ctx.params_get(net).astype(np.float64) for net in (na.STATNET, na.RRSNET)]
    ctx.close()
    return p


def _q_fp64(na, ora, cfg, hv, film_np, ps, pr):
    stats = ora.net_forward(cfg, na.STATNET, hv, film_np, None, ps).T.copy()
    raw = ora.net_forward(cfg, na.RRSNET, hv, film_np, stats, pr)[:, 0]
    return ora.q_f64(cfg, na.Q_RRSNET_CHAINED, hv, raw, film_np)


def oracle_trajectory(na, ora, cfg, p0, frames, warm, n, n_probe):
    """FP64 q of the probe set at the start and after every frame's training step, and the
    batch losses of every step ([frames][2], NaN where a net did not train)."""
    film_np = _film(cfg.n_pixels)
    hp = ora.HostVertices(ora.synth_vertices(n_probe, PROBE_SEED, cfg.n_pixels))
    ps, pr = p0[0].copy(), p0[1].copy()
    ms, vs, mr, vr = (np.zeros_like(ps), np.zeros_like(ps), np.zeros_like(pr), np.zeros_like(pr))
    qs = [_q_fp64(na, ora, cfg, hp, film_np, ps, pr)]
    losses = np.full((frames, 2), np.nan)
    t_rrs = 0
    error_sum = 0.7 * cfg.n_pixels
    for f in range(frames):
        hv = ora.HostVertices(ora.synth_vertices(n, SEED + f, cfg.n_pixels))
        tL, side = _frame_inputs(f, n)
        stats = ora.net_forward(cfg, na.STATNET, hv, film_np, None, ps).T.copy() if f >= warm else None
        l0, g0, _ = ora.train_grads(cfg, na.STATNET, hv, ps, na.LOSS_LL2, film=film_np, target_L=tL)
        ora.adam(ps, ms, vs, g0, cfg.lr_statnet, cfg.adam_beta1, cfg.adam_beta2, cfg.adam_eps, cfg.adam_l2_reg,
                 f + 1)
        losses[f, 0] = l0
        if f >= warm:
            rs = ora.RrsSide(None, side["error"], side["sample_weight"], side["q_target"], error_sum)
            l1, g1, _ = ora.train_grads(cfg, na.RRSNET, hv, pr, na.LOSS_RRS, film=film_np, stats=stats, side=rs)
            t_rrs += 1
            ora.adam(pr, mr, vr, g1, cfg.lr_rrsnet, cfg.adam_beta1, cfg.adam_beta2, cfg.adam_eps,
                     cfg.adam_l2_reg, t_rrs)
            losses[f, 1] = l1
        qs.append(_q_fp64(na, ora, cfg, hp, film_np, ps, pr))
    return np.array(qs), losses


def device_trajectory(torch, na, cfg, p0, frames, warm, n, n_probe, bounces=4):
    """Device q of the probe set at the start and after every frame, and the batch losses
    ([frames][2], NaN where a net did not train)."""
    ctx = na.Context(cfg)
    for net in (na.STATNET, na.RRSNET):
        ctx.params_set(net, p0[net].astype(np.float32))
    P = cfg.n_pixels
    film = torch.from_numpy(_film(P)).cuda()
    frame = None
    if cfg.precision != na.FP32:
        a, bufs = ctx.frame_buffers(P, bounces, capacity=2 * P)
        a.film_dev = film.data_ptr()
        frame = (a, bufs)
    probe = na.synth_vertices(n_probe, PROBE_SEED, P)
    dev = dict(device="cuda")
    qb = dict(q=torch.zeros(n_probe, dtype=torch.float32, **dev), qn=torch.zeros(n_probe, dtype=torch.float32, **dev),
              split=torch.zeros(n_probe, dtype=torch.uint8, **dev),
              survivors=torch.zeros(8 * n_probe, dtype=torch.int32, **dev),
              n_surv=torch.zeros(1, dtype=torch.int32, **dev), overflow=torch.zeros(1, dtype=torch.int32, **dev))

    def q_probe():
        args = ctx.decide_args(probe, film, qb["q"], qb["survivors"], qb["n_surv"], qb["overflow"],
                               q_mode=na.Q_RRSNET_CHAINED, seed=1, stream_id=0, qn=qb["qn"], split=qb["split"])
        na.check(ctx.rrs_infer_decide(args), "rrs_infer_decide")
        torch.cuda.synchronize()
        return qb["q"].cpu().numpy().astype(np.float64)

    qs = [q_probe()]
    loss = torch.zeros(2, dtype=torch.float32, **dev)
    losses = np.full((frames, 2), np.nan)
    for f in range(frames):
        if frame is not None:
            frame[0].seed = 77 + f
            na.check(ctx.frame_run(frame[0]), "frame_run")
        vb = na.synth_vertices(n, SEED + f, P)
        tL, side = _frame_inputs(f, n)
        # the side struct holds device pointers only: the tensors stay referenced until the step ends
        keep = {k: torch.from_numpy(v).cuda() for k, v in side.items()}
        keep["L"] = torch.from_numpy(tL).cuda()
        sd = ctx.side(error=keep["error"], sample_weight=keep["sample_weight"], q_target=keep["q_target"],
                      error_sum=0.7 * P)
        mask = 1 if f < warm else 3
        na.check(ctx.train_step(vb, mask, film=film, target_L=keep["L"], side=sd, loss=loss), "train_step")
        torch.cuda.synchronize()
        l = loss.cpu().numpy()
        losses[f, 0] = l[0]
        if mask == 3:
            losses[f, 1] = l[1]
        qs.append(q_probe())
    status = 0 if frame is None else int(frame[1]["status"].item())
    ctx.close()
    return np.array(qs), losses, status


def drift_curve(q_dev, q_ref):
    """Per frame (row -1 first): mean, 99th percentile and max of |q_dev - q_ref| / q_ref."""
    r = np.abs(q_dev - q_ref) / q_ref
    return np.stack([r.mean(axis=1), np.percentile(r, 99, axis=1), r.max(axis=1)], axis=1)
