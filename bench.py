"""Benchmark of the neural RRS stage on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Workload: 2^20 synthetic path vertices resident in HBM (16-level hash grid with
T = 2^19, SH-4 direction/normal encodings, 64-wide StatNet 4x64->6 and RRSNet
3x64->1). One step is one online-training iteration of the stage:
  1. the RRS stage of one bounce: StatNet -> RRSNet chained inference, q, the exact
     normalisation sum, the split decision and the stable survivor compaction;
  2. one joint training step on the same 2^20 vertices: StatNet (NRRS_LL2) and
     RRSNet (NRRS_RRS) forward, backward and Adam.
`value` is RRS queries per second of that whole step in BF16 (MFMA, FP32
accumulation, FP32 master weights, dynamic loss scaling); the per-phase rates
and the same step in FP32 and FP16 are reported beside it.

`ops` times every operation of the BF16 step (the library's event pairs around each
launcher, nrrs_set_op_timing, over further timed steps) and rates each against its
roofline; `roofline` is the entry of the step's dominant kernel, the paired
level-major grid-encoding pass (k_encode_levels<BF16, true>, two launches per step).
`matched_error` compares each precision's RRS stage (StatNet statistics, q) and both
nets' gradients with the FP64 oracle on the first 2^16 vertices of the batch.
`shipped_nets` times the same step with the reference's shipped nets (nrrs_shipped_config) in
BF16 and FP32, with the BF16 q error against the FP64 oracle.
`wavefront_frame` and `online_loop` time the 10-bounce persistent frame launch and the
256-frame online loop; `online_loop.q_drift_vs_fp64` is the q drift of a reduced online loop
trained in BF16 against the FP64 oracle trained on the same seed stream.
`cpu_baseline` is the FP32 C++ port of the same step on the host's threads, with the
single-threaded FP64 oracle beside it.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nrrs-code_amd"))

N_VERTICES = 1 << 20
N_PIXELS = 1 << 20
SEED = 0x4E525253 ^ 2
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
# HBM bytes per launch of the paired encode pass at 2^20 vertices, from rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE (profiles/encode_pair_bf16_2p20_pmc.md), or None when not measured.
ENCODE_TRAFFIC_BYTES = 1204578005
MACS = {"statnet": 80 * 64 + 3 * 64 * 64 + 64 * 6, "rrsnet": 80 * 64 + 2 * 64 * 64 + 64 * 1}
# Algorithmic work of one launch of each operation of the step (bytes it must move, or MFMA
# flops), for a batch of n vertices and nets of n_grid + n_mlp parameters:
#  - encode_pair(_counts): 12 B position + 16 levels x 8 corners x 8 B (both nets' BF16 feature
#    pairs) + 2 x 64 B level-major encodings written = 1164 B per vertex;
#  - mlp_chained: StatNet and RRSNet forward, 2 x (17 792 + 13 376) flops per query;
#  - train_statnet / train_rrsnet: forward, dX and dW, 3 x 2 x MACs per sample;
#  - bin_records_pair: per vertex and level 12 B position + 2 x 4 B dL/denc (FP16 pairs) read and
#    4.5 records x (4 B key + 2 x 4 B values) written = 74 B;
#  - apply_statnet / apply_rrsnet: per vertex and level 4.5 records x (4 B key + 4 B value)
#    read = 36 B, and the FP16 table gradient read and written (4 B per grid parameter);
#  - adam_statnet / adam_rrsnet: 28 B per parameter (gradient, p, m, v read; p, m, v written).
def op_work(name, n, n_grid, n_mlp):
    levels = 16
    if name in ("encode_pair", "encode_pair_counts"):
        return "bytes", 1164 * n
    if name == "mlp_chained":
        return "flops", 2 * (MACS["statnet"] + MACS["rrsnet"]) * n
    if name in ("train_statnet", "train_rrsnet"):
        return "flops", 6 * MACS["statnet" if name == "train_statnet" else "rrsnet"] * n
    if name == "bin_records_pair":
        return "bytes", 74 * levels * n
    if name in ("apply_statnet", "apply_rrsnet"):
        return "bytes", 36 * levels * n + 4 * n_grid
    if name == "adam_statnet":
        return "bytes", 28 * (n_grid + n_mlp["statnet"])
    if name == "adam_rrsnet":
        return "bytes", 28 * (n_grid + n_mlp["rrsnet"])
    return None, 0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--cpu-sample", type=int, default=1 << 19)
    ap.add_argument("--precisions", default="bf16,fp16,fp32,fp32_plain")
    ap.add_argument("--no-frame", action="store_true", help="skip the multi-bounce wavefront and online loop")
    ap.add_argument("--no-peaks", action="store_true", help="skip the copy and GEMM measurements of the GPU")
    ap.add_argument("--no-oracle", action="store_true", help="skip the FP64-oracle error report")
    ap.add_argument("--no-shipped", action="store_true", help="skip the step with the reference's shipped nets")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="only run a 256 MiB device copy and, after 8 warm-up steps, 3 BF16 steps at 2^20 "
                         "(for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs; the profile summaries use the "
                         "dispatches of the last 3 steps)")
    ap.add_argument("--probe-shipped", action="store_true",
                    help="with --pmc-probe: the steps use the reference's shipped nets (nrrs_shipped_config)")
    return ap.parse_args()


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def measured_peaks(torch):
    """What this GPU reaches on plain library work, beside the published peaks the roofline
    fractions use: a 1 GiB device-to-device copy (read + write bytes per second) and a
    8192^3 BF16 GEMM (torch.matmul, dense TFLOP/s), each the best of 5 after a warm-up."""
    def best(fn, reps=5):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1) * 1e-3)
        return min(ts)

    src = torch.empty(1 << 28, dtype=torch.float32, device="cuda")  # 1 GiB
    src.fill_(1.0)
    dst = torch.empty_like(src)
    t_copy = best(lambda: dst.copy_(src))
    copy_gbs = 2 * src.numel() * 4 / t_copy / 1e9
    del src, dst
    n = 8192
    a = torch.randn(n, n, device="cuda").to(torch.bfloat16)
    b = torch.randn(n, n, device="cuda").to(torch.bfloat16)
    t_mm = best(lambda: torch.matmul(a, b))
    del a, b
    torch.cuda.empty_cache()
    return {"copy_GBps": round(copy_gbs, 1), "copy_frac_of_peak": round(copy_gbs / HBM_PEAK_GBS, 3),
            "bf16_gemm_TFLOPs": round(2 * n ** 3 / t_mm / 1e12, 1),
            "bf16_gemm_frac_of_peak": round(2 * n ** 3 / t_mm / 1e12 / BF16_PEAK_TFLOPS, 3)}


def cpu_baseline(sample):
    """CPU paths of the same step (chained inference + decision, then StatNet and RRSNet
    forward/backward and Adam) on the host: the FP32 multi-threaded port on `sample`
    vertices (median of 5 runs after one warm-up) and the FP64 oracle on one core over
    sample / 8 vertices (median of 3)."""
    import statistics

    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import nrrs_amd as na
    import oracle_ref as ora

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)

    def inputs(m):
        cfg = na.default_config(n_pixels=N_PIXELS, max_queue=m, max_train=m)
        hv = ora.HostVertices(ora.synth_vertices(m, SEED, N_PIXELS))
        rng = np.random.default_rng(1)
        film = rng.lognormal(-1, 1, (3, N_PIXELS)).astype(np.float32)
        ps = ora.init_params(cfg, na.STATNET, 1, 0.5)
        pr = ora.init_params(cfg, na.RRSNET, 2, 0.5)
        tL = ora.synth_targets(m, 3)
        err, sw, qt = rng.uniform(0, 2, m), np.ones(m), rng.uniform(0.05, 5, m)
        return cfg, hv, film, ps, pr, tL, err, sw, qt

    cfg, hv, film, ps, pr, tL, err, sw, qt = inputs(sample)
    runs = []
    for _ in range(6):
        r = ora.cpu_step_f32(cfg, hv, film, ps, pr, tL, err, sw, qt, sample * 0.7, threads, grads=False)
        runs.append((r["total"], r["times"][0], r["times"][1]))
    runs = runs[1:]
    med = statistics.median(t[0] for t in runs)
    t_inf = statistics.median(t[1] for t in runs)
    t_tr = statistics.median(t[2] for t in runs)

    m = max(sample // 8, 1)
    cfg, hv, film, ps, pr, tL, err, sw, qt = inputs(m)
    side = ora.RrsSide(None, err, sw, qt, m * 0.7)
    t64 = []
    for _ in range(3):
        t0 = time.perf_counter()
        stats = np.ascontiguousarray(ora.net_forward(cfg, na.STATNET, hv, film, None, ps).T)
        raw = ora.net_forward(cfg, na.RRSNET, hv, film, stats, pr)[:, 0]
        q = ora.q_f64(cfg, na.Q_RRSNET_CHAINED, hv, raw, film)
        ora.rrs_decide(cfg, q.astype(np.float32), seed=1, stream=0)
        for net, p, kind in ((na.STATNET, ps, na.LOSS_LL2), (na.RRSNET, pr, na.LOSS_RRS)):
            _, g, _ = ora.train_grads(cfg, net, hv, p, kind, film=film, stats=stats if net == na.RRSNET else None,
                                      target_L=tL, side=side)
            pp, mm, vv = p.copy(), np.zeros_like(p), np.zeros_like(p)
            ora.adam(pp, mm, vv, g, 1e-2, 0.9, 0.99, 1e-15, 0.0, 1)
        t64.append(time.perf_counter() - t0)
    med64 = statistics.median(t64)
    host = f"{os.cpu_count()} logical CPUs, {_cpu_model()}"
    return {"value": round(sample / med, 1), "unit": "queries/s", "cores": threads, "kind": "port",
            "sample": f"{sample} vertices of the same step, FP32 C++ port on {threads} threads, median of 5 runs "
                      f"after a warm-up: {med:.3f} s (inference + decision {t_inf:.3f} s, training {t_tr:.3f} s); "
                      f"host: {host}",
            # the two parts of the same runs, in the units of the GPU's rrs_queries_per_s and
            # train_samples_per_s
            "rrs_queries_per_s": round(sample / t_inf, 1), "train_samples_per_s": round(sample / t_tr, 1),
            "fp64_oracle": {"value": round(m / med64, 1), "unit": "queries/s", "cores": 1,
                            "sample": f"{m} vertices, FP64 oracle on one thread, median of 3: {med64:.2f} s"}}


class Path:
    """One precision: context, inputs and the timed step. `shipped`: the reference's shipped nets
    (nrrs_shipped_config: 8-level grid, 32-wide LeakyReLU nets, OneBlob inputs, grid-less RRSNet,
    Ema of the weights) instead of the benchmark nets."""

    def __init__(self, precision, torch, na, np, shipped=False, **cfg):
        self.torch, self.na = torch, na
        n = N_VERTICES
        self.n = n
        config = na.shipped_config if shipped else na.default_config
        self.ctx = na.Context(config(precision=precision, n_pixels=N_PIXELS, max_queue=n, max_train=n, **cfg))
        rng = np.random.default_rng(0)
        for net in (na.STATNET, na.RRSNET):
            p = self.ctx.params_get(net)
            g, _ = self.ctx.param_count(net)
            p[:g] = rng.uniform(-0.3, 0.3, g).astype(np.float32)
            self.ctx.params_set(net, p)
        self.vb = na.synth_vertices(n, SEED, N_PIXELS)
        gen = torch.Generator(device="cuda")
        gen.manual_seed(SEED)
        dev = dict(device="cuda", dtype=torch.float32)
        self.film = torch.exp(torch.randn(3, N_PIXELS, generator=gen, **dev) - 1.0)
        L = torch.exp(torch.randn(3, n, generator=gen, **dev) * 1.5 - 1.0)
        fire = torch.rand(n, generator=gen, **dev) < 0.05
        self.target_L = torch.where(fire, L * 100.0, L).contiguous()
        self.err = torch.rand(n, generator=gen, **dev) * 2.0
        self.sw = torch.ones(n, **dev)
        self.qt = torch.rand(n, generator=gen, **dev) * 5.0 + 0.05
        self.side = self.ctx.side(error=self.err, sample_weight=self.sw, q_target=self.qt, error_sum=0.7 * N_PIXELS)
        self.q = torch.empty(n, **dev)
        self.qn = torch.empty(n, **dev)
        self.split = torch.empty(n, dtype=torch.uint8, device="cuda")
        self.surv = torch.empty(8 * n, dtype=torch.int32, device="cuda")
        self.n_surv = torch.zeros(1, dtype=torch.int32, device="cuda")
        self.ovf = torch.zeros(1, dtype=torch.int32, device="cuda")
        self.loss = torch.zeros(2, **dev)
        self.args = self.ctx.decide_args(self.vb, self.film, self.q, self.surv, self.n_surv, self.ovf,
                                         q_mode=na.Q_RRSNET_CHAINED, qn=self.qn, split=self.split)

    def infer(self, k):
        self.args.pcg_stream = k
        self.na.check(self.ctx.rrs_infer_decide(self.args), "rrs_infer_decide")

    def train(self):
        self.na.check(self.ctx.train_step(self.vb, 3, film=self.film, target_L=self.target_L, side=self.side,
                                          loss=self.loss), "train_step")

    def run(self, steps, warmup):
        torch = self.torch
        stream = torch.cuda.current_stream()
        for k in range(warmup):
            self.infer(k)
            self.train()
        torch.cuda.synchronize()
        ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
        t0 = time.perf_counter()
        for k in range(steps):
            ev[k][0].record(stream)
            self.infer(warmup + k)
            ev[k][1].record(stream)
            self.train()
            ev[k][2].record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        inf = sum(e[0].elapsed_time(e[1]) for e in ev) / steps
        trn = sum(e[1].elapsed_time(e[2]) for e in ev) / steps
        return {"ms_per_step": 1e3 * wall / steps, "infer_ms": inf, "train_ms": trn,
                "rrs_queries_per_s": self.n / (inf * 1e-3), "train_samples_per_s": self.n / (trn * 1e-3),
                "step_queries_per_s": self.n * steps / wall, "loss": self.loss.cpu().tolist(),
                "survivors": int(self.n_surv.item())}

    def op_table(self, steps):
        """Mean duration (ms) per step of every operation of the step, from the library's event
        pairs around each launcher (nrrs_set_op_timing) over `steps` further steps."""
        self.ctx.set_op_timing(True)
        self.ctx.op_times()  # empty collection
        acc, launches = {}, {}
        for k in range(steps):
            self.infer(100 + k)
            self.train()
            for name, ms in self.ctx.op_times():
                acc[name] = acc.get(name, 0.0) + ms
                launches[name] = launches.get(name, 0) + 1
        self.ctx.set_op_timing(False)
        return {k: {"ms_per_step": v / steps, "launches_per_step": launches[k] / steps} for k, v in acc.items()}


def matched_errors(torch, na, np, precisions, m=1 << 16):
    """Each precision against the FP64 oracle on the first m vertices of the benchmark batch,
    with the parameters the contexts start from (the same in every precision): StatNet
    statistics (nrrs_statnet_infer) and q of the chained RRS stage (relative to the statistic
    range and to q), and both nets' gradients (nrrs_backward with NRRS_LL2, and NRRS_RRS fed with
    the oracle's statistics: norm-wise for the table, largest MLP entry error relative to the
    largest MLP gradient). The gradients are taken at a loss scale that keeps the FP16 table
    gradient finite and divided by it."""
    import copy

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ref as ora

    ora.load()
    out, ref = {}, None
    for name, prec in precisions.items():
        path = Path(prec, torch, na, np)
        ctx = path.ctx
        sub = copy.copy(path.vb)
        sub.n = m
        if ref is None:  # the oracle on this batch and these parameters (FP32 masters in FP64)
            hv = ora.HostVertices(sub.to_numpy())
            film = path.film.cpu().numpy()
            tL = path.target_L[:, :m].cpu().numpy()
            side = ora.RrsSide(None, path.err[:m].cpu().numpy(), path.sw[:m].cpu().numpy(),
                               path.qt[:m].cpu().numpy(), 0.7 * N_PIXELS)
            ps, pr = (ctx.params_get(net).astype(np.float64) for net in (na.STATNET, na.RRSNET))
            s64 = np.ascontiguousarray(ora.net_forward(ctx.cfg, na.STATNET, hv, film, None, ps).T)
            raw64 = ora.net_forward(ctx.cfg, na.RRSNET, hv, film, s64, pr)[:, 0]
            q64 = ora.q_f64(ctx.cfg, na.Q_RRSNET_CHAINED, hv, raw64, film)
            _, g0, _ = ora.train_grads(ctx.cfg, na.STATNET, hv, ps, na.LOSS_LL2, film=film, target_L=tL)
            _, g1, _ = ora.train_grads(ctx.cfg, na.RRSNET, hv, pr, na.LOSS_RRS, film=film, stats=s64, side=side)
            ref = dict(s64=s64, q64=q64, g=(g0, g1))
        q = torch.empty(m, device="cuda")
        surv = torch.empty(8 * m, dtype=torch.int32, device="cuda")
        ns = torch.zeros(1, dtype=torch.int32, device="cuda")
        args = ctx.decide_args(sub, path.film, q, surv, ns, q_mode=na.Q_RRSNET_CHAINED)
        na.check(ctx.rrs_infer_decide(args), "rrs_infer_decide")
        stats = ctx.statnet_infer(sub).cpu().numpy().astype(np.float64)
        qd = q.cpu().numpy().astype(np.float64)
        rq = np.abs(qd - ref["q64"]) / ref["q64"]
        r = {"vertices": m,
             "stats_err_max_rel_to_range": float(np.max(np.abs(stats - ref["s64"])) / np.max(np.abs(ref["s64"]))),
             "q_rel_err_max": float(rq.max()), "q_rel_err_mean": float(rq.mean()),
             "q_rel_err_p99_99": float(np.percentile(rq, 99.99))}
        stats_ref = torch.from_numpy(ref["s64"].astype(np.float32)).cuda()
        sd = ctx.side(error=path.err, sample_weight=path.sw, q_target=path.qt, error_sum=0.7 * N_PIXELS)
        for net, key in ((na.STATNET, "statnet"), (na.RRSNET, "rrsnet")):
            g_ref = ref["g"][net]
            g, _ = ctx.param_count(net)
            scale = 1.0 if prec == na.FP32 else float(2.0 ** np.clip(np.floor(np.log2(3e4 / np.abs(g_ref).max())), 0, 15))
            ctx.grads_zero(net)
            if net == na.STATNET:
                rc = ctx.backward(net, sub, na.LOSS_LL2, target_L=path.target_L, loss_scale=scale)
            else:
                rc = ctx.backward(net, sub, na.LOSS_RRS, film=path.film, stats=stats_ref, side=sd, loss_scale=scale)
            na.check(rc, "backward")
            gd = ctx.grads_get(net).astype(np.float64) / scale
            r[f"grad_{key}_grid_norm_rel_err"] = float(np.linalg.norm(gd[:g] - g_ref[:g]) / np.linalg.norm(g_ref[:g]))
            r[f"grad_{key}_mlp_max_rel_err"] = float(np.max(np.abs(gd[g:] - g_ref[g:])) / np.max(np.abs(g_ref[g:])))
        out[name] = r
        del path, ctx
        torch.cuda.empty_cache()
    return out


def shipped_nets(torch, na, np, steps, warmup, m=1 << 16):
    """The same step with the reference's shipped nets (nrrs_shipped_config) in BF16 and FP32:
    chained StatNet -> RRSNet inference (the grid-less RRSNet reads the StatNet statistics) and
    decision, then the joint training step with the Ema update; q and the StatNet statistics of
    the BF16 stage against the FP64 oracle on the first m vertices of the batch."""
    import copy

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ref as ora

    out = {}
    for name, prec in (("bf16", na.BF16), ("fp32", na.FP32)):
        path = Path(prec, torch, na, np, shipped=True)
        if name == "bf16":
            ctx = path.ctx
            sub = copy.copy(path.vb)
            sub.n = m
            hv = ora.HostVertices(sub.to_numpy())
            film = path.film.cpu().numpy()
            ps, pr = (ctx.params_get_inference(net).astype(np.float64) for net in (na.STATNET, na.RRSNET))
            s64 = np.ascontiguousarray(ora.net_forward(ctx.cfg, na.STATNET, hv, film, None, ps).T)
            raw64 = ora.net_forward(ctx.cfg, na.RRSNET, hv, film, s64, pr)[:, 0]
            q64 = ora.q_f64(ctx.cfg, na.Q_RRSNET_CHAINED, hv, raw64, film)
            q = torch.empty(m, device="cuda")
            surv = torch.empty(8 * m, dtype=torch.int32, device="cuda")
            ns = torch.zeros(1, dtype=torch.int32, device="cuda")
            na.check(ctx.rrs_infer_decide(ctx.decide_args(sub, path.film, q, surv, ns, q_mode=na.Q_RRSNET_CHAINED)),
                     "rrs_infer_decide")
            stats = ctx.statnet_infer(sub).cpu().numpy().astype(np.float64)
            rq = np.abs(q.cpu().numpy().astype(np.float64) - q64) / q64
            err = {"vertices": m, "q_rel_err_max": float(rq.max()), "q_rel_err_mean": float(rq.mean()),
                   "stats_err_max_rel_to_range": float(np.max(np.abs(stats - s64)) / np.max(np.abs(s64)))}
        r = path.run(steps if name == "bf16" else max(2, steps // 2), warmup)
        out[name] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}
        if name == "bf16":
            out[name]["vs_fp64_oracle"] = err
            out[name]["ops_ms"] = {k: round(v["ms_per_step"], 4) for k, v in path.op_table(3).items()}
        del path
        torch.cuda.empty_cache()
    out["config"] = {"grid": "8 levels x 2 features, T=2^19, base 16, scale 2.0",
                     "mlp": "StatNet 32 inputs (grid 16 + OneBlob 16) -> 32x3 LeakyReLU -> 6; RRSNet 16 inputs "
                            "(NRRSNetworkInput1, no grid) -> 32x3 LeakyReLU -> 1",
                     "optimizer": "Ema(0.95) of Adam (lr 5e-3 / 5e-4, l2 3e-6)", "vertices": N_VERTICES}
    out["bf16_over_fp32"] = {"rrs_queries": round(out["bf16"]["rrs_queries_per_s"] / out["fp32"]["rrs_queries_per_s"], 3),
                             "train_samples": round(out["bf16"]["train_samples_per_s"] /
                                                    out["fp32"]["train_samples_per_s"], 3),
                             "step": round(out["bf16"]["step_queries_per_s"] / out["fp32"]["step_queries_per_s"], 3)}
    return out


def frame_wavefront(torch, na, np, n_frames=8):
    """Batched uneven wavefront: 10 bounces from 1920x1080 primary vertices, one
    persistent launch per frame (autotuned tile shape) replayed from a hipGraph,
    against the same frame driven bounce by bounce from the host (queue size read
    back every bounce, as the reference integrator does)."""
    P = 1920 * 1080
    nb = 10
    ctx = na.Context(precision=na.BF16, n_pixels=P, max_queue=2 * P, max_train=1 << 18)
    rng = np.random.default_rng(0)
    for net in (na.STATNET, na.RRSNET):
        p = ctx.params_get(net)
        g, _ = ctx.param_count(net)
        p[:g] = rng.uniform(-0.3, 0.3, g).astype(np.float32)
        ctx.params_set(net, p)
    film = torch.from_numpy(rng.lognormal(-1, 1, (3, P)).astype(np.float32)).cuda()
    a, b = ctx.frame_buffers(P, nb, capacity=2 * P)
    a.film_dev = film.data_ptr()
    a.seed = SEED
    shape, tune_ms = ctx.frame_autotune(a)
    seed_dev = torch.tensor([SEED], dtype=torch.int64, device="cuda")
    a.seed_dev = seed_dev.data_ptr()
    a.advance_seed = 1
    stream = torch.cuda.Stream()
    ctx.frame_graph_capture(a, stream)
    ctx.frame_graph_launch(stream)
    stream.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(n_frames):
        ctx.frame_graph_launch(stream)
    e1.record(stream)
    stream.synchronize()
    ms = e0.elapsed_time(e1) / n_frames
    sizes = b["sizes"].cpu().numpy().astype(np.int64)
    status = int(b["status"].item())
    queries = int(sizes[:nb].sum())
    # the same frame bounce by bounce: per-bounce entry points, size read back each bounce
    queue = torch.arange(P, dtype=torch.int32, device="cuda")
    q = torch.empty(2 * P, device="cuda")
    surv = torch.empty(2 * P, dtype=torch.int32, device="cuda")
    ns = torch.zeros(1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    n = P
    for bnc in range(nb):
        vb = na.frame_vertices(SEED, bnc, queue, n)
        args = ctx.decide_args(vb, film, q, surv, ns, q_mode=na.Q_RRSNET_CHAINED, seed=SEED, stream_id=1000 + bnc)
        na.check(ctx.rrs_infer_decide(args), "rrs_infer_decide")
        n = int(ns.item())
        queue = surv[:n].clone()
    torch.cuda.synchronize()
    host_ms = 1e3 * (time.perf_counter() - t0)
    return {"resolution": "1920x1080", "bounces": nb, "tile_shape": {"waves": shape[0], "m": shape[1]},
            "autotune_best_ms": round(tune_ms, 3), "ms_per_frame_graph": round(ms, 3),
            "queries_per_frame": queries, "rrs_queries_per_s": round(queries / (ms * 1e-3), 1),
            "queue_sizes_last_frame": sizes.tolist(), "status": status,
            "ms_per_frame_host_driven_bounces": round(host_ms, 3)}


def online_loop(torch, na, np, n_frames=256, warm_frames=32):
    """256-frame online loop: every frame one persistent wavefront launch (10 bounces of
    1920x1080) then one training step on 2^18 sampled vertices: StatNet alone for the
    first 32 frames, StatNet + RRSNet afterwards. Per-frame latency from HIP events."""
    P = 1920 * 1080
    nb = 10
    B = 1 << 18
    ctx = na.Context(precision=na.BF16, n_pixels=P, max_queue=2 * P, max_train=B)
    film = torch.exp(torch.randn(3, P, device="cuda") - 1.0)
    a, bufs = ctx.frame_buffers(P, nb, capacity=2 * P)
    a.film_dev = film.data_ptr()
    seed_dev = torch.tensor([SEED], dtype=torch.int64, device="cuda")
    a.seed_dev = seed_dev.data_ptr()
    a.advance_seed = 1
    stream = torch.cuda.Stream()
    ctx.frame_graph_capture(a, stream)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(SEED)
    dev = dict(device="cuda", dtype=torch.float32)
    err = torch.rand(B, generator=gen, **dev) * 2.0
    sw = torch.ones(B, **dev)
    qt = torch.rand(B, generator=gen, **dev) * 5.0 + 0.05
    side = ctx.side(error=err, sample_weight=sw, q_target=qt, error_sum=0.7 * P)
    loss = torch.zeros(2, **dev)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(n_frames)]
    with torch.cuda.stream(stream):
        t0 = time.perf_counter()
        for f in range(n_frames):
            ev[f][0].record(stream)
            ctx.frame_graph_launch(stream)
            vb = na.synth_vertices(B, SEED + 7919 * f, P, stream=stream)
            L = torch.exp(torch.randn(3, B, generator=gen, **dev) * 1.5 - 1.0)
            L = torch.where(torch.rand(B, generator=gen, **dev) < 0.05, L * 100.0, L)
            mask = 1 if f < warm_frames else 3
            na.check(ctx.train_step(vb, mask, film=film, target_L=L, side=side, loss=loss, stream=stream),
                     "train_step")
            ev[f][1].record(stream)
        stream.synchronize()
        wall = time.perf_counter() - t0
    lat = np.array([e[0].elapsed_time(e[1]) for e in ev])
    return {"frames": n_frames, "frames_per_s": round(n_frames / wall, 2),
            "latency_ms_p50": round(float(np.percentile(lat, 50)), 3),
            "latency_ms_p99": round(float(np.percentile(lat, 99)), 3),
            "train_batch": B, "statnet_only_frames": warm_frames,
            "final_loss": [round(x, 5) for x in loss.cpu().tolist()],
            "queue_sizes_last_frame": bufs["sizes"].cpu().tolist(), "status": int(bufs["status"].item())}


def online_drift(torch, na, np, frames=16, warm=4, n=4096):
    """Drift of q against the FP64 oracle trained on the same seed stream, on a reduced online
    loop (the FP64 oracle cannot follow 2^18 samples per frame): `frames` frames of 4096 pixels,
    each one 4-bounce persistent frame launch and one BF16 training step on `n` samples (StatNet
    alone for `warm` frames), against the oracle's FP64 steps on the same inputs from the same
    initial parameters; q of a 4096-vertex probe set after every frame, each side on its own
    parameters (tests/online_drift.py)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import online_drift as od
    import oracle_ref as ora

    cfg = od.config(na, na.BF16, n, 4096)
    p0 = od.initial_params(na, cfg)
    t0 = time.perf_counter()
    q_ref, l_ref = od.oracle_trajectory(na, ora, cfg, p0, frames, warm, n, 4096)
    t_ora = time.perf_counter() - t0
    q_dev, _, status = od.device_trajectory(torch, na, cfg, p0, frames, warm, n, 4096)
    c = od.drift_curve(q_dev, q_ref)
    return {"frames": frames, "statnet_only_frames": warm, "train_batch": n, "pixels": 4096, "probe": 4096,
            "mean_per_frame": [float(f"{x:.3e}") for x in c[:, 0]],
            "p99_per_frame": [float(f"{x:.3e}") for x in c[:, 1]],
            "max_per_frame": [float(f"{x:.3e}") for x in c[:, 2]],
            "first_row": "untrained start (same parameters on both sides)", "frame_status": status,
            "fp64_oracle_s": round(t_ora, 1)}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported"}))
        sys.exit(2)
    import numpy as np
    import torch

    import nrrs_amd as na

    torch.cuda.set_device(0)
    if args.pmc_probe:
        src = torch.ones(64 << 20, device="cuda")  # 256 MiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        path = Path(na.BF16, torch, na, np, shipped=args.probe_shipped)
        for k in range(8 + 3):  # warm-up (clocks, caches), then 3 measured steps
            path.infer(k)
            path.train()
        torch.cuda.synchronize()
        print(json.dumps({"pmc_probe": "done", "copy_bytes": src.numel() * 4}))
        return
    precs = {"bf16": na.BF16, "fp16": na.FP16, "fp32": na.FP32, "fp32_plain": na.FP32}
    # fp32: level-major grid encoding pass before the network kernel (the default); fp32_plain:
    # every level gathered per vertex inside the network kernel
    extra = {"fp32_plain": dict(fp32_level_major=0)}
    results, q_by, ops, n_mlp, n_grid = {}, {}, {}, {}, 0
    for name in args.precisions.split(","):
        path = Path(precs[name], torch, na, np, **extra.get(name, {}))
        r = path.run(args.steps if name == "bf16" else max(2, args.steps // 2), args.warmup)
        if name == "bf16":
            ops = path.op_table(max(3, args.steps // 2))
            n_grid, n_mlp["statnet"] = path.ctx.param_count(na.STATNET)
            n_mlp["rrsnet"] = path.ctx.param_count(na.RRSNET)[1]
        # q of the same (pre-training) parameters at the full batch, against the FP32 path
        if name != "fp32_plain":
            fresh = Path(precs[name], torch, na, np)
            fresh.infer(0)
            q_by[name] = fresh.q.cpu().numpy()
            del fresh
        results[name] = r
        del path
        torch.cuda.empty_cache()
        print(f"[bench] {name}: {json.dumps({k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()})}",
              file=sys.stderr, flush=True)

    bf = results["bf16"]
    n = N_VERTICES
    # every operation of the BF16 step against its roofline (HBM bytes, or MFMA flops)
    op_rows = {}
    for name, o in ops.items():
        kind, work = op_work(name, n, n_grid, n_mlp)
        row = {"ms_per_step": round(o["ms_per_step"], 4), "launches_per_step": o["launches_per_step"]}
        if kind:
            t = o["ms_per_step"] * 1e-3 / o["launches_per_step"]
            if kind == "bytes":
                row.update(bound="hbm", achieved=round(work / t / 1e9, 1), unit="GB/s",
                           frac=round(work / t / 1e9 / HBM_PEAK_GBS, 4), work_per_launch=work)
            else:
                row.update(bound="mfma", achieved=round(work / t / 1e12, 2), unit="TFLOP/s",
                           frac=round(work / t / 1e12 / BF16_PEAK_TFLOPS, 4), work_per_launch=work)
        op_rows[name] = row
    # dominant kernel of the step: the paired level-major encode pass (k_encode_levels<BF16, true>),
    # launched by the RRS stage (encode_pair) and by the training step (encode_pair_counts, which
    # also counts the table-gradient records); the entry uses the RRS stage's launch
    enc = op_rows.get("encode_pair", {})
    roofline = {"bound": "hbm", "achieved": enc.get("achieved"), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": enc.get("frac"), "traffic": ENCODE_TRAFFIC_BYTES,
                "kernel": "k_encode_levels<BF16, true>: both nets' level-major hash-grid encodings from the paired "
                          "table, 2^20 vertices (two launches per step: the RRS stage's and the training step's, the "
                          "latter also counting the table-gradient records)",
                "kernel_us": round(ops["encode_pair"]["ms_per_step"] * 1e3, 1) if "encode_pair" in ops else None,
                "algorithmic_bytes_per_launch": 1164 * n,
                "bytes_per_vertex": "12 B position + 16 levels x 8 corners x 8 B + 2 x 64 B encodings = 1164 B"}
    matched = {}
    if "fp32" in q_by:
        for name in q_by:
            if name != "fp32":
                rel = np.abs(q_by[name] - q_by["fp32"]) / q_by["fp32"]
                matched[name] = {"q_rel_err_vs_fp32_2p20_max": float(rel.max()),
                                 "q_rel_err_vs_fp32_2p20_mean": float(rel.mean())}
    if not args.no_oracle:
        fp64 = matched_errors(torch, na, np, {k: precs[k] for k in args.precisions.split(",") if k != "fp32_plain"})
        for name, r in fp64.items():
            matched.setdefault(name, {}).update({"vs_fp64_oracle": r})
        print(f"[bench] matched error: {json.dumps(matched)}", file=sys.stderr, flush=True)
    out = {
        "metric": "RRS queries/s + train samples/s at 2^20 vertices, BF16 vs FP32 at matched error",
        "value": round(bf["step_queries_per_s"], 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(bf["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic (PCG32-seeded vertices, log-normal radiance targets, random-init grid, Xavier MLP)",
        "config": {"workload": "online RRS-stage step at 2^20 vertices: StatNet->RRSNet chained inference + "
                               "normalise + split + compaction, then one joint StatNet+RRSNet training step",
                   "vertices": n, "pixels": N_PIXELS, "grid": "16 levels x 2 features, T=2^19, base 16, scale 1.5",
                   "mlp": "StatNet 80-64x4-6, RRSNet 80-64x3-1"},
        "per_precision": {k: {kk: (round(vv, 2) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                          for k, v in results.items()},
        "rrs_queries_per_s": round(bf["rrs_queries_per_s"], 1),
        "train_samples_per_s": round(bf["train_samples_per_s"], 1),
        "matched_error": matched,
        "roofline": roofline,
        "ops": op_rows,
    }
    for ref, key in (("fp32", "bf16_over_fp32"), ("fp32_plain", "bf16_over_fp32_plain")):
        if ref in results:
            out[key] = {
                "rrs_queries": round(bf["rrs_queries_per_s"] / results[ref]["rrs_queries_per_s"], 3),
                "train_samples": round(bf["train_samples_per_s"] / results[ref]["train_samples_per_s"], 3),
                "step": round(bf["step_queries_per_s"] / results[ref]["step_queries_per_s"], 3)}
    if not args.no_shipped:
        out["shipped_nets"] = shipped_nets(torch, na, np, args.steps, args.warmup)
        print(f"[bench] shipped nets: {json.dumps(out['shipped_nets'])}", file=sys.stderr, flush=True)
    if not args.no_peaks:
        pk = measured_peaks(torch)
        out["measured_peaks"] = pk
        if roofline["achieved"]:
            roofline["frac_of_measured_copy"] = round(roofline["achieved"] / pk["copy_GBps"], 4)
        print(f"[bench] peaks: {json.dumps(pk)}", file=sys.stderr, flush=True)
    if not args.no_frame:
        out["wavefront_frame"] = frame_wavefront(torch, na, np)
        print(f"[bench] wavefront: {json.dumps(out['wavefront_frame'])}", file=sys.stderr, flush=True)
        out["online_loop"] = online_loop(torch, na, np)
        if not args.no_oracle:
            out["online_loop"]["q_drift_vs_fp64"] = online_drift(torch, na, np)
        print(f"[bench] online: {json.dumps(out['online_loop'])}", file=sys.stderr, flush=True)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_sample)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
