"""Benchmark of the neural RRS stage on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp32] [--no-cpu]

A step is one bounce of the RRS stage over 2^20 synthetic path vertices that
are already resident in HBM: StatNet inference, RRSNet inference on the StatNet
statistics, q, the exact normalisation sum, the split decision and the stable
survivor compaction. `value` is RRS queries per second of the whole step.

The roofline entry covers the dominant kernel (the fused StatNet encode+MLP
kernel), timed with HIP events on the stream it runs on. The CPU baseline is the
FP64 oracle run single-threaded on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nrrs-code_amd"))

N_VERTICES = 1 << 20
N_PIXELS = 1024 * 1024
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
MLP_MACS = {"statnet": 80 * 64 + 3 * 64 * 64 + 64 * 6, "rrsnet": 80 * 64 + 2 * 64 * 64 + 64 * 1}


def statnet_bytes_per_query(table_bytes):
    # input record (pos, dir, normal, rough, albedo_d, albedo_s) + 16 levels x 8 corners
    # x 2 features + 6 FP32 outputs
    return 64 + 16 * 8 * 2 * table_bytes + 24


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="fp32", choices=["fp32"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU-baseline leg")
    ap.add_argument("--cpu-sample", type=int, default=1 << 15)
    return ap.parse_args()


def cpu_baseline(sample):
    """FP64 oracle, one thread, on `sample` vertices of the same recipe."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ref as ora
    import nrrs_amd as na

    cfg = na.default_config(n_pixels=N_PIXELS, max_queue=sample)
    d = ora.synth_vertices(sample, 0x4E525253 ^ 2, N_PIXELS)
    hv = ora.HostVertices(d)
    film = np.random.default_rng(1).lognormal(-1, 1, (3, N_PIXELS)).astype(np.float32)
    ps = ora.init_params(cfg, na.STATNET, 1, 0.5)
    pr = ora.init_params(cfg, na.RRSNET, 2, 0.5)
    t0 = time.perf_counter()
    stats = ora.net_forward(cfg, na.STATNET, hv, film, None, ps).T
    raw = ora.net_forward(cfg, na.RRSNET, hv, film, np.ascontiguousarray(stats), pr)[:, 0]
    q = ora.q_f64(cfg, na.Q_RRSNET_CHAINED, hv, raw, film)
    ora.rrs_decide(cfg, q.astype(np.float32), seed=1, stream=0)
    dt = time.perf_counter() - t0
    return {"value": sample / dt, "unit": "queries/s", "cores": 1, "kind": "port",
            "sample": f"{sample} vertices of the same workload, FP64 oracle, 1 thread, {dt:.2f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported"}))
        sys.exit(2)
    import numpy as np
    import torch

    import nrrs_amd as na

    torch.cuda.set_device(0)
    n = N_VERTICES
    ctx = na.Context(n_pixels=N_PIXELS, max_queue=n)
    rng = np.random.default_rng(0)
    for net in (na.STATNET, na.RRSNET):
        p = ctx.params_get(net)
        g, _ = ctx.param_count(net)
        p[:g] = rng.uniform(-0.5, 0.5, g).astype(np.float32)
        ctx.params_set(net, p)
    vb = na.synth_vertices(n, 0x4E525253 ^ 2, N_PIXELS)
    film = torch.from_numpy(rng.lognormal(-1, 1, (3, N_PIXELS)).astype(np.float32)).cuda()
    q = torch.empty(n, device="cuda")
    qn = torch.empty(n, device="cuda")
    split = torch.empty(n, dtype=torch.uint8, device="cuda")
    surv = torch.empty(8 * n, dtype=torch.int32, device="cuda")
    n_surv = torch.zeros(1, dtype=torch.int32, device="cuda")
    ovf = torch.zeros(1, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()

    args_chain = ctx.decide_args(vb, film, q, surv, n_surv, ovf, q_mode=na.Q_RRSNET_CHAINED, qn=qn, split=split)

    def step(k):
        args_chain.pcg_stream = k
        na.check(ctx.rrs_infer_decide(args_chain), "rrs_infer_decide")

    for k in range(args.warmup):
        step(k)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(args.warmup + k)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = 1e3 * dt / args.steps
    qps = n * args.steps / dt

    # dominant kernel: fused StatNet encode + MLP, HIP events on its stream
    reps = max(args.steps, 5)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    stats = ctx.statnet_infer(vb)
    e0.record(stream)
    for _ in range(reps):
        ctx.statnet_infer(vb)
    e1.record(stream)
    torch.cuda.synchronize()
    kern_s = e0.elapsed_time(e1) / 1e3 / reps
    del stats
    bytes_q = statnet_bytes_per_query(4)
    flops_q = 2 * MLP_MACS["statnet"]
    gbs = bytes_q * n / kern_s / 1e9
    tflops = flops_q * n / kern_s / 1e12
    # the FP32 path is VALU-bound: report against the FP32 peak
    roofline = {"bound": "mfma", "achieved": round(tflops, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(tflops / FP32_PEAK_TFLOPS, 4), "traffic": None,
                "kernel": "k_net_fp32<5,6,EPI_RAW> (StatNet encode+MLP)",
                "kernel_us": round(kern_s * 1e6, 2), "hbm_gbs_algorithmic": round(gbs, 1)}

    out = {
        "metric": "RRS queries/s + train samples/s at 2^20 vertices, BF16 vs FP32 at matched error",
        "value": round(qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic (PCG32 seeded vertices, random-init grid and Xavier MLP weights)",
        "config": {"workload": "RRS stage, StatNet->RRSNet chained inference + normalise + split + "
                               "compaction, 2^20 vertices, 16-level hash grid T=2^19, 64-wide MLPs",
                   "vertices": n, "pixels": N_PIXELS, "precision": args.precision},
        "roofline": roofline,
        "survivors_last_step": int(n_surv.item()),
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_sample)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
