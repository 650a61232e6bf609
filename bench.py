"""Benchmark of the LMM pricing path on one MI355X.

Workload: the 40-rate, 5-factor LMM caplet portfolio (2^20 scrambled Sobol
paths, 40 steps, 40 caplets at K = 0.045) priced with all 80 adjoint Greeks
(40 deltas, 40 vegas), rate state checkpointed every 8 steps.  One "step" of
the benchmark is that whole job: fused Sobol/bridge generation, the forward
price kernel writing checkpoints, and the reverse (adjoint) kernel, with all
inputs resident in HBM.  ``value`` is P*S / (time of one job) in path-steps/s.

Also reported: the forward-only pipeline (generation + price kernel) and the
adjoint ratio (job time / forward-only time), price and Greek errors against
the FP64 oracle on the same Sobol points (on a 2^14-path prefix of the same
workload), a roofline entry for the dominant kernel (the reverse sweep,
VALU-bound, algorithmic flops / HIP-event duration), and the FP64 oracle timed
on the host cores as a CPU baseline.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "LMM path-steps/s; adjoint all-Greeks time ÷ price time; error vs FP64"
FP32_VECTOR_PEAK_TFLOPS = 157.3  # MI355X published FP32 vector peak


def flops_per_rate_update(F):
    """Algorithmic flops of one rate update (FMA = 2), counted from the
    kernels' arithmetic: forward PC step 13F + 20; reverse step 21F + 40."""
    return 13 * F + 20, 21 * F + 40


def rate_updates_per_path(N):
    return N * (N + 1) // 2  # sum over steps s of alive rates N - s


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--paths", type=int, default=1 << 20)
    ap.add_argument("--checkpoint-every", type=int, default=8)
    ap.add_argument("--cpu-sample-paths", type=int, default=1 << 17)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-error-check", action="store_true")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after the benchmark, copy a buffer of this many MiB twice (device to "
                         "device): a plain copy of known size that calibrates HBM byte counters")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        sys.exit("bench.py measures one GPU only: the pricing path runs on a single MI355X in "
                 "this build (multi-GPU sharding by Sobol path offset is described, not built); "
                 "run with --gpus 1.")

    import torch

    import oracle as O
    from useful_code_amd import configs, pipeline

    torch.cuda.set_device(0)
    w = configs.caplet_portfolio(paths=args.paths, checkpoint_every=args.checkpoint_every)
    P, N, F, S = w.paths, w.n_rates, w.factors, w.steps
    pr = pipeline.LMMPricer(w)

    ev = lambda: torch.cuda.Event(enable_timing=True)  # noqa: E731

    def job(record=None):
        if record is None:
            pr.generate()
            pr.price_kernel(checkpoints=True)
            pr.greeks_kernel()
            return
        e0, e1, e2, e3 = record
        e0.record()
        pr.generate()
        e1.record()
        pr.price_kernel(checkpoints=True)
        e2.record()
        pr.greeks_kernel()
        e3.record()

    def forward(record=None):
        if record is None:
            pr.generate()
            pr.price_kernel(checkpoints=False)
            return
        e0, e1, e2 = record
        e0.record()
        pr.generate()
        e1.record()
        pr.price_kernel(checkpoints=False)
        e2.record()

    for _ in range(args.warmup):
        job()
        forward()
    torch.cuda.synchronize()

    # timed region: K whole jobs (price + 80 Greeks)
    recs = [(ev(), ev(), ev(), ev()) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in recs:
        job(r)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_job = wall * 1e3 / args.steps
    gen_ms = np.mean([r[0].elapsed_time(r[1]) for r in recs])
    fwdck_ms = np.mean([r[1].elapsed_time(r[2]) for r in recs])
    bwd_ms = np.mean([r[2].elapsed_time(r[3]) for r in recs])
    job_ev_ms = np.mean([r[0].elapsed_time(r[3]) for r in recs])

    # forward-only pipeline (price, no checkpoints), same count
    frecs = [(ev(), ev(), ev()) for _ in range(args.steps)]
    torch.cuda.synchronize()
    for r in frecs:
        forward(r)
    torch.cuda.synchronize()
    fwd_pipe_ms = np.mean([r[0].elapsed_time(r[2]) for r in frecs])
    fwd_ms = np.mean([r[1].elapsed_time(r[2]) for r in frecs])
    price = pr.price_out.item()

    # Same job with a checkpoint at every step (the forward also keeps each
    # step's drift sums, the reverse sweep recomputes nothing).
    k1 = {}
    if w.checkpoint_every != 1:
        w1 = configs.caplet_portfolio(paths=args.paths, checkpoint_every=1)
        pr1 = pipeline.LMMPricer(w1)
        pr1.z = pr.z  # same variates
        for _ in range(args.warmup):
            pr1.price_kernel(checkpoints=True)
            pr1.greeks_kernel()
        r1 = [(ev(), ev(), ev()) for _ in range(args.steps)]
        torch.cuda.synchronize()
        for e0, e1, e2 in r1:
            e0.record()
            pr1.price_kernel(checkpoints=True)
            e1.record()
            pr1.greeks_kernel()
            e2.record()
        torch.cuda.synchronize()
        k1_fwd = np.mean([a.elapsed_time(b) for a, b, _ in r1])
        k1_bwd = np.mean([b.elapsed_time(c) for _, b, c in r1])
        k1 = {"price_fwd_checkpoints_ms": k1_fwd, "price_bwd_ms": k1_bwd,
              "job_ms": gen_ms + k1_fwd + k1_bwd,
              "adjoint_ratio": (gen_ms + k1_fwd + k1_bwd) / fwd_pipe_ms,
              "greeks_equal_k8": bool(torch.allclose(pr1.greeks_out, pr.greeks_out,
                                                     rtol=2e-5, atol=0.0))}
        del pr1

    f_fwd, f_bwd = flops_per_rate_update(F)
    upd = rate_updates_per_path(N) * P
    bwd_flops = (f_fwd + f_bwd) * upd  # segment recompute + reverse step
    achieved_tflops = bwd_flops / (bwd_ms * 1e-3) / 1e12

    out = {
        "metric": METRIC,
        "value": P * S / (ms_job * 1e-3),
        "unit": "path-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_job,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: seeded Sobol points (Joe-Kuo directions, MT19937 scramblers), "
                "seeded initial curve",
        "config": {"workload": "lmm_caplets_40x5: 2^20 paths, 40 rates = 40 steps, 5 factors, "
                               "40 caplets K=0.045, price + 40 deltas + 40 vegas by adjoint, "
                               f"checkpoint every {w.checkpoint_every} steps",
                   "paths": P, "rates": N, "steps": S, "factors": F,
                   "checkpoint_every": w.checkpoint_every},
        "forward_path_steps_per_s": P * S / (fwd_pipe_ms * 1e-3),
        "forward_ms": fwd_pipe_ms,
        "adjoint_ratio": job_ev_ms / fwd_pipe_ms,
        "price": price,
        "kernels_ms": {"sobol_bridge": gen_ms, "price_fwd_checkpoints": fwdck_ms,
                       "price_bwd": bwd_ms, "price_fwd": fwd_ms},
        "checkpoint_every_1": k1,
        "roofline": {
            "kernel": "k_price_bwd (reverse sweep + segment recompute)",
            "bound": "valu",
            "achieved": achieved_tflops,
            "peak": FP32_VECTOR_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / FP32_VECTOR_PEAK_TFLOPS,
            "traffic": None,
        },
    }

    if not args.no_error_check:
        # Error vs FP64 on the first 2^14 paths of the same Sobol sequence.
        ws = configs.caplet_portfolio(paths=1 << 14, checkpoint_every=w.checkpoint_every)
        ps = pipeline.LMMPricer(ws)
        p_gpu, g_gpu = ps.price_and_greeks()
        g_gpu = g_gpu.cpu().numpy()
        ints = O.scramble(O.sobol_ints(ws.paths, ws.dims, 0), ws.scramblers, ws.dims)
        m = O.caplet_model(ws.f0, ws.taus, ws.disp(), ws.A, ws.vols, ws.strike, ws.paths)
        p64, d64, v64 = O.caplet_pipeline(m, ints, adjoint=True)
        g64 = np.concatenate([d64, v64])
        out["price_abs_err_vs_fp64"] = abs(p_gpu.item() - p64)
        out["greeks_max_rel_err_vs_fp64"] = float(np.max(np.abs(g_gpu - g64) / np.abs(g64)))
        out["error_sample"] = "first 2^14 paths of the workload's Sobol sequence"

    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        wc = configs.caplet_portfolio(paths=args.cpu_sample_paths)
        ints = O.scramble(O.sobol_ints(wc.paths, wc.dims, 0), wc.scramblers, wc.dims)
        m = O.caplet_model(wc.f0, wc.taus, wc.disp(), wc.A, wc.vols, wc.strike, wc.paths)
        t = time.perf_counter()
        O.caplet_pipeline(m, ints, adjoint=True)
        dt = time.perf_counter() - t
        out["cpu_baseline"] = {
            "value": wc.paths * S / dt,
            "unit": "path-steps/s",
            "cores": threads,
            "kind": "port",
            "sample": f"FP64 oracle (oracle/koo_oracle.c, OpenMP over paths, gcc -O2) pricing "
                      f"the same workload with all 80 adjoint Greeks on its first "
                      f"{wc.paths} paths: {dt:.2f} s",
        }

    if args.copy_probe_mib > 0:
        n = args.copy_probe_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        out["copy_probe_bytes"] = {"read": n * 4, "write": n * 4}

    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
