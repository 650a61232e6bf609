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
on the host cores as a CPU baseline.  The ``bermudan_ls`` object reports the
Longstaff-Schwartz workload (Bermudan payer swaption, 20 rates, 3 factors,
2^18 paths, 10 exercise dates): stage times, its own roofline entry for the
per-date regression kernel (HBM-bound) and its price error against the FP64
oracle at full size.

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
HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth


def flops_per_rate_update(F):
    """Algorithmic flops of one rate update (one path, step and alive rate),
    an FMA counted as 2 and a hardware exp or reciprocal as 1, counted from
    the arithmetic of the shared step code (lmm_step.h):
      forward PC step (pc_step): 5F FMA (e, mu, dW, e^, mu^ sums), 3 MUL,
        4 ADD, 4 transcendental (exp and reciprocal of g and g^) = 10F + 11;
      reverse step (reverse_step): 9F + 6 FMA (mu, mu^, dW recomputed, the
        adjoint sums E^, E and their contractions, e and e^ un-accumulated,
        the lambda, lambda^ and vega terms), 7 MUL, 7 ADD (three of them the
        DPP fold of the vega term), 4 transcendental = 18F + 30.
    At F = 5: 61 + 120 = 181 for the backward kernel (segment recompute plus
    reverse).  The SQ counters of k_price_bwd<5> read 195 executed FP32 flops
    per update (85.1 FMA, 6.6 MUL, 10.1 ADD, 8.2 transcendental): the
    difference is per-step work spread over the rates and multiply-adds the
    compiler forms from a MUL and an ADD."""
    return 10 * F + 11, 18 * F + 30


def rate_updates_per_path(N):
    return N * (N + 1) // 2  # sum over steps s of alive rates N - s


# Bytes the LS first pass must move per path and exercise date (FP32): the two basis
# variables, exercise value and numeraire, the later date's two variables,
# exercise value and numeraire for its exercise decision, the continuation
# value read and written.
LS_DATE_BYTES_PER_PATH = 4 * (2 + 1 + 1 + 2 + 1 + 1 + 1 + 1)


def profiled_traffic(kernel, grid):
    """HBM bytes per dispatch of `kernel` at `grid` (total threads) from the
    rocprofv3 counter summaries committed under profiles/ (their
    `traffic_json:` lines: FETCH_SIZE + WRITE_SIZE scaled by a plain-copy
    calibration), with the summary's file name; (None, None) when absent."""
    import glob
    for fn in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(fn) as f:
            for line in f:
                if not line.startswith("traffic_json: "):
                    continue
                rec = json.loads(line[len("traffic_json: "):])
                if rec.get("kernel") == kernel and rec.get("grid") == grid and rec.get("hbm_bytes"):
                    return float(rec["hbm_bytes"]), os.path.relpath(fn, ROOT)
    return None, None


def profiled_mfma(kernel, grid):
    """MFMA counter record of `kernel` at `grid` from the `mfma_json:` lines of
    the counter summaries under profiles/ (executed MFMA flops = 512 x
    SQ_INSTS_VALU_MFMA_MOPS_F32/_F64, busy = SQ_VALU_MFMA_BUSY_CYCLES over
    GRBM_GUI_ACTIVE / 8 XCDs x 1024 SIMDs), with the file name; (None, None) when absent."""
    import glob
    for fn in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(fn) as f:
            for line in f:
                if line.startswith("mfma_json: "):
                    rec = json.loads(line[len("mfma_json: "):])
                    if rec.get("kernel") == kernel and rec.get("grid") == grid:
                        return rec, os.path.relpath(fn, ROOT)
    return None, None


def with_mfma(roof, kernel, grid, kernel_ms):
    """Adds `mfma`: the profiled executed MFMA flops per dispatch over this
    run's live kernel time, against the dense MFMA peak of the type the
    counters saw, and the profiled MFMA busy fraction."""
    rec, src = profiled_mfma(kernel, grid)
    if rec is None:
        roof["mfma"] = None
        return roof
    if kernel_ms is None:  # no live time of this kernel alone: the trace's average
        kernel_ms = rec["avg_us"] * 1e-3
    tf = rec["mfma_flops"] / (kernel_ms * 1e-3) / 1e12
    roof["mfma"] = {"kernel": kernel, "executed_flops": rec["mfma_flops"],
                    "achieved": tf, "peak": rec["peak_tflops"], "unit": "TFLOP/s",
                    "frac": tf / rec["peak_tflops"], "busy_profiled": rec["mfma_busy"],
                    "source": src}
    return roof


def with_traffic(roof, kernel, grid):
    t, src = profiled_traffic(kernel, grid)
    roof["traffic"] = t
    if src:
        roof["traffic_source"] = src
    return roof


def bermudan_section(args, torch, O, configs, pipeline, ev):
    """Config-3 Longstaff-Schwartz timings, roofline and error."""
    w = configs.bermudan_swaption(paths=args.bermudan_paths)
    P, S, E = w.paths, w.steps, len(w.exercise_steps)
    pr = pipeline.BermudanPricer(w)

    def job(r=None):
        if r is not None:
            r[0].record()
        pr.generate()
        if r is not None:
            r[1].record()
        pr.paths_kernel()
        if r is not None:
            r[2].record()
        pr.regress()
        if r is not None:
            r[3].record()
        pr.second_pass()
        if r is not None:
            r[4].record()

    for _ in range(args.warmup):
        job()
    recs = [tuple(ev() for _ in range(5)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for r in recs:
        job(r)
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    gen = np.mean([r[0].elapsed_time(r[1]) for r in recs])
    paths = np.mean([r[1].elapsed_time(r[2]) for r in recs])
    regress = np.mean([r[2].elapsed_time(r[3]) for r in recs])
    second = np.mean([r[3].elapsed_time(r[4]) for r in recs])
    # the same job recorded once as a HIP graph and replayed
    pr.job(capture=True)
    grecs = [(ev(), ev()) for _ in range(args.steps)]
    torch.cuda.synchronize()
    for a, b in grecs:
        a.record()
        pr.job(capture=True)
        b.record()
    torch.cuda.synchronize()
    graph_ms = float(np.mean([a.elapsed_time(b) for a, b in grecs]))
    price1, price2 = pr.price1.item(), pr.price2.item()
    itm = float((pr.ev.view(E, P) > 0).float().mean().item())
    # The whole first pass (E + 1 round launches: every block sums the
    # previous round's partials and solves, then its paths' date work; the
    # price): algorithmic bytes of the E dates over its event time.
    stage_bytes = E * LS_DATE_BYTES_PER_PATH * P
    gbs = stage_bytes / (regress * 1e-3) / 1e9
    k = 6
    mfma_flops = E * 2.0 * itm * P * (k + 1) ** 2
    res = {
        "workload": "bermudan_payer_20x3: 2^18 Sobol paths, 20 rates = 20 steps, 3 factors, "
                    "10 exercise dates (steps 0,2,..,18), ITM-only LS on 6 quadratic functions "
                    "of (co-terminal swap rate, first alive forward)",
        "paths": P,
        "path_steps_per_s": P * S / ((gen + paths + regress) * 1e-3),
        "job_ms": wall_ms,
        "graph_job_ms": graph_ms,
        "kernels_ms": {"sobol_bridge": gen, "bermudan_paths": paths, "ls_first_pass": regress,
                       "ls_second_pass": second},
        "price_first_pass": price1,
        "price_second_pass_same_paths": price2,
        "itm_fraction_mean": itm,
        "roofline": with_traffic({
            "kernel": "LS first pass (ls_first_pass_gpu): one k_ls_round_single<2> launch per "
                      "exercise date and one for the price (per date: every block sums the "
                      "previous date's FP64 partials and solves, decision update, MFMA normal "
                      "equations with FP64 accumulation), then k_finalize_cols",
            "bound": "hbm",
            "achieved": gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": gbs / HBM_PEAK_GBS,
            "traffic": None,
            "algorithmic_bytes": stage_bytes,
            "mfma_useful_tflops": mfma_flops / (regress * 1e-3) / 1e12,
        }, "ls_first_pass_gpu", P),
    }
    with_mfma(res["roofline"], "ls_first_pass_gpu", P, regress)
    if not args.no_error_check:
        ints = O.scramble(O.sobol_ints(P, w.dims, 0), w.scramblers, w.dims)
        z = O.bridge_transform(O.invnorm_u32(ints), S, w.factors)
        m = O.caplet_model(w.f0, w.taus, w.disp(), w.A, w.vols, w.strike, P)
        t = time.perf_counter()
        bv, evo, num = O.bermudan_paths(m, z, w.exercise_steps)
        # the variables as the GPU regresses on them, (x - shift) * scale
        bvn = (bv - w.basis_shift[:, :, None]) * w.basis_scale[:, :, None]
        p64, _ = O.ls_first_pass(bvn, np.full(E, 2, np.int32), np.zeros_like(evo), evo, num,
                                 itm_only=True)
        dt = time.perf_counter() - t
        res["price_abs_err_vs_fp64"] = abs(price1 - p64)
        res["price_fp64"] = p64
        res["cpu_fp64_paths_and_ls_s"] = dt
    return res


def multi_ls_section(args, torch, O, configs, pipeline, ev):
    """The reference's multi-regression LS first pass (nested re-regressions
    with the fraction selector, per-regression normalisation) on the config-3
    Bermudan path quantities, against the FP64 oracle on the same inputs."""
    from useful_code_amd import koo
    w = configs.bermudan_swaption(paths=args.bermudan_paths)
    P, E = w.paths, len(w.exercise_steps)
    R, cutoff, sel = 3, 10, (0.25, 0.26, 2.5, 0.8)  # LSmulti_test.cu:270-273
    pr = pipeline.BermudanPricer(w)
    pr.generate()
    pr.paths_kernel()
    maxd = O.basis_count(2, True)
    ws = torch.empty(koo.ls_multi_workspace(P), dtype=torch.uint8, device="cuda")
    co = torch.zeros(E * R * maxd, dtype=torch.float64, device="cuda")
    lo = torch.zeros(E * R, dtype=torch.float64, device="cuda")
    up = torch.zeros_like(lo)
    price = torch.zeros(1, dtype=torch.float64, device="cuda")
    cash = torch.zeros(E * P, dtype=torch.float32, device="cuda")

    def run():
        koo.ls_multi_first_pass_gpu(E, pr.bv, pr.ves, 2, cash, pr.ev, pr.num, P, 1, 1, R, cutoff,
                                    *sel, ws, co, lo, up, None, None, price)

    for _ in range(args.warmup):
        run()
    recs = [(ev(), ev()) for _ in range(args.steps)]
    torch.cuda.synchronize()
    for a, b in recs:
        a.record()
        run()
        b.record()
    torch.cuda.synchronize()
    ms = float(np.mean([a.elapsed_time(b) for a, b in recs]))
    res = {
        "workload": "bermudan_payer_20x3 path quantities (2^18 paths, 10 exercise dates, 2 "
                    "variables, cross basis): up to 3 nested regressions per date, "
                    "RegressionSelectorFraction(0.25, 0.26, 2.5, 0.8), normalised, FP64",
        "paths": P,
        "ms_per_pass": ms,
        "price_first_pass": price.item(),
        "nested_bands": int((up.view(E, R)[:, :2] > 0).sum().item()),
    }
    with_mfma(res, "k_multi_gram<2>", P, None)
    if not args.no_error_check:
        bv = pr.bv.view(E, 2, P).cpu().double().numpy()
        evo = pr.ev.view(E, P).cpu().double().numpy()
        num = pr.num.view(E, P).cpu().double().numpy()
        t = time.perf_counter()
        p64 = O.ls_multi_first_pass(bv, pr.ves, np.zeros_like(evo), evo, num, 1, R, cutoff, *sel,
                                    normalise=1)[0]
        res["cpu_fp64_oracle_s"] = time.perf_counter() - t
        res["price_fp64"] = p64
        res["price_rel_err_vs_fp64"] = abs(price.item() - p64) / abs(p64)
    return res


def portfolio_section(args, torch, configs, pipeline, ev):
    """Config 5: 64 Bermudan swaptions (10-40 rates, 4-20 exercise dates,
    1-5 factors, 2^17 paths each), in whichever launch form the autotuner
    measured faster: the whole portfolio in one persistent launch (teams of
    resident blocks per product), or the generators over four streams then
    the batched launches (path quantities per rate class, the LS first
    passes one launch per date rank, frozen forward and adjoint backward per
    rate class) -- captured into one HIP graph with the autotuner's
    checkpoint interval; one step = one graph replay (first-pass price,
    frozen price and 2N Greeks for every product)."""
    from useful_code_amd import autotune
    prods = configs.bermudan_portfolio(paths=args.portfolio_paths)
    pf = pipeline.PortfolioPricer(prods, checkpoint_every=4, n_streams=4)
    t = time.perf_counter()
    tuning = autotune.tune_portfolio(pf)
    tune_s = time.perf_counter() - t
    pf.apply_tuning(tuning)
    # eager reference timing (launch-bound on the host)
    torch.cuda.synchronize()
    t = time.perf_counter()
    pf.run(capture=False)
    torch.cuda.synchronize()
    eager_ms = (time.perf_counter() - t) * 1e3
    pf.run(capture=True)  # warm-up + capture + first replay
    torch.cuda.synchronize()
    recs = [(ev(), ev()) for _ in range(args.steps)]
    for a, b in recs:
        a.record()
        pf.run(capture=True)
        b.record()
    torch.cuda.synchronize()
    ms = float(np.mean([a.elapsed_time(b) for a, b in recs]))
    res = pf.results()
    path_steps = sum(w.paths * w.steps for w in prods)
    return {
        "workload": "bermudan_portfolio_64: 64 Bermudan payer swaptions, 10-40 rates, 4-20 "
                    "exercise dates, K 0.03-0.07, 1-5 factors, 2^17 Sobol paths each; LS first pass "
                    "(ITM, MFMA normal equations, FP64 Cholesky) + frozen-strategy adjoint deltas "
                    "and vegas per product; the autotuner's launch form, captured in one graph",
        "graph_replay_ms": ms,
        "eager_ms": eager_ms,
        "path_steps_per_s": path_steps / (ms * 1e-3),
        "products": len(prods),
        "tuning_s": tune_s,
        "checkpoint_every": tuning["k"],
        "ls_blocks_per_product": tuning.get("ls_blocks"),
        "ls_blocks_per_cu_ms": tuning.get("ls_blocks_per_cu_ms"),
        "launch": "persistent" if pf.persistent else "batched",
        "launch_choice_ms": {"persistent (one launch, teams sized by measured work)":
                             tuning.get("persistent_ms"),
                             "batched (per rate class and stage)": tuning.get("batched_ms"),
                             "persistent team-sizing rounds": tuning.get("persistent_rounds_ms")},
        "stream_load_ms": tuning["stream_load_ms"],
        "frozen_price_sum": float(sum(r[1] for r in res)),
        "batched": bool(pf.batched),
    }


def swap_section(args, torch, O, configs, pipeline, ev):
    """Plain-LMM payer swap (20 rates, 3 factors, 2^15 paths) through the
    staged reference-named kernels, with the reference's own CPU pipeline
    (oracle/_ref, one core) timed beside it when it is present."""
    from useful_code_amd import koo
    w = configs.payer_swap()
    P, S = w.paths, w.steps
    pr = pipeline.SwapPricer(w)
    for _ in range(args.warmup):
        pr.price()
    recs = [(ev(), ev()) for _ in range(args.steps)]
    torch.cuda.synchronize()
    for a, b in recs:
        a.record()
        pr.price()
        b.record()
    torch.cuda.synchronize()
    ms = float(np.mean([a.elapsed_time(b) for a, b in recs]))
    price = pr.price_out.item()
    model_free = float(np.sum(w.tau * (w.f0 - w.strike) * np.cumprod(1 / (1 + w.tau * w.f0))))
    res = {
        "workload": "lmm_payer_swap_20x3: 2^15 scrambled Sobol paths, 20 rates = 20 steps, "
                    "3 factors, payer swap K=0.05, staged pipeline of reference-named kernels",
        "paths": P,
        "gpu_ms": ms,
        "path_steps_per_s": P * S / (ms * 1e-3),
        "price": price,
        "model_free_value": model_free,
    }
    if O.ref_available():
        pw = int(np.ceil(np.log2(S)))
        scr = np.random.RandomState(7).randint(0, 2**32, size=(1 << pw) * w.factors,
                                               dtype=np.uint64).astype(np.uint32)
        t = time.perf_counter()
        rp = O.ref_swap_pipeline(P, w.f0, w.taus, w.A, koo.fixed_drifts(w.A), w.strike, scr, pw)
        dt = time.perf_counter() - t
        res["cpu_reference"] = {
            "value": P * S / dt, "unit": "path-steps/s", "cores": 1, "kind": "reference",
            "seconds": dt, "price": rp,
            "sample": "the reference's gold LMMEvolutionRoutineGold + curve, swap-flow and "
                      "discounting routines (compiled from the reference sources into "
                      "oracle/_ref), the same workload with its own Sobol/bridge choices",
        }
    return res


def basket_section(args, torch, O, configs, pipeline, ev):
    """Multi-asset Black-Scholes basket Asian (32 assets, 64 bridge steps,
    2^20 paths): price + 32 deltas + 32 vegas per job; roofline of the fused
    path kernel (HBM-bound: it reads the variates once)."""
    w = configs.basket_asian(paths=args.basket_paths)
    P, A, S = w.paths, w.assets, w.steps
    pr = pipeline.BasketPricer(w)
    for _ in range(args.warmup):
        pr.price_and_greeks()
    recs = [(ev(), ev(), ev()) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for a, b, c in recs:
        a.record()
        pr.generate()
        b.record()
        pr.greeks_kernel()
        c.record()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / args.steps
    gen = float(np.mean([a.elapsed_time(b) for a, b, _ in recs]))
    kern = float(np.mean([b.elapsed_time(c) for _, b, c in recs]))
    out = pr.out.cpu().numpy()
    zbytes = 4.0 * S * A * P
    gbs = zbytes / (kern * 1e-3) / 1e9
    tri = A * (A + 1) // 2  # multiply-adds of the triangular correlation per path-step
    mfma_tflops = 2.0 * tri * S * P / (kern * 1e-3) / 1e12
    res = {
        "workload": "basket_asian_32x64: 32 assets, 64 Sobol-bridge steps (2048 dims), 2^20 paths, "
                    "arithmetic-average basket Asian call K=100, price + 32 deltas + 32 vegas",
        "paths": P,
        "path_steps_per_s": P * S / ((gen + kern) * 1e-3),
        "job_ms": wall,
        "kernels_ms": {"sobol_bridge": gen, "basket_asian": kern},
        "price": float(out[0]),
        "roofline": with_traffic({
            "kernel": "k_basket_asian<2> (MFMA correlation, log-Euler, exp, adjoint sums)",
            "bound": "hbm",
            "achieved": gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": gbs / HBM_PEAK_GBS,
            "traffic": None,
            "algorithmic_bytes": zbytes,
            "mfma_useful_tflops": mfma_tflops,
        }, "k_basket_asian<2>", P // 4),
    }
    with_mfma(res["roofline"], "k_basket_asian<2>", P // 4, kern)
    if not args.no_error_check:
        ws = configs.basket_asian(paths=1 << 14)
        ps = pipeline.BasketPricer(ws)
        g = ps.price_and_greeks().cpu().numpy()
        ints = O.scramble(O.sobol_ints(ws.paths, ws.dims, 0), ws.scramblers, ws.dims)
        z = O.bridge_transform(O.invnorm_u32(ints), ws.steps, ws.assets)
        m = O.basket_model(np.log(ws.spots), ws.vols, ws.chol, ws.weights, ws.rate, ws.dt,
                           ws.strike, ws.steps, ws.paths)
        pp, d, v = O.basket_asian_paths(m, z)
        g64 = np.concatenate([d, v]) / ws.paths
        res["price_rel_err_vs_fp64"] = abs(g[0] / pp.mean() - 1)
        res["greeks_max_rel_err_vs_fp64"] = float(np.max(np.abs(g[1:] - g64) / np.abs(g64)))
        res["error_sample"] = "first 2^14 paths of the workload's Sobol sequence"
    del pr
    torch.cuda.empty_cache()
    return res


def stage_section(args, torch, configs, pipeline, ev):
    """The reference-ABI stage kernels at the config-2 shape (P = 2^20,
    N = S = 40, F = 5, D = 200): achieved HBM GB/s from algorithmic bytes over
    HIP-event time (mean of 5 launches after one warm-up, each after a read
    that evicts the last-level cache), against 8 TB/s."""
    from useful_code_amd import koo
    w = configs.caplet_portfolio(paths=args.paths)
    P, N, F, S = w.paths, w.n_rates, w.factors, w.steps
    D = S * F
    dev = lambda x, dt: torch.as_tensor(np.ascontiguousarray(x)).to(dt).cuda()  # noqa: E731
    dirs = dev(koo.sobol_directions(D).view(np.int32), torch.int32)
    scr = dev(np.asarray(w.scramblers, np.uint32).view(np.int32), torch.int32)
    alloc = dev(koo.bridge_allocate_diagonal(F, S), torch.int32)
    ints = torch.empty(D * P, dtype=torch.int32, device="cuda")
    normals = torch.empty(D * P, dtype=torch.float32, device="cuda")
    z = torch.empty(D * P, dtype=torch.float32, device="cuda")
    rates = torch.empty(S * N * P, dtype=torch.float32, device="cuda")
    logs = torch.empty_like(rates)
    disc = torch.empty(S * (N + 1) * P, dtype=torch.float32, device="cuda")
    num = torch.empty(S * P, dtype=torch.float32, device="cuda")
    flows = torch.empty(S * P, dtype=torch.float32, device="cuda")
    flows2 = torch.empty(S * P, dtype=torch.float32, device="cuda")
    dflows = torch.empty(S * P, dtype=torch.float32, device="cuda")
    summed = torch.empty(P, dtype=torch.float32, device="cuda")
    r1 = torch.empty(S * P, dtype=torch.float32, device="cuda")
    A = dev(w.A, torch.float32)
    f0 = dev(w.f0, torch.float32)
    lf0 = dev(np.log(w.f0 + w.disp()), torch.float32)
    taus = dev(w.taus, torch.float32)
    dsp = dev(w.disp(), torch.float32)
    mu0 = dev(koo.spot_drifts(w.A[0], w.f0, w.taus, w.disp(), 0), torch.float32)
    fx = dev(koo.fixed_drifts(w.A), torch.float32)
    alive = dev(w.alive(), torch.int32)
    first = dev(np.arange(1, S + 1), torch.int32)
    theta = dev(np.zeros(S), torch.float32)
    aux = dev(np.concatenate([[w.strike], np.full(S, w.tau)]), torch.float32)
    ws = torch.empty(koo.sum_workspace_doubles(S * P), dtype=torch.float64, device="cuda")
    out64 = torch.empty(1, dtype=torch.float64, device="cuda")
    U = N * (N + 1) // 2  # alive rate updates per path
    four = 4.0
    # the LS normal equations of 6 basis functions and a target over P paths
    # (the reference's reduceOuterProductSymmetric_gpu on [X; y])
    kb = 6
    basis = torch.empty(kb * P, dtype=torch.float32, device="cuda").normal_()
    target = torch.empty(P, dtype=torch.float32, device="cuda").normal_()
    lsws = torch.empty(koo.ls_workspace(P), dtype=torch.uint8, device="cuda")
    gram = torch.empty((kb + 1) ** 2, dtype=torch.float64, device="cuda")
    stages = [
        ("sobolGPUInts", lambda: koo.sobolGPUInts(P, D, 0, dirs, ints), four * D * P),
        ("scramble_gpu (all D rows)", lambda: koo.scramble_gpu(ints, scr, D, P), 2 * four * D * P),
        ("inverse_normal_u32_gpu", lambda: koo.inverse_normal_u32_gpu(ints, normals),
         2 * four * D * P),
        ("brownian_bridge_gpu", lambda: koo.brownian_bridge_gpu(S, alloc, normals, z, P, F),
         2 * four * D * P),
        ("sobol_bridge_gpu (fused a+bridge)",
         lambda: koo.sobol_bridge_gpu(P, 0, S, F, dirs, scr, D, alloc, z), four * D * P),
        ("LMM_evolver_pc_all_gpu (rates, logs, discounts)",
         lambda: koo.LMM_evolver_pc_all_gpu(f0, lf0, taus, z, A, mu0, fx, dsp, alive, P, F, S, N,
                                            rates, logs, disc),
         four * (D * P + 2 * U * P + (U + S) * P)),
        ("spot_measure_numeraires_computation_gpu",
         lambda: koo.spot_measure_numeraires_computation_gpu(disc, num, P, N),
         four * (2 * S * P + S * P)),
        ("forward_rate_extraction_gpu",
         lambda: koo.forward_rate_extraction_gpu(rates, np.arange(S), P, S, N, r1),
         2 * four * S * P),
        ("cashFlowGeneratorCallerCaplet",
         lambda: koo.cash_flow_generator("caplet", flows, flows2, aux, P, S, r1),
         3 * four * S * P),
        ("cashFlowDiscounting_gpu",
         lambda: koo.cashFlowDiscounting_gpu(first, first, theta, disc, flows, num, P, S, dflows,
                                             summed),
         four * (5 * S * P + P)),
        ("koo_sum_f32_to_f64 (S*P values)", lambda: koo.sum_f32_to_f64(dflows, ws, out64),
         four * S * P),
        ("ls_normal_eq_gpu (X^T X and X^T y, 6 basis functions + target, MFMA)",
         lambda: koo.ls_normal_eq_gpu(P, kb, basis, target, None, lsws, gram), four * (kb + 1) * P),
    ]
    # a plain device-to-device copy of the same size as the quantile's cube
    # (read + write 2 * 4 B * D * P): what the memory system gives a streaming
    # kernel with nothing else to do, the practical ceiling for these stages
    copy_src = torch.empty(D * P, dtype=torch.float32, device="cuda").fill_(1.0)
    copy_dst = torch.empty_like(copy_src)
    stages.append(("device copy (torch, reference point)", lambda: copy_dst.copy_(copy_src),
                   2 * four * D * P))
    # Before each timed call a sum over 512 MiB evicts the stage's data from
    # the 256 MB last-level (Infinity) cache, so its bytes come from HBM (a
    # read, so no dirty lines are left to be written back during the timed
    # call), and keeps the device busy while the host enqueues the call, so
    # the events time the kernels and not the host's launch overhead.
    scrub = torch.ones(128 << 20, dtype=torch.float32, device="cuda")
    res = {}
    for name, fn, nbytes in stages:
        fn()
        recs = [(ev(), ev()) for _ in range(5)]
        for a, b in recs:
            scrub.sum()
            a.record()
            fn()
            b.record()
        torch.cuda.synchronize()
        ms = float(np.mean([a.elapsed_time(b) for a, b in recs]))
        gbs = nbytes / (ms * 1e-3) / 1e9
        res[name] = {"ms": ms, "algorithmic_bytes": nbytes, "GB_per_s": gbs,
                     "frac_of_hbm_peak": gbs / HBM_PEAK_GBS}
        if name.startswith("ls_normal_eq_gpu"):
            with_mfma(res[name], "k_gram<2>", P // 4, ms)
    return res


def host_cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores():
    """Cores this process may run on (its affinity mask), capped by
    OMP_NUM_THREADS when that is set."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return n


def cpu_baseline_section(args, O, configs, S):
    """The FP64 oracle on the host cores, same workload and Sobol points:
    (1) one thread on a 2^16-path prefix, scaled linearly in paths to the
    full path count (the oracle's cost per path does not depend on P);
    (2) OpenMP over paths on every core of the affinity mask at the sample
    size (default the full 2^20 paths)."""
    cores = host_cores()

    def run(paths, threads):
        wc = configs.caplet_portfolio(paths=paths)
        ints = O.scramble(O.sobol_ints(wc.paths, wc.dims, 0), wc.scramblers, wc.dims)
        m = O.caplet_model(wc.f0, wc.taus, wc.disp(), wc.A, wc.vols, wc.strike, wc.paths)
        O.set_threads(threads)
        t = time.perf_counter()
        O.caplet_pipeline(m, ints, adjoint=True)
        return time.perf_counter() - t

    p1 = min(1 << 16, args.cpu_sample_paths)
    dt1 = run(p1, 1)
    dtn = run(args.cpu_sample_paths, cores)
    O.set_threads(cores)
    return {
        "value": args.cpu_sample_paths * S / dtn,
        "unit": "path-steps/s",
        "cores": cores,
        "kind": "port",
        "cpu_model": host_cpu_model(),
        "compiler": "gcc -O3 -fopenmp (oracle/Makefile)",
        "sample": f"FP64 oracle (oracle/koo_oracle.c, OpenMP over paths on {cores} cores of the "
                  f"affinity mask): quantiles, bridge, evolution, caplets and all 80 adjoint "
                  f"Greeks of the same workload on {args.cpu_sample_paths} paths (the Sobol "
                  f"integers generated beforehand, not timed): {dtn:.2f} s",
        "single_thread": {
            "value": p1 * S / dt1,
            "unit": "path-steps/s",
            "cores": 1,
            "paths": p1,
            "seconds": dt1,
            "seconds_scaled_to_full_paths": dt1 * args.cpu_sample_paths / p1,
            "scaling": f"linear in paths: {p1} timed, x{args.cpu_sample_paths // p1} for "
                       f"{args.cpu_sample_paths}",
        },
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--paths", type=int, default=1 << 20)
    ap.add_argument("--checkpoint-every", type=int, default=8)
    ap.add_argument("--cpu-sample-paths", type=int, default=1 << 20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-error-check", action="store_true")
    ap.add_argument("--no-bermudan", action="store_true")
    ap.add_argument("--no-basket", action="store_true")
    ap.add_argument("--no-swap", action="store_true")
    ap.add_argument("--no-portfolio", action="store_true")
    ap.add_argument("--no-multi-ls", action="store_true")
    ap.add_argument("--portfolio-paths", type=int, default=1 << 17)
    ap.add_argument("--basket-paths", type=int, default=1 << 20)
    ap.add_argument("--no-stages", action="store_true",
                    help="skip the per-stage HBM measurements of the reference-ABI kernels")
    ap.add_argument("--no-k1", action="store_true",
                    help="skip the checkpoint-every-step comparison")
    ap.add_argument("--bermudan-paths", type=int, default=1 << 18)
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
    if w.checkpoint_every != 1 and not args.no_k1:
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
        "roofline": with_traffic({
            "kernel": f"k_price_bwd<{F}> (reverse sweep + segment recompute)",
            "bound": "valu",
            "achieved": achieved_tflops,
            "peak": FP32_VECTOR_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / FP32_VECTOR_PEAK_TFLOPS,
            "traffic": None,
        }, f"k_price_bwd<{F}>", P),
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
        out["cpu_baseline"] = cpu_baseline_section(args, O, configs, S)

    if not args.no_bermudan:
        out["bermudan_ls"] = bermudan_section(args, torch, O, configs, pipeline, ev)

    if not args.no_multi_ls:
        out["bermudan_multi_ls"] = multi_ls_section(args, torch, O, configs, pipeline, ev)
        torch.cuda.empty_cache()

    if not args.no_portfolio:
        out["bermudan_portfolio"] = portfolio_section(args, torch, configs, pipeline, ev)
        torch.cuda.empty_cache()

    if not args.no_swap:
        out["payer_swap_config1"] = swap_section(args, torch, O, configs, pipeline, ev)

    if not args.no_basket:
        out["basket_asian"] = basket_section(args, torch, O, configs, pipeline, ev)

    if not args.no_stages:
        del pr
        torch.cuda.empty_cache()
        out["stage_kernels_hbm"] = stage_section(args, torch, configs, pipeline, ev)

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
