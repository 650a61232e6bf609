"""Autotuner for the Bermudan portfolio (measured on the device).

Per-product mode: for every product it times, with HIP events, the
frozen-strategy forward and reverse sweep at each candidate checkpoint
interval k (1, 2, 4, 8, 16, bounded by a checkpoint-memory budget) and keeps
the fastest; then it times the product's whole sequence at that k and
schedules the products over the streams by longest-processing-time-first
(each product goes to the stream with the least accumulated time).

Batched mode: the candidate intervals are timed on the batched forward and
backward launches of the whole portfolio (the products share the device, so
the sum is what counts); the LS tile -- 1024-lane blocks per product in the
batched regression launches, a share of 1, 2, 4 or 8 blocks per CU over the
portfolio -- is timed on the batched path-quantity and regression launches;
and the generators are scheduled over the streams by LPT.

Persistent launch (the whole portfolio in one kernel, a team of resident
blocks per product): with the batched mode's interval, the teams are sized
by the library's work model, the launch is timed, and the teams are resized
in proportion to each product's measured block-time (the kernel records
every block's run time) for a few rounds, keeping the fastest sizing; then
the persistent launch and the batched launches are compared and the faster
one is what the pricer runs.

The result -- k, stream and LS tile per product, the team weights, the
launch form and the measured times -- is a plain dict that PortfolioPricer
takes and that can be stored as JSON next to the results it produced.
"""
import json

import numpy as np
import torch

from . import koo

CANDIDATE_K = (1, 2, 4, 8, 16)
CANDIDATE_LS_BLOCKS_PER_CU = (1, 2, 4, 8)


def _time(fn, reps=3):
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(reps)]
    fn()
    torch.cuda.synchronize()
    for a, b in ev:
        a.record()
        fn()
        b.record()
    torch.cuda.synchronize()
    return float(np.median([a.elapsed_time(b) for a, b in ev]))


def checkpoint_bytes(w, k):
    ck, sc, pb, _ = koo.bermudan_adjoint_workspace(w.paths, w.n_rates, w.factors, k)
    return ck + sc + pb


def _lpt(ms, n_streams):
    load = [0.0] * n_streams
    stream = [0] * len(ms)
    for j in sorted(range(len(ms)), key=lambda j: -ms[j]):
        s = int(np.argmin(load))
        stream[j] = s
        load[s] += ms[j]
    return stream, load


def tune_batched(pricer, memory_budget_bytes=64 << 30, n_streams=None):
    """Batched portfolio: one checkpoint interval for every product, the one
    whose batched forward + backward launches are fastest (the products share
    the device, so their sum is what counts); generators scheduled over the
    streams by LPT on their measured times."""
    n_streams = n_streams or len(pricer.streams)
    for pr in pricer.pricers:
        pr.generate()
    pricer.k = [1] * len(pricer.pricers)
    pricer.plan = None
    pricer._issue()  # path quantities and regressions, so the strategy is realistic
    torch.cuda.synchronize()
    k_ms = {}
    for k in CANDIDATE_K:
        if any(k > pr.w.n_rates for pr in pricer.pricers):
            continue
        if sum(checkpoint_bytes(pr.w, k) for pr in pricer.pricers) > memory_budget_bytes:
            continue
        pricer.k = [k] * len(pricer.pricers)
        pricer._prepare_batch()
        k_ms[k] = _time(lambda: (koo.bermudan_portfolio_fwd(pricer.plan, pricer.descriptors),
                                 koo.bermudan_portfolio_bwd(pricer.plan, pricer.descriptors)))
        for pr in pricer.pricers:  # release this interval's buffers
            del pr.adj
        torch.cuda.empty_cache()
    best = min(k_ms, key=k_ms.get)
    # LS tile: 1024-lane blocks per product of the batched LS date launches,
    # a share of the CUs over the portfolio (2 per CU is the library default)
    n = len(pricer.pricers)
    cus = torch.cuda.get_device_properties(torch.cuda.current_device()).multi_processor_count
    ls_ms = {}
    for per_cu in CANDIDATE_LS_BLOCKS_PER_CU:
        pricer.ls_blocks = [max(1, -(-per_cu * cus // n))] * n
        pricer.k = [best] * n
        pricer._prepare_batch()
        ls_ms[per_cu] = _time(lambda: koo.bermudan_ls_portfolio_run(pricer.ls_plan,
                                                                    pricer.ls_descriptors))
    best_ls = min(ls_ms, key=ls_ms.get)
    ls_blocks = [max(1, -(-best_ls * cus // n))] * n
    pricer.ls_blocks = ls_blocks
    pricer.plan = None
    ms = [_time(pr.generate, reps=2) for pr in pricer.pricers]
    stream, load = _lpt(ms, n_streams)
    return {"k": [best] * n, "stream": stream, "ms": ms,
            "k_ms": [{str(a): b for a, b in k_ms.items()}] * n,
            "ls_blocks": ls_blocks, "ls_blocks_per_cu_ms": {str(a): b for a, b in ls_ms.items()},
            "stream_load_ms": load, "batched": True}


def tune_persistent(pricer, rounds=3):
    """Team weights of the persistent launch: the model's first, then each
    round the measured block-time per product of the previous run; returns
    (best ms, its weights, [ms per round])."""
    pricer.work = None
    best, best_w, hist = None, None, []
    for _ in range(rounds):
        pricer.pplan = None
        pricer.graph = None
        ms = _time(pricer._issue, reps=2)
        hist.append(ms)
        if best is None or ms < best:
            best, best_w = ms, pricer.work
        pricer.work = pricer.measured_work()
    pricer.work = best_w
    pricer.pplan = None
    return best, best_w, hist


def tune_portfolio(pricer, memory_budget_bytes=64 << 30, n_streams=None):
    """pricer: a PortfolioPricer (its buffers are reused).  Returns
    {"k": [...], "stream": [...], "ms": [...], "k_ms": [{k: ms}, ...]}.
    A batched pricer is tuned by tune_batched, and a persistent one also by
    tune_persistent, which decides between the two launch forms."""
    if getattr(pricer, "batched", False):
        persistent = getattr(pricer, "persistent", False)
        pricer.persistent = False
        res = tune_batched(pricer, memory_budget_bytes, n_streams)
        if persistent:
            pricer.k = list(res["k"])
            pricer.ls_blocks = list(res["ls_blocks"])
            pricer.stream_of = list(res["stream"])
            pricer.plan = None
            batched_ms = _time(pricer._issue, reps=2)
            pricer.persistent = True
            ms, work, hist = tune_persistent(pricer)
            res.update({"persistent": ms < batched_ms, "work": work,
                        "persistent_ms": ms, "persistent_rounds_ms": hist,
                        "batched_ms": batched_ms})
        pricer.persistent = persistent
        return res
    n_streams = n_streams or len(pricer.streams)
    ks, k_ms = [], []
    budget_each = memory_budget_bytes / max(1, len(pricer.pricers))
    for pr in pricer.pricers:
        w = pr.w
        pr.generate()
        pr.paths_kernel()
        pr.regress()
        times = {}
        for k in CANDIDATE_K:
            if k > w.n_rates or checkpoint_bytes(w, k) > budget_each:
                continue
            times[k] = _time(lambda: pr.greeks(checkpoint_every=k))
        if not times:  # the largest interval always fits the least memory
            k = max(c for c in CANDIDATE_K if c <= w.n_rates)
            times[k] = _time(lambda: pr.greeks(checkpoint_every=k))
        best = min(times, key=times.get)
        ks.append(best)
        k_ms.append(times)
        # release the buffers of the losing intervals before the next product
        pr.greeks(checkpoint_every=best)
        torch.cuda.empty_cache()
    ms = []
    for pr, k in zip(pricer.pricers, ks):
        def seq(pr=pr, k=k):
            pr.generate()
            pr.paths_kernel()
            pr.regress()
            pr.greeks(checkpoint_every=k)
        ms.append(_time(seq, reps=2))
    stream, load = _lpt(ms, n_streams)
    return {"k": ks, "stream": stream, "ms": ms, "k_ms": [{str(a): b for a, b in t.items()}
                                                          for t in k_ms],
            "stream_load_ms": load}


def save(tuning, path):
    with open(path, "w") as f:
        json.dump(tuning, f, indent=1)


def load(path):
    with open(path) as f:
        return json.load(f)
