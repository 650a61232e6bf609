"""Autotuner for the batched Bermudan portfolio (measured on the device).

For every product it times, with HIP events, the frozen-strategy forward and
reverse sweep at each candidate checkpoint interval k (1, 2, 4, 8, 16, bounded
by a checkpoint-memory budget) and keeps the fastest; then it times the
product's whole sequence at that k and schedules the products over the
streams by longest-processing-time-first (each product goes to the stream
with the least accumulated time).  The result -- k and stream per product and
the measured times -- is a plain dict that PortfolioPricer takes and that can
be stored as JSON next to the results it produced.
"""
import json

import numpy as np
import torch

from . import koo

CANDIDATE_K = (1, 2, 4, 8, 16)


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


def tune_portfolio(pricer, memory_budget_bytes=64 << 30, n_streams=None):
    """pricer: a PortfolioPricer (its buffers are reused).  Returns
    {"k": [...], "stream": [...], "ms": [...], "k_ms": [{k: ms}, ...]}."""
    from .pipeline import BermudanPricer  # noqa: F401  (type of pricer.pricers)
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
    load = [0.0] * n_streams
    stream = [0] * len(ms)
    for j in sorted(range(len(ms)), key=lambda j: -ms[j]):
        s = int(np.argmin(load))
        stream[j] = s
        load[s] += ms[j]
    return {"k": ks, "stream": stream, "ms": ms, "k_ms": [{str(a): b for a, b in t.items()}
                                                          for t in k_ms],
            "stream_load_ms": load}


def save(tuning, path):
    with open(path, "w") as f:
        json.dump(tuning, f, indent=1)


def load(path):
    with open(path) as f:
        return json.load(f)
