"""Test helper: prices built from the reference's own gold routines (compiled
from its sources into oracle/_ref by oracle/Makefile), used as a second,
independent check of the oracle and of the GPU results."""
import numpy as np

import oracle as O


def gold_caplet_price(w, z, f0=None, A=None):
    """The caplet strip priced by the reference's own gold routines (built
    from its sources into oracle/_ref) on the bridged variates z (S, F, P):
    correlate_drift_paths_gold with the fixed drifts -1/2 sum_f A^2
    (gold/LMM_evolver_classes_gold.h:193-205), LMM_evolver_pc_gold,
    discount_ratios_computation_gold, spot_measure_numeraires_computation_offset_gold,
    forward_rate_extraction_gold (rate s at step s) and cashFlowDiscounting_gold
    (paid at T_{s+1}).  Only the payoff tau (f - K)^+ is not the reference's:
    it has no caplet product."""
    R = O.ref()
    f0 = w.f0 if f0 is None else f0
    A = w.A if A is None else A
    S, N, F, P = w.steps, w.n_rates, w.factors, w.paths
    taus, disp = np.full(N, w.tau), np.ascontiguousarray(w.disp(), np.float64)
    zt = np.ascontiguousarray(np.transpose(z.reshape(S, F, P), (0, 2, 1))).reshape(-1)
    fixed = np.ascontiguousarray(-0.5 * np.sum(A * A, axis=2)).reshape(-1)
    Af = np.ascontiguousarray(A).reshape(-1)
    incr = np.zeros(S * N * P)
    R.ref_correlate_drift(zt, incr, Af, N * F, fixed, F, N, P, S)
    rates, logs = np.zeros(S * N * P), np.zeros(S * N * P)
    R.ref_lmm_pc(np.ascontiguousarray(f0, np.float64), taus, incr, Af, disp, P, F, S, N, rates,
                 logs)
    disc = np.zeros(S * (N + 1) * P)
    R.ref_discount_ratios(rates, taus, np.arange(S, dtype=np.int32), P, S, N, disc)
    num = np.zeros(S * P)
    R.ref_numeraires(disc, P, N, num)
    r1 = np.zeros(S * P)
    R.ref_forward_extraction(rates, np.arange(S, dtype=np.int32), P, S, N, r1)
    flows = np.ascontiguousarray(taus[:S, None] * np.maximum(r1.reshape(S, P) - w.strike, 0.0))
    first = np.arange(1, S + 1, dtype=np.int32)
    out, summed = np.zeros(S * P), np.zeros(P)
    R.ref_cash_flow_discounting(first, first, np.zeros(S), disc, flows.reshape(-1), num, P, S, out,
                                summed)
    return summed.mean()
