"""Host-side set-up of the product (C++ in libkoo_amd.so) against the oracle
and the reference gold fixtures: direction numbers, bridge allocation,
correlation, PCA pseudo-roots and drifts."""
import os

import numpy as np

import oracle as O
from useful_code_amd import configs, koo

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_direction_numbers_match_reference_gold():
    from test_oracle_golden import sobol_dims_fold
    g = np.load(os.path.join(G, "sobol_ref.npz"))
    assert np.array_equal(koo.sobol_directions(210), g["dirs210"])
    assert np.array_equal(koo.sobol_directions(2048), g["dirs2048"])
    assert np.array_equal(sobol_dims_fold(koo.sobol_directions(10200)), g["dirs_fold_all"])


def test_bridge_allocation_matches_oracle():
    for F, S in [(1, 1), (3, 20), (5, 40), (32, 64)]:
        assert np.array_equal(koo.bridge_allocate_diagonal(F, S), O.bridge_allocate_diagonal(F, S))


def test_pseudo_roots_match_reference_gold():
    p = np.load(os.path.join(G, "pseudo_roots_ref.npz"))
    N, _, F = p["A"].shape
    A = koo.flat_vol_pseudo_roots(p["rate_times"], p["rate_times"][:N], p["vols"], F, 0.0, 0.1)
    for s in range(N):
        assert np.allclose(A[s] @ A[s].T, p["A"][s] @ p["A"][s].T, rtol=1e-9, atol=1e-15)
        assert np.allclose(np.abs(A[s]), np.abs(p["A"][s]), rtol=1e-7, atol=1e-12)
    # same sign convention as the oracle: identical factors
    Ao = O.flat_vol_pseudo_roots(p["rate_times"], p["rate_times"][:N], p["vols"], F, 0.0, 0.1)
    assert np.allclose(A, Ao, atol=1e-12)


def test_factor_pseudo_roots_match_oracle_and_total_variance():
    rt = 0.5 * np.arange(1, 42)
    vols = 0.15 + 0.1 * np.exp(-0.2 * rt[:40])
    A, B = koo.factor_pseudo_roots(rt, vols, 5, 0.0, 0.05)
    Ao, Bo = O.factor_pseudo_roots(rt, vols, 5, 0.0, 0.05)
    assert np.allclose(A, Ao, atol=1e-12) and np.allclose(B, Bo, atol=1e-12)
    # sum over steps s <= i of |A(s,i,.)|^2 = sigma_i^2 T_i (Black total variance)
    tv = np.array([np.sum(A[: i + 1, i, :] ** 2) for i in range(40)])
    assert np.allclose(tv, vols**2 * rt[:40], rtol=1e-12)


def test_drifts_match_reference_gold():
    g = np.load(os.path.join(G, "lmm_gold_small.npz"))
    assert np.allclose(koo.spot_drifts(g["A"][0], g["f0"], g["taus"], g["disp"]), g["mu0"],
                       rtol=1e-14, atol=1e-18)
    assert np.allclose(koo.fixed_drifts(g["A"]), g["fixed"], rtol=1e-15)


def test_caplet_workload_is_reproducible():
    a = configs.caplet_portfolio(paths=1024)
    b = configs.caplet_portfolio(paths=1024)
    assert np.array_equal(a.f0, b.f0) and np.array_equal(a.scramblers, b.scramblers)
    assert a.A.shape == (40, 40, 5) and a.dims == 200
