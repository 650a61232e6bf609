"""Synthetic workloads of the LMM pricing path (host-side inputs, FP64).

All randomness is seeded: numpy's MT19937 (``np.random.RandomState``) draws
the initial curve perturbations and the 32-bit scramblers, so a workload is
reproducible from its seed.  Rate times are T_i = tau * (i + 1), i = 0..N
(rate i accrues over [T_i, T_{i+1}]); evolution times are the reset times
T_0..T_{N-1}, so steps == rates and alive(s) = s.
"""
from dataclasses import dataclass, field

import numpy as np

from . import koo


@dataclass
class LMMWorkload:
    name: str
    paths: int
    n_rates: int
    factors: int
    tau: float
    f0: np.ndarray
    vols: np.ndarray
    rate_times: np.ndarray  # N + 1
    A: np.ndarray  # (S, N, F) FP64
    strike: float
    product: str
    scramblers: np.ndarray  # uint32, one per Sobol dimension
    checkpoint_every: int = 8
    displacements: np.ndarray = field(default=None)
    path_offset: int = 0
    exercise_steps: np.ndarray = field(default=None)  # Bermudan: exercise evolution steps
    basis_shift: np.ndarray = field(default=None)  # (E, 2): LS variables are (x - shift) * scale
    basis_scale: np.ndarray = field(default=None)

    @property
    def steps(self):
        return self.n_rates

    @property
    def dims(self):
        return self.steps * self.factors

    @property
    def taus(self):
        return np.full(self.n_rates, self.tau)

    def disp(self):
        return np.zeros(self.n_rates) if self.displacements is None else self.displacements

    def alive(self):
        return np.arange(self.steps, dtype=np.int32)


def _scramblers(rng, dims):
    return rng.randint(0, 2**32, size=dims, dtype=np.uint64).astype(np.uint32)


def caplet_portfolio(paths=2**20, seed=2024, checkpoint_every=8):
    """LMM price plus adjoint Greeks: 40 forwards, tau = 0.5,
    f0_i = 0.03 + 0.0005 i + N(0, 0.001), sigma_i = 0.15 + 0.1 exp(-0.2 T_i),
    5 PCA factors of exp(-0.05 |T_i - T_j|), 40 caplets at K = 0.045,
    2^20 Sobol paths, 40 steps, checkpoint every 8 steps."""
    N, F, tau = 40, 5, 0.5
    rng = np.random.RandomState(seed)
    f0 = 0.03 + 0.0005 * np.arange(N) + rng.normal(0.0, 0.001, N)
    rate_times = tau * np.arange(1, N + 2)
    vols = 0.15 + 0.1 * np.exp(-0.2 * rate_times[:N])
    A, _ = koo.factor_pseudo_roots(rate_times, vols, F, 0.0, 0.05)
    return LMMWorkload("lmm_caplets_40x5", paths, N, F, tau, f0, vols, rate_times, A, 0.045,
                       "caplets", _scramblers(rng, N * F), checkpoint_every)


def payer_swap(paths=2**15, seed=2023):
    """Plain LMM, CPU-runnable case: 20 forwards, tau = 0.5,
    f0 = 0.05 + U(-0.005, 0.005), flat vol 0.20, exp(-0.1 |T_i - T_j|)
    reduced to 3 factors by PCA per step, 2^15 paths, payer swap at 0.05."""
    N, F, tau = 20, 3, 0.5
    rng = np.random.RandomState(seed)
    f0 = 0.05 + rng.uniform(-0.005, 0.005, N)
    rate_times = tau * np.arange(1, N + 2)
    vols = np.full(N, 0.20)
    A = koo.flat_vol_pseudo_roots(rate_times, rate_times[:N], vols, F, 0.0, 0.1)
    return LMMWorkload("lmm_payer_swap_20x3", paths, N, F, tau, f0, vols, rate_times, A, 0.05,
                       "swap", _scramblers(rng, N * F))


def bermudan_swaption(paths=2**18, seed=2025):
    """Bermudan payer swaption by Longstaff-Schwartz: 20 forwards, tau = 0.5,
    f0 = 0.05 + U(-0.005, 0.005), flat vol 0.2, exp(-0.1 |T_i - T_j|) reduced
    to 3 factors by PCA per step, 2^18 Sobol paths.  Exercise into the swap to
    T_20 at strike 0.05 on 10 dates, the resets T_0, T_2, ..., T_18 (steps
    0, 2, ..., 18).  Basis: quadratic with cross term in the co-terminal swap
    rate and the first alive forward (6 functions), regressed on
    in-the-money paths; the GPU works in the variables (x - 0.05) * 100."""
    N, F, tau = 20, 3, 0.5
    rng = np.random.RandomState(seed)
    f0 = 0.05 + rng.uniform(-0.005, 0.005, N)
    rate_times = tau * np.arange(1, N + 2)
    vols = np.full(N, 0.20)
    A = koo.flat_vol_pseudo_roots(rate_times, rate_times[:N], vols, F, 0.0, 0.1)
    ex = np.arange(0, N, 2, dtype=np.int32)
    K = 0.05
    return LMMWorkload("bermudan_payer_20x3", paths, N, F, tau, f0, vols, rate_times, A, K,
                       "bermudan", _scramblers(rng, N * F), exercise_steps=ex,
                       basis_shift=np.full((ex.size, 2), K), basis_scale=np.full((ex.size, 2), 100.0))


def bermudan_portfolio(n_products=64, paths=2**17, seed=2028):
    """Batched callable portfolio: Bermudan payer swaptions with 10-40 rates
    (uniform from seed), 4-20 exercise dates (evenly spread over the resets up
    to the second-to-last, fewer when the deal is short), strikes U(0.03, 0.07),
    1-5 factors, flat vols U(0.15, 0.25), forwards 0.05 + U(-0.005, 0.005),
    2^17 Sobol paths each.  LS variables are normalised as (x - 0.05) * 100."""
    rng = np.random.RandomState(seed)
    prods = []
    for j in range(n_products):
        N = int(rng.randint(10, 41))
        E = int(min(rng.randint(4, 21), N - 1))
        K = float(rng.uniform(0.03, 0.07))
        F = int(min(rng.randint(1, 6), N))
        f0 = 0.05 + rng.uniform(-0.005, 0.005, N)
        vols = np.full(N, rng.uniform(0.15, 0.25))
        rate_times = 0.5 * np.arange(1, N + 2)
        A = koo.flat_vol_pseudo_roots(rate_times, rate_times[:N], vols, F, 0.0, 0.1)
        ex = np.unique(np.round(np.linspace(0, N - 2, E)).astype(np.int32))
        prods.append(LMMWorkload(f"bermudan_{j}_{N}x{F}_e{ex.size}", paths, N, F, 0.5, f0, vols,
                                 rate_times, A, K, "bermudan", _scramblers(rng, N * F),
                                 exercise_steps=ex, basis_shift=np.full((ex.size, 2), 0.05),
                                 basis_scale=np.full((ex.size, 2), 100.0)))
    return prods


@dataclass
class BasketWorkload:
    """Multi-asset Black-Scholes basket; variates are (steps, assets, paths)."""
    name: str
    paths: int
    assets: int
    steps: int
    maturity: float
    rate: float
    spots: np.ndarray
    vols: np.ndarray
    corr: np.ndarray
    chol: np.ndarray  # lower Cholesky factor of corr
    weights: np.ndarray
    strike: float
    scramblers: np.ndarray
    path_offset: int = 0

    @property
    def dt(self):
        return self.maturity / self.steps

    @property
    def factors(self):
        return self.assets

    @property
    def dims(self):
        return self.steps * self.assets

    def scaled_chol(self):
        """C[a][f] = vol_a sqrt(dt) L[a][f]."""
        return self.vols[:, None] * np.sqrt(self.dt) * self.chol


def basket_asian(paths=2**20, seed=2027, assets=32, steps=64):
    """Multi-dimensional Black-Scholes: 32 assets, spots 100 exp(N(0, 0.1)),
    vols U(0.15, 0.35), rate 0.03, correlation 0.3 constant plus a seeded
    rank-2 perturbation (unit diagonal), Cholesky-factored; 64 Sobol-bridge
    steps over one year, 2^20 paths; arithmetic-average basket Asian call at
    K = 100 with equal weights; 32 deltas and 32 vegas by adjoint."""
    rng = np.random.RandomState(seed)
    spots = 100.0 * np.exp(rng.normal(0.0, 0.1, assets))
    vols = rng.uniform(0.15, 0.35, assets)
    V = rng.normal(0.0, 0.2, size=(assets, 2))
    C = 0.3 * np.ones((assets, assets)) + 0.7 * np.eye(assets) + V @ V.T
    d = np.sqrt(np.diag(C))
    corr = C / np.outer(d, d)
    chol = koo.cholesky_lower(corr)
    return BasketWorkload("basket_asian_32x64", paths, assets, steps, 1.0, 0.03, spots, vols, corr,
                          chol, np.full(assets, 1.0 / assets), 100.0,
                          _scramblers(rng, steps * assets))


def bridge_allocation(w):
    return koo.bridge_allocate_diagonal(w.factors, w.steps)
