"""Host orchestration of the LMM pricing path on one GPU.

``LMMPricer`` owns the device copies of a workload and runs
  * ``generate``: fused Sobol -> scramble -> normal -> bridge (one kernel),
  * ``price``: the fused forward price kernel (no rate cube written),
  * ``price_and_greeks``: forward with checkpoints plus the reverse sweep.
``BermudanPricer`` runs the Longstaff-Schwartz path (fused evolution to the
exercise quantities, backward induction with MFMA normal equations and FP64
Cholesky, and the frozen-strategy pass).
``lmm_evolution_routine`` mirrors the reference's staged pipeline
(LMMEvolutionRoutineRawSingleKernel, LMM_evolver_full.cu:264-425) through the
reference-named entry points.  Buffers are allocated once per pricer; calls
only launch kernels on the current stream.
"""
import numpy as np
import torch

from . import koo


def _dev(x, dtype):
    return torch.as_tensor(np.ascontiguousarray(x), dtype=dtype).cuda()


class LMMPricer:
    def __init__(self, w, device="cuda"):
        self.w = w
        P, N, F, S = w.paths, w.n_rates, w.factors, w.steps
        self.dirs = _dev(koo.sobol_directions(w.dims).view(np.int32), torch.int32)
        self.scr = _dev(np.asarray(w.scramblers, np.uint32).view(np.int32), torch.int32)
        self.alloc = _dev(koo.bridge_allocate_diagonal(F, S), torch.int32)
        self.f0 = _dev(w.f0, torch.float32)
        self.taus = _dev(w.taus, torch.float32)
        self.disp = _dev(w.disp(), torch.float32)
        self.A = _dev(w.A, torch.float32)
        self.vols = _dev(w.vols, torch.float32)
        self.z = torch.empty(S * F * P, dtype=torch.float32, device=device)
        ck, sc, pa = koo.lmm_price_workspace(P, N, F, w.checkpoint_every)
        self.ckpt = torch.empty(ck // 4, dtype=torch.float32, device=device)
        self.scratch = torch.empty(sc // 4, dtype=torch.float32, device=device)
        self.partials = torch.empty(pa // 8, dtype=torch.float64, device=device)
        self.price_out = torch.zeros(1, dtype=torch.float64, device=device)
        self.greeks_out = torch.zeros(2 * N, dtype=torch.float64, device=device)
        self.aux = np.array([w.strike], np.float32)

    def generate(self):
        w = self.w
        koo.sobol_bridge_gpu(w.paths, w.path_offset, w.steps, w.factors, self.dirs, self.scr,
                             w.dims, self.alloc, self.z)

    def price_kernel(self, checkpoints=False):
        w = self.w
        koo.lmm_price_fwd(w.paths, w.n_rates, w.factors, w.checkpoint_every, self.z, self.f0,
                          self.taus, self.disp, self.A, koo.KOO_PRODUCT_CAPLETS, self.aux,
                          self.ckpt if checkpoints else None, self.partials, self.price_out)

    def greeks_kernel(self):
        w = self.w
        koo.lmm_price_bwd(w.paths, w.n_rates, w.factors, w.checkpoint_every, self.z, self.f0,
                          self.taus, self.disp, self.A, self.vols, koo.KOO_PRODUCT_CAPLETS,
                          self.aux, self.ckpt, self.scratch, self.partials, self.greeks_out)

    def price(self):
        self.generate()
        self.price_kernel(checkpoints=False)
        return self.price_out

    def price_and_greeks(self):
        self.generate()
        self.price_kernel(checkpoints=True)
        self.greeks_kernel()
        return self.price_out, self.greeks_out


class BermudanPricer:
    """Bermudan swaption by Longstaff-Schwartz on one GPU.

    ``generate(offset)``: variates of Sobol points offset+1 ...; ``paths_kernel``:
    basis variables, exercise values and numeraires per exercise date;
    ``regress``: first pass (coefficients and in-sample price); ``second_pass``:
    price under the frozen coefficients on the current paths."""

    def __init__(self, w, device="cuda"):
        self.w = w
        P, N, F, S = w.paths, w.n_rates, w.factors, w.steps
        E = len(w.exercise_steps)
        self.E = E
        self.dirs = _dev(koo.sobol_directions(w.dims).view(np.int32), torch.int32)
        self.scr = _dev(np.asarray(w.scramblers, np.uint32).view(np.int32), torch.int32)
        self.alloc = _dev(koo.bridge_allocate_diagonal(F, S), torch.int32)
        self.f0 = _dev(w.f0, torch.float32)
        self.taus = _dev(w.taus, torch.float32)
        self.disp = _dev(w.disp(), torch.float32)
        self.A = _dev(w.A, torch.float32)
        self.z = torch.empty(S * F * P, dtype=torch.float32, device=device)
        self.bv = torch.empty(E * 2 * P, dtype=torch.float32, device=device)
        self.ev = torch.empty(E * P, dtype=torch.float32, device=device)
        self.num = torch.empty(E * P, dtype=torch.float32, device=device)
        self.ws = torch.empty(koo.ls_workspace(P), dtype=torch.uint8, device=device)
        self.coeffs = torch.zeros(E * 6, dtype=torch.float64, device=device)
        self.price1 = torch.zeros(1, dtype=torch.float64, device=device)
        self.price2 = torch.zeros(1, dtype=torch.float64, device=device)
        self.ves = np.full(E, 2, np.int32)
        self.shift = np.asarray(w.basis_shift, np.float32)
        self.scale = np.asarray(w.basis_scale, np.float32)

    def generate(self, offset=None):
        w = self.w
        koo.sobol_bridge_gpu(w.paths, w.path_offset if offset is None else offset, w.steps,
                             w.factors, self.dirs, self.scr, w.dims, self.alloc, self.z)

    def paths_kernel(self):
        w = self.w
        koo.bermudan_paths_gpu(w.paths, w.n_rates, w.factors, self.z, self.f0, self.taus,
                               self.disp, self.A, w.exercise_steps, w.strike, self.bv, self.ev,
                               self.num)

    def regress(self):
        w = self.w
        koo.ls_first_pass_gpu(self.E, self.bv, self.ves, 2, None, self.ev, self.num, w.paths, True,
                              True, self.shift, self.scale, self.ws, self.coeffs, self.price1)

    def second_pass(self):
        w = self.w
        koo.ls_second_pass_gpu(self.E, self.bv, self.ves, 2, None, self.ev, self.num, w.paths, True,
                               True, self.shift, self.scale, self.coeffs, self.ws, self.price2)

    def price(self):
        """First-pass (in-sample) price, then the frozen-strategy price on the
        next block of Sobol points; returns both as device tensors."""
        self.generate()
        self.paths_kernel()
        self.regress()
        self.generate(self.w.path_offset + self.w.paths)
        self.paths_kernel()
        self.second_pass()
        return self.price1, self.price2


class BasketPricer:
    """Basket Asian on one GPU: the fused generator (Sobol -> scramble ->
    normal -> bridge) into (steps, assets, paths) variates, then the fused
    path/payoff/adjoint kernel.  ``price_and_greeks`` returns the device
    vector [price, deltas (A), vegas (A)]."""

    def __init__(self, w, device="cuda"):
        self.w = w
        P, A, S = w.paths, w.assets, w.steps
        self.dirs = _dev(koo.sobol_directions(w.dims).view(np.int32), torch.int32)
        self.scr = _dev(np.asarray(w.scramblers, np.uint32).view(np.int32), torch.int32)
        self.alloc = _dev(koo.bridge_allocate_diagonal(A, S), torch.int32)
        self.z = torch.empty(S * A * P, dtype=torch.float32, device=device)
        self.x0 = _dev(np.log(w.spots), torch.float32)
        self.vols = _dev(w.vols, torch.float32)
        self.lt = _dev(w.scaled_chol(), torch.float32)
        self.weights = _dev(w.weights, torch.float32)
        self.partials = torch.empty(koo.basket_asian_workspace(P, A) // 8, dtype=torch.float64,
                                    device=device)
        self.out = torch.zeros(1 + 2 * A, dtype=torch.float64, device=device)

    def generate(self):
        w = self.w
        koo.sobol_bridge_gpu(w.paths, w.path_offset, w.steps, w.assets, self.dirs, self.scr,
                             w.dims, self.alloc, self.z)

    def greeks_kernel(self):
        w = self.w
        koo.basket_asian_gpu(w.paths, w.assets, w.steps, self.z, self.x0, self.vols, self.lt,
                             self.weights, w.rate, w.dt, w.strike, self.partials, self.out)

    def price_and_greeks(self):
        self.generate()
        self.greeks_kernel()
        return self.out


def lmm_evolution_routine(w, use_fused_generator=False):
    """Staged pipeline of LMMEvolutionRoutineRawSingleKernel: Sobol integers
    for D = S*F dimensions, scramble of the first S rows (as the reference
    driver does, LMM_evolver_full.cu:324-327), inverse normal, diagonal-
    allocation bridge, then the all-steps PC evolver with discount ratios.
    Returns (rates, log_rates, discounts) as device tensors (S,N,P),
    (S,N,P), (S,N+1,P)."""
    P, N, F, S = w.paths, w.n_rates, w.factors, w.steps
    D = S * F
    dirs = _dev(koo.sobol_directions(D).view(np.int32), torch.int32)
    scr = _dev(np.asarray(w.scramblers, np.uint32).view(np.int32), torch.int32)
    normals = torch.empty(D * P, dtype=torch.float32, device="cuda")
    if use_fused_generator:
        koo.sobol_scramble_normal_gpu(P, D, w.path_offset, dirs, scr, S, normals)
    else:
        ints = torch.empty(D * P, dtype=torch.int32, device="cuda")
        koo.sobolGPUInts(P, D, w.path_offset, dirs, ints)
        koo.scramble_gpu(ints, scr, S, P)
        koo.inverse_normal_u32_gpu(ints, normals)
    alloc = _dev(koo.bridge_allocate_diagonal(F, S), torch.int32)
    z = torch.empty(D * P, dtype=torch.float32, device="cuda")
    koo.brownian_bridge_gpu(S, alloc, normals, z, P, F)
    A = w.A.astype(np.float32)
    mu0 = koo.spot_drifts(w.A[0], w.f0, w.taus, w.disp(), 0)
    fx = koo.fixed_drifts(w.A)
    rates = torch.empty(S * N * P, dtype=torch.float32, device="cuda")
    logs = torch.empty_like(rates)
    disc = torch.empty(S * (N + 1) * P, dtype=torch.float32, device="cuda")
    koo.LMM_evolver_pc_all_gpu(_dev(w.f0, torch.float32), _dev(np.log(w.f0 + w.disp()), torch.float32),
                               _dev(w.taus, torch.float32), z, _dev(A, torch.float32),
                               _dev(mu0, torch.float32), _dev(fx, torch.float32),
                               _dev(w.disp(), torch.float32), _dev(w.alive(), torch.int32), P, F,
                               S, N, rates, logs, disc)
    return rates.view(S, N, P), logs.view(S, N, P), disc.view(S, N + 1, P)
