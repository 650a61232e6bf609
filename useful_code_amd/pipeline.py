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

    def job(self, capture=True):
        """The whole pricing (generate, path quantities, first pass, second
        pass).  With ``capture`` the first call runs it eagerly once (lazy
        module loads), records it into a HIP graph and every call replays the
        graph: one launch of the graph instead of 16 kernel launches (with
        10 exercise dates)."""
        if not capture:
            self.generate()
            self.paths_kernel()
            self.regress()
            self.second_pass()
            return
        if getattr(self, "graph", None) is None:
            self.job(capture=False)
            torch.cuda.synchronize()
            self.graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.graph):
                self.job(capture=False)
        self.graph.replay()

    def adjoint_buffers(self, checkpoint_every):
        """(k, checkpoints, segment scratch, path buffers, partials) of the
        frozen adjoint at checkpoint interval k, allocated on first use."""
        w = self.w
        P, N, F = w.paths, w.n_rates, w.factors
        if not hasattr(self, "adj") or self.adj[0] != checkpoint_every:
            ck, sc, pb, pa = koo.bermudan_adjoint_workspace(P, N, F, checkpoint_every)
            dev = self.z.device
            self.adj = (checkpoint_every,
                        torch.empty(ck // 4, dtype=torch.float32, device=dev),
                        torch.empty(sc // 4, dtype=torch.float32, device=dev),
                        torch.empty(pb // 4, dtype=torch.float32, device=dev),
                        torch.empty(pa // 8, dtype=torch.float64, device=dev))
            self.vols = _dev(w.vols, torch.float32)
            self.frozen_price = torch.zeros(1, dtype=torch.float64, device=dev)
            self.frozen_greeks = torch.zeros(2 * N, dtype=torch.float64, device=dev)
        return self.adj

    def gen_product(self):
        """This product's koo.GenProduct (the fused generator's inputs, as
        generate())."""
        w = self.w
        return koo.GenProduct(w.path_offset, w.dims, self.dirs.data_ptr(), self.scr.data_ptr(),
                              self.alloc.data_ptr())

    def ls_batch_product(self, ls_blocks=0):
        """This product's koo.LsProduct (path quantities and LS first pass,
        as paths_kernel() + regress()); ls_blocks as koo_ls_product."""
        w = self.w
        self._ex_ls = np.ascontiguousarray(w.exercise_steps, np.int32)
        self._sh_ls = np.ascontiguousarray(self.shift, np.float32)
        self._sc_ls = np.ascontiguousarray(self.scale, np.float32)
        return koo.LsProduct(
            w.paths, w.n_rates, w.factors, self.z.data_ptr(), self.f0.data_ptr(),
            self.taus.data_ptr(), self.disp.data_ptr(), self.A.data_ptr(), float(w.strike),
            len(self._ex_ls), self._ex_ls.ctypes.data, self._sh_ls.ctypes.data,
            self._sc_ls.ctypes.data, 1, self.bv.data_ptr(), self.ev.data_ptr(),
            self.num.data_ptr(), self.ws.data_ptr(), self.coeffs.data_ptr(),
            self.price1.data_ptr(), int(ls_blocks))

    def batch_product(self, checkpoint_every):
        """This product's koo.BermProduct for the batched portfolio kernels
        (buffers as in greeks(); host arrays kept alive on the pricer)."""
        w = self.w
        k, ck, sc, pb, pa = self.adjoint_buffers(checkpoint_every)
        self._ex = np.ascontiguousarray(w.exercise_steps, np.int32)
        self._sh = np.ascontiguousarray(self.shift, np.float64)
        self._sc = np.ascontiguousarray(self.scale, np.float64)
        return koo.BermProduct(
            w.paths, w.n_rates, w.factors, k, self.z.data_ptr(), self.f0.data_ptr(),
            self.taus.data_ptr(), self.disp.data_ptr(), self.A.data_ptr(), self.vols.data_ptr(),
            float(w.strike), len(self._ex), self._ex.ctypes.data, self.coeffs.data_ptr(),
            self._sh.ctypes.data, self._sc.ctypes.data, 1, ck.data_ptr(), pb.data_ptr(),
            sc.data_ptr(), pa.data_ptr(), self.frozen_price.data_ptr(),
            self.frozen_greeks.data_ptr())

    def greeks(self, checkpoint_every=4):
        """Frozen-strategy price and pathwise adjoint Greeks (N deltas, N vegas)
        on the current variates, with the coefficients of the last regress()."""
        w = self.w
        P, N, F = w.paths, w.n_rates, w.factors
        k, ck, sc, pb, pa = self.adjoint_buffers(checkpoint_every)
        koo.bermudan_frozen_fwd(P, N, F, k, self.z, self.f0, self.taus, self.disp, self.A,
                                w.strike, w.exercise_steps, self.coeffs, self.shift, self.scale, True,
                                ck, pb, pa, self.frozen_price)
        koo.bermudan_frozen_bwd(P, N, F, k, self.z, self.f0, self.taus, self.disp, self.A,
                                self.vols, w.strike, ck, pb, sc, pa, self.frozen_greeks)
        return self.frozen_price, self.frozen_greeks

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


class PortfolioPricer:
    """A portfolio of Bermudan swaptions priced with Longstaff-Schwartz and
    frozen-strategy adjoint Greeks (generator, path quantities, first pass,
    frozen forward and reverse sweep per product).  ``persistent`` (the
    default): the whole portfolio in one persistent launch
    (bermudan_portfolio_persistent_run); else ``batched``: the generators on
    ``n_streams`` streams, then batched launches per stage; else each
    product's sequence on its stream.  With ``capture`` the launches are
    recorded once into a HIP graph that ``run`` replays.
    ``checkpoint_every`` maps a product index to its checkpoint interval
    (the autotuner's choice) or is one interval for all."""

    def __init__(self, products, checkpoint_every=4, n_streams=4, device="cuda", batched=True,
                 persistent=True):
        self.products = products
        self.pricers = [BermudanPricer(w, device=device) for w in products]
        self.k = [checkpoint_every(j) if callable(checkpoint_every) else checkpoint_every
                  for j in range(len(products))]
        self.streams = [torch.cuda.Stream(device=device) for _ in range(n_streams)]
        self.stream_of = [j % n_streams for j in range(len(products))]
        self.graph = None
        self.device = device
        # batched: the generators on the streams, then the path quantities and
        # LS first passes of every product (one launch for the paths, one per
        # date rank), every frozen forward in one launch and every backward in
        # one launch; otherwise each product's sequence on its stream
        self.batched = batched and all(w.factors <= 5 and w.n_rates <= 45 for w in products)
        self.ls_blocks = [0] * len(products)  # LS tile per product (0: the library's share)
        self.plan = None
        self.persistent = persistent and self.batched
        self.pplan = None
        self.work = None  # team work weights of the persistent launch (None: the model)

    def apply_tuning(self, tuning):
        """Checkpoint interval, stream and LS tile per product, and the
        persistent launch's team weights and launch form, from
        autotune.tune_portfolio."""
        self.k = list(tuning["k"])
        self.stream_of = list(tuning["stream"])
        self.ls_blocks = list(tuning.get("ls_blocks", [0] * len(self.pricers)))
        if "persistent" in tuning:
            self.persistent = bool(tuning["persistent"]) and self.batched
            self.work = tuning.get("work")
        self.graph = None
        self.plan = None
        self.pplan = None

    def _prepare_persistent(self):
        gen = [pr.gen_product() for pr in self.pricers]
        ls = [pr.ls_batch_product() for pr in self.pricers]
        adj = [pr.batch_product(self.k[j]) for j, pr in enumerate(self.pricers)]
        nbytes = koo.bermudan_portfolio_persistent_bytes([w.paths for w in self.products])
        self.pbuf = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        self.pplan = koo.bermudan_portfolio_persistent_prepare(gen, ls, adj, self.pbuf,
                                                               self.work)

    def _prepare_batch(self):
        n = len(self.pricers)
        ls_items = [pr.ls_batch_product(self.ls_blocks[j]) for j, pr in enumerate(self.pricers)]
        self.ls_descriptors = torch.empty(koo.bermudan_ls_portfolio_descriptor_bytes(n),
                                          dtype=torch.uint8, device=self.device)
        self.ls_plan = koo.bermudan_ls_portfolio_prepare(ls_items, self.ls_descriptors)
        items = [pr.batch_product(self.k[j]) for j, pr in enumerate(self.pricers)]
        self.descriptors = torch.empty(koo.bermudan_portfolio_descriptor_bytes(n),
                                       dtype=torch.uint8, device=self.device)
        self.plan = koo.bermudan_portfolio_prepare(items, self.descriptors)

    def _issue(self):
        if self.persistent:
            if self.pplan is None:
                self._prepare_persistent()
            koo.bermudan_portfolio_persistent_run(self.pplan, self.pbuf)
            return
        if self.batched and self.plan is None:
            self._prepare_batch()
        main = torch.cuda.current_stream()
        for s in self.streams:
            s.wait_stream(main)
        order = sorted(range(len(self.pricers)), key=lambda j: self.stream_of[j])
        for j in order:
            pr = self.pricers[j]
            with torch.cuda.stream(self.streams[self.stream_of[j]]):
                pr.generate()
                if not self.batched:
                    pr.paths_kernel()
                    pr.regress()
                    pr.greeks(checkpoint_every=self.k[j])
        for s in self.streams:
            main.wait_stream(s)
        if self.batched:
            koo.bermudan_ls_portfolio_run(self.ls_plan, self.ls_descriptors)
            koo.bermudan_portfolio_fwd(self.plan, self.descriptors)
            koo.bermudan_portfolio_bwd(self.plan, self.descriptors)

    def run(self, capture=True):
        """Prices the portfolio; with ``capture`` the first call records the
        graph (after one eager warm-up that allocates every buffer) and every
        call replays it."""
        if not capture:
            self._issue()
            return
        if self.graph is None:
            self._issue()  # warm-up: allocations, tables, lazy module loads
            torch.cuda.synchronize()
            self.graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.graph):
                self._issue()
        self.graph.replay()

    def measured_work(self):
        """Per-product block-time of the last persistent run (the sum of its
        team's block run times): the weights that size the teams to the
        work as it ran."""
        t = koo.persistent_block_ticks(self.pplan, self.pbuf).astype(np.float64)
        return [float(t[f:f + b].sum()) for f, b in
                zip(self.pplan.team_first[:len(self.pricers)],
                    self.pplan.team_blocks[:len(self.pricers)])]

    def results(self):
        """[(first-pass price, frozen-strategy price, greeks (2N))] per product."""
        return [(pr.price1.item(), pr.frozen_price.item(), pr.frozen_greeks.cpu().numpy())
                for pr in self.pricers]


class SwapPricer:
    """Plain-LMM payer swap through the reference-named stage kernels, the
    reference's staged pipeline (LMMEvolutionRoutineRawSingleKernel and its
    product chain, LMM_evolver_full.cu:264-425, Market_model_pricer.cu): Sobol
    integers, scramble of the first S rows, inverse normal, bridge, all-steps
    PC evolver with discount ratios, spot numeraires, extraction of each
    step's fixing, swap flows, discounting of both flow legs, FP64 sum."""

    def __init__(self, w, device="cuda"):
        self.w = w
        P, N, F, S = w.paths, w.n_rates, w.factors, w.steps
        D = S * F
        f32 = lambda n: torch.empty(n, dtype=torch.float32, device=device)  # noqa: E731
        self.dirs = _dev(koo.sobol_directions(D).view(np.int32), torch.int32)
        self.scr = _dev(np.asarray(w.scramblers, np.uint32).view(np.int32), torch.int32)
        self.alloc = _dev(koo.bridge_allocate_diagonal(F, S), torch.int32)
        self.ints = torch.empty(D * P, dtype=torch.int32, device=device)
        self.normals, self.z = f32(D * P), f32(D * P)
        self.rates, self.logs = f32(S * N * P), f32(S * N * P)
        self.disc, self.num, self.r1 = f32(S * (N + 1) * P), f32(S * P), f32(S * P)
        self.f1, self.f2, self.d1, self.d2 = f32(S * P), f32(S * P), f32(S * P), f32(S * P)
        self.s1, self.s2 = f32(P), f32(P)
        self.f0 = _dev(w.f0, torch.float32)
        self.lf0 = _dev(np.log(w.f0 + w.disp()), torch.float32)
        self.taus = _dev(w.taus, torch.float32)
        self.A = _dev(w.A, torch.float32)
        self.mu0 = _dev(koo.spot_drifts(w.A[0], w.f0, w.taus, w.disp(), 0), torch.float32)
        self.fx = _dev(koo.fixed_drifts(w.A), torch.float32)
        self.dsp = _dev(w.disp(), torch.float32)
        self.alive = _dev(w.alive(), torch.int32)
        self.first = _dev(np.arange(1, S + 1), torch.int32)
        self.theta = _dev(np.zeros(S), torch.float32)
        self.aux = _dev(np.concatenate([[w.strike, 1.0], w.taus, w.taus]), torch.float32)
        self.ws = torch.empty(koo.sum_workspace_doubles(P), dtype=torch.float64, device=device)
        self.price_out = torch.zeros(1, dtype=torch.float64, device=device)

    def price(self):
        w = self.w
        P, N, F, S = w.paths, w.n_rates, w.factors, w.steps
        koo.sobolGPUInts(P, S * F, w.path_offset, self.dirs, self.ints)
        koo.scramble_gpu(self.ints, self.scr, S, P)
        koo.inverse_normal_u32_gpu(self.ints, self.normals)
        koo.brownian_bridge_gpu(S, self.alloc, self.normals, self.z, P, F)
        koo.LMM_evolver_pc_all_gpu(self.f0, self.lf0, self.taus, self.z, self.A, self.mu0, self.fx,
                                   self.dsp, self.alive, P, F, S, N, self.rates, self.logs,
                                   self.disc)
        koo.spot_measure_numeraires_computation_gpu(self.disc, self.num, P, N)
        koo.forward_rate_extraction_gpu(self.rates, np.arange(S), P, S, N, self.r1)
        koo.cash_flow_generator("swap", self.f1, self.f2, self.aux, P, S, self.r1, self.r1,
                                self.r1)
        koo.cashFlowDiscounting_gpu(self.first, self.first, self.theta, self.disc, self.f1,
                                    self.num, P, S, self.d1, self.s1)
        koo.cashFlowDiscounting_gpu(self.first, self.first, self.theta, self.disc, self.f2,
                                    self.num, P, S, self.d2, self.s2)
        self.s1.add_(self.s2)
        koo.sum_f32_to_f64(self.s1, self.ws, self.price_out)
        self.price_out.div_(P)
        return self.price_out


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
