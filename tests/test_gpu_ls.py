"""Stage (d), Longstaff-Schwartz, on the GPU against the FP64 oracle.

* The MFMA Gram kernel (X^T X and X^T y) and the outer products against FP64
  numpy products of the same FP32 inputs (relative 1e-5: FP32 accumulation
  over at most a few hundred paths per wavefront, FP64 across them).
* FP64 Cholesky against numpy.linalg.solve (1e-12).
* Bermudan path quantities against the oracle evolving the same variates in
  FP64 (swap rates and forwards 2e-5 relative, exercise values 1e-5 of
  their maximum, numeraires 2e-6 relative).
* The LS first pass against the oracle on the same FP32 inputs: the first
  non-trivial regression (no earlier rounding involved) to 1e-4 in
  continuation value, the price to 1e-5 absolute; the frozen-strategy pass under the same
  coefficients to 1e-6.
* The config-3 workload end to end: GPU price within 1e-5 of the oracle that
  generates its own FP64 normals from the same Sobol integers.
* The reference-named small kernels against their stated formulas; the
  pointwise products bit-for-bit against FP32 numpy products."""
import numpy as np
import pytest

import oracle as O
from useful_code_amd import configs, koo, pipeline

pytestmark = pytest.mark.gpu


def dev(torch, x, dtype=None):
    dtype = dtype or torch.float32
    return torch.as_tensor(np.ascontiguousarray(x)).to(dtype).cuda()


# ---------------------------------------------------------------- MFMA Gram

@pytest.mark.parametrize("n_basis,P,masked", [(6, 10000, True), (6, 64, False), (12, 5000, True),
                                              (15, 4099, False), (1, 300, True)])
def test_normal_equations_gram_matches_numpy(torch_cuda, n_basis, P, masked):
    torch = torch_cuda
    rng = np.random.RandomState(n_basis * 1000 + P)
    X = rng.normal(size=(n_basis, P)).astype(np.float32)
    X[0] = 1.0
    y = rng.normal(size=P).astype(np.float32)
    mask = (rng.uniform(size=P) < 0.6).astype(np.float32) if masked else None
    ws = torch.empty(koo.ls_workspace(P), dtype=torch.uint8, device="cuda")
    out = torch.zeros((n_basis + 1) ** 2, dtype=torch.float64, device="cuda")
    print(f"gram n_basis={n_basis} P={P} masked={masked}", flush=True)
    koo.ls_normal_eq_gpu(P, n_basis, dev(torch, X), dev(torch, y),
                         None if mask is None else dev(torch, mask), ws, out)
    torch.cuda.synchronize()
    Z = np.vstack([X, y[None]]).astype(np.float64)
    if mask is not None:
        Z = Z * mask[None]
    want = Z @ Z.T
    got = out.cpu().numpy().reshape(n_basis + 1, n_basis + 1)
    scale = np.sqrt(np.outer(np.diag(want), np.diag(want)))
    assert np.max(np.abs(got - want) / scale) < 1e-5


@pytest.mark.parametrize("R,C", [(6, 6), (6, 3), (11, 10), (1, 17)])
def test_outer_products_match_numpy(torch_cuda, R, C):
    torch = torch_cuda
    P = 3000
    rng = np.random.RandomState(R * 100 + C)
    a = rng.normal(size=(R, P)).astype(np.float32)
    b = rng.normal(size=(C, P)).astype(np.float32)
    wsf = koo.outer_product_workspace_floats(P)
    ws = torch.empty(wsf, dtype=torch.float32, device="cuda")
    ans = torch.zeros(R * C, dtype=torch.float32, device="cuda")
    print(f"outer R={R} C={C}", flush=True)
    koo.reduceOuterProduct_gpu(P, R, C, False, 256, wsf // (R * C), dev(torch, a), dev(torch, b),
                               ans, ws)
    want = a.astype(np.float64) @ b.astype(np.float64).T
    got = ans.cpu().numpy().reshape(R, C)
    assert np.max(np.abs(got - want)) < 1e-5 * np.sqrt(P) * 4
    sym = torch.zeros(R * R, dtype=torch.float32, device="cuda")
    koo.reduceOuterProduct_gpu(P, R, R, True, 256, wsf // (R * R), dev(torch, a), None, sym, ws)
    want = a.astype(np.float64) @ a.astype(np.float64).T
    assert np.allclose(sym.cpu().numpy().reshape(R, R), want, rtol=1e-5, atol=1e-5 * P)
    # cuBLAS-named variant: column-major, alpha * A A^T + beta * C
    old = rng.normal(size=(R, R)).astype(np.float32)
    c = dev(torch, old.T.copy().reshape(-1))  # column-major storage of `old`
    koo.reduceOuterProductSymmetric_cublas(P, R, 0.5, 2.0, dev(torch, a), c, ws)
    got = c.cpu().numpy().reshape(R, R).T
    assert np.allclose(got, 0.5 * want + 2.0 * old, rtol=1e-5, atol=1e-5 * P)


def test_reference_sized_workspace_is_enough(torch_cuda):
    torch = torch_cuda
    P, R = 20000, 6
    a = np.random.RandomState(1).normal(size=(R, P)).astype(np.float32)
    blocks = 64
    ws = torch.empty(blocks * R * R, dtype=torch.float32, device="cuda")  # reference contract
    ans = torch.zeros(R * R, dtype=torch.float32, device="cuda")
    koo.reduceOuterProduct_gpu(P, R, R, True, blocks, blocks, dev(torch, a), None, ans, ws)
    want = a.astype(np.float64) @ a.astype(np.float64).T
    assert np.allclose(ans.cpu().numpy().reshape(R, R), want, rtol=1e-5, atol=1e-5 * P)


@pytest.mark.parametrize("P,off", [(4096, 0), (4099, 0), (4096, 1), (1, 0)])
def test_pointwise_products_are_exact(torch_cuda, P, off):
    """FP32 products are exact roundings of the FP64 ones: bit-for-bit with
    numpy's FP32 multiply, on the float4 path (P % 4 == 0, aligned) and the
    scalar path (ragged P or a row starting one float in)."""
    torch = torch_cuda
    R = 5
    rng = np.random.RandomState(P + off)
    a = rng.normal(size=(R, P)).astype(np.float32)
    b = rng.normal(size=P).astype(np.float32)
    da = torch.zeros(R * P + off, dtype=torch.float32, device="cuda")
    da[off:] = dev(torch, a.reshape(-1))
    db = torch.zeros(P + off, dtype=torch.float32, device="cuda")
    db[off:] = dev(torch, b)
    out = torch.zeros(P + off, dtype=torch.float32, device="cuda")
    koo.PointwiseProduct_gpu(64, 256, P, da[off:], db[off:], out[off:])
    assert np.array_equal(out[off:].cpu().numpy(), a[0] * b)
    outm = torch.zeros(R * P + off, dtype=torch.float32, device="cuda")
    koo.PointwiseProductMultipleFirst_gpu(64, 256, P, da[off:], db[off:], outm[off:], R)
    assert np.array_equal(outm[off:].cpu().numpy().reshape(R, P), a * b[None, :])
    sentinel = -7.0
    outs = torch.full((R * R * P + off,), sentinel, dtype=torch.float32, device="cuda")
    koo.PointwiseProductSymmetricUpperTriangular_gpu(64, 256, P, R, da[off:], outs[off:])
    got = outs[off:].cpu().numpy().reshape(R, R, P)
    for r in range(R):
        for c in range(R):
            if c >= r:
                assert np.array_equal(got[r, c], a[r] * a[c])
            else:
                assert np.all(got[r, c] == sentinel)


def test_cholesky_solve_gpu(torch_cuda):
    torch = torch_cuda
    rng = np.random.RandomState(4)
    for n in (1, 6, 15):
        G = rng.normal(size=(n + 4, n))
        A = G.T @ G + 0.5 * np.eye(n)
        b = rng.normal(size=n)
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        info = torch.full((1,), -1, dtype=torch.int32, device="cuda")
        koo.cholesky_solve_f64_gpu(n, dev(torch, A, torch.float64), dev(torch, b, torch.float64),
                                   x, info)
        assert info.item() == 0
        assert np.allclose(x.cpu().numpy(), np.linalg.solve(A, b), rtol=1e-12, atol=1e-13)
    A = -np.eye(3)
    x = torch.zeros(3, dtype=torch.float64, device="cuda")
    info = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    koo.cholesky_solve_f64_gpu(3, dev(torch, A, torch.float64),
                               dev(torch, np.ones(3), torch.float64), x, info)
    assert info.item() == 1 and np.all(x.cpu().numpy() == 0.0)


# ---------------------------------------------------------------- Bermudan LS

def _oracle_model(w):
    return O.caplet_model(w.f0, w.taus, w.disp(), w.A, w.vols, w.strike, w.paths)


# Frozen-strategy Greeks are held to 1e-4 relative per Greek, plus 1e-6 of the
# largest Greek: a Greek whose path contributions cancel to far below the
# largest (a first delta, a far vega) is resolved only to the FP32 precision
# of those contributions.  That is 1e-7 to 1e-6 of the largest: FP32 log-rates
# near log(0.05) carry 2.4e-7 of rounding per step, accumulated over up to 40
# steps and products of up to 40 discount factors (measured on the 64-product
# portfolio: up to 5e-7 of the largest, on the first deltas of the one-factor
# 34-rate swaption, whose price the FP32 paths move by 1.8e-6 relative on the
# same exercise decisions).  They are compared with the
# FP64 oracle on the GPU's own exercise decisions, so a decision that an FP32
# continuation value takes the other way near the boundary does not enter
# the comparison; the decisions themselves are checked against the oracle's
# own (FP64) strategy separately.
GREEK_REL = 1e-4
GREEK_FLOOR = 1e-6
FLIPS_PER_PATH = 1e-4


def _gpu_exercise_steps(torch, pr, w):
    N, P = w.n_rates, w.paths
    return pr.adj[3].view(torch.int32)[N * P:(N + 1) * P].cpu().numpy()


def _check_frozen(torch, w, pr, z, price, greeks, tag):
    """Frozen price and every Greek of one product against the FP64 oracle
    on the same Sobol points and coefficients; returns (flips, max rel err)."""
    N, P = w.n_rates, w.paths
    coeffs = pr.coeffs.cpu().numpy().reshape(pr.E, 6)
    fs = O.frozen_strategy(w.exercise_steps, coeffs, w.basis_shift, w.basis_scale)
    m = _oracle_model(w)
    vals, ex_o, _, _ = O.bermudan_frozen_paths(m, fs, z, adjoint=False)
    tau_gpu = _gpu_exercise_steps(torch, pr, w)
    flips = int(np.sum(tau_gpu != ex_o))
    gv, gd, gvv = O.bermudan_given_paths(m, fs, z, tau_gpu)
    g = greeks.cpu().numpy() if hasattr(greeks, "cpu") else np.asarray(greeks)
    g64 = np.concatenate([gd, gvv]) / P
    # error over its allowance: 1e-4 of the Greek plus 1e-6 of the largest
    rel = np.abs(g - g64) / (GREEK_REL * np.abs(g64) + GREEK_FLOOR * np.max(np.abs(g64)))
    print(f"{tag}: frozen price gpu={price:.10f} fp64(same decisions)={gv.mean():.10f} "
          f"fp64(own decisions)={vals.mean():.10f}; decisions differing {flips}/{P}; "
          f"greeks: error / allowance max {rel.max():.2e}, max rel err "
          f"{np.max(np.abs(g - g64) / np.abs(g64)):.2e}", flush=True)
    for i in np.argsort(-rel)[:3]:
        print(f"  greek {i}: gpu {g[i]:.10e} fp64 {g64[i]:.10e} "
              f"rel {abs(g[i] / g64[i] - 1):.2e} of allowance {rel[i]:.2e}", flush=True)
    assert flips <= FLIPS_PER_PATH * P, tag
    assert abs(price - gv.mean()) < 1e-6, tag
    assert abs(price - vals.mean()) < 1e-5, tag
    assert np.all(rel <= 1.0), tag
    return flips, float(rel.max())


@pytest.fixture(scope="module")
def berm(torch_cuda):
    w = configs.bermudan_swaption(paths=1 << 15)
    pr = pipeline.BermudanPricer(w)
    pr.generate()
    pr.paths_kernel()
    torch_cuda.cuda.synchronize()
    E, P = pr.E, w.paths
    z = pr.z.cpu().numpy().astype(np.float64).reshape(w.steps, w.factors, P)
    bv = pr.bv.cpu().numpy().reshape(E, 2, P)
    ev = pr.ev.cpu().numpy().reshape(E, P)
    num = pr.num.cpu().numpy().reshape(E, P)
    return w, pr, z, bv, ev, num


def test_bermudan_paths_match_fp64_oracle(berm):
    w, pr, z, bv, ev, num = berm
    obv, oev, onum = O.bermudan_paths(_oracle_model(w), z, w.exercise_steps)
    err_sr = np.max(np.abs(bv[:, 0] / obv[:, 0] - 1))
    err_f = np.max(np.abs(bv[:, 1] / obv[:, 1] - 1))
    err_ev = np.max(np.abs(ev - oev)) / np.max(oev)
    err_num = np.max(np.abs(num / onum - 1))
    print(f"swap rate {err_sr:.2e} forward {err_f:.2e} exercise {err_ev:.2e} numeraire {err_num:.2e}")
    assert err_sr < 2e-5 and err_f < 2e-5 and err_ev < 1e-5 and err_num < 2e-6
    assert np.array_equal(ev > 0, oev > 0) or np.mean((ev > 0) != (oev > 0)) < 1e-4


def _norm(w, bv):
    """The LS variables as the GPU regresses on them, (x - shift) * scale per
    date: well scaled, so the 1e-8 SVD cut of the solve drops nothing."""
    bv = np.asarray(bv, np.float64)
    return (bv - w.basis_shift[:, :, None]) * w.basis_scale[:, :, None]


def _gpu_first_pass(torch, pr):
    pr.regress()
    torch.cuda.synchronize()
    return pr.price1.item(), pr.coeffs.cpu().numpy().reshape(pr.E, 6)


def _to_raw(c, shift, scale):
    """Coefficients of the quadratic cross basis in x~ = (x - s) * a, b ->
    the same function in raw x, y (exact algebra, FP64)."""
    (sx, sy), (ax, ay) = shift, scale
    b0, b1, b2, b3, b4, b5 = c
    # x~ = ax x - ax sx,  y~ = ay y - ay sy
    px, qx = ax, -ax * sx
    py, qy = ay, -ay * sy
    r = np.zeros(6)
    r[0] = b0 + b1 * qx + b2 * qy + b3 * qx * qx + b4 * qx * qy + b5 * qy * qy
    r[1] = b1 * px + 2 * b3 * px * qx + b4 * px * qy
    r[2] = b2 * py + b4 * qx * py + 2 * b5 * py * qy
    r[3] = b3 * px * px
    r[4] = b4 * px * py
    r[5] = b5 * py * py
    return r


def test_ls_first_pass_matches_fp64_oracle_on_same_inputs(torch_cuda, berm):
    w, pr, z, bv, ev, num = berm
    E, P = ev.shape
    price, coeffs = _gpu_first_pass(torch_cuda, pr)
    ves = np.full(E, 2, np.int32)
    bvn = _norm(w, bv)
    oprice, ocoeffs = O.ls_first_pass(bvn, ves, np.zeros((E, P)), ev.astype(np.float64),
                                      num.astype(np.float64), itm_only=True)
    # Date E-2: the first regression with a non-zero target (the last date's
    # target is identically zero and its decision is exercise-if-in-the-money,
    # so no earlier rounding enters); compare continuation values.
    e = E - 2
    itm = ev[e] > 0
    X = np.stack([O.basis_functions(bvn[e, :, p]) for p in np.nonzero(itm)[0][:2000]])
    c_gpu, c_orc = X @ coeffs[e], X @ ocoeffs[e]
    rel = np.max(np.abs(c_gpu - c_orc)) / np.max(np.abs(c_orc))
    print(f"first pass price gpu={price:.10f} fp64={oprice:.10f} diff={price - oprice:.3e}; "
          f"date E-2 continuation rel err {rel:.2e}")
    assert rel < 1e-4
    assert abs(price - oprice) < 1e-5


@pytest.mark.parametrize("paths", [1 << 18, 5000, 77])
def test_persistent_first_pass_equals_per_date_launches(torch_cuda, paths):
    """The first pass as one persistent launch (KOO_LS_LAUNCH_PER_DATE=0:
    the partials exchanged through sentinel-armed slots between dates) and as
    one launch per date (the default) run the same per-date block code over
    the same blocks, so coefficients and price agree bit for bit; repeated
    persistent runs (the slots re-armed each time) agree too."""
    import os
    torch = torch_cuda
    w = configs.bermudan_swaption(paths=paths)
    pr = pipeline.BermudanPricer(w)
    pr.generate()
    pr.paths_kernel()
    print(f"persistent first pass, {paths} paths", flush=True)
    runs = []
    for mode in ("0", "1", "0"):
        os.environ["KOO_LS_LAUNCH_PER_DATE"] = mode
        try:
            pr.coeffs.fill_(float("nan"))
            pr.price1.fill_(float("nan"))
            pr.regress()
            torch.cuda.synchronize()
        finally:
            os.environ.pop("KOO_LS_LAUNCH_PER_DATE", None)
        runs.append((pr.coeffs.clone(), pr.price1.item()))
    (c0, p0), (c1, p1), (c2, p2) = runs
    print(f"price persistent {p0:.12f} per date {p1:.12f}", flush=True)
    assert np.isfinite(p0) and p0 == p1 == p2
    assert torch.equal(c0, c1) and torch.equal(c0, c2)


def test_ls_second_pass_matches_oracle_under_same_coefficients(torch_cuda, berm):
    w, pr, z, bv, ev, num = berm
    E, P = ev.shape
    _, coeffs = _gpu_first_pass(torch_cuda, pr)
    pr.second_pass()
    torch_cuda.cuda.synchronize()
    got = pr.price2.item()
    raw = np.stack([_to_raw(coeffs[e], w.basis_shift[e], w.basis_scale[e]) for e in range(E)])
    want = O.ls_second_pass(bv.astype(np.float64), np.full(E, 2, np.int32), np.zeros((E, P)),
                            ev.astype(np.float64), num.astype(np.float64), raw, itm_only=True)
    print(f"second pass gpu={got:.10f} fp64={want:.10f} diff={got - want:.3e}")
    assert abs(got - want) < 1e-6


def test_ls_all_paths_and_plain_basis(torch_cuda, berm):
    # the regression over every path and the basis without cross terms
    w, pr, z, bv, ev, num = berm
    torch = torch_cuda
    E, P = ev.shape
    for cross in (True, False):
        k = O.basis_count(2, cross)
        coeffs = torch.zeros(E * k, dtype=torch.float64, device="cuda")
        out = torch.zeros(1, dtype=torch.float64, device="cuda")
        koo.ls_first_pass_gpu(E, pr.bv, pr.ves, 2, None, pr.ev, pr.num, P, cross, False, pr.shift,
                              pr.scale, pr.ws, coeffs, out)
        torch.cuda.synchronize()
        oprice, _ = O.ls_first_pass(_norm(w, bv), np.full(E, 2, np.int32),
                                    np.zeros((E, P)), ev.astype(np.float64),
                                    num.astype(np.float64), use_cross=cross, itm_only=False)
        print(f"all paths cross={cross}: gpu={out.item():.10f} fp64={oprice:.10f}")
        assert abs(out.item() - oprice) < 1e-5


def test_ls_with_cash_flows_and_fewer_variables(torch_cuda, berm):
    # cancellable-swap style flows and a one-variable last date (three basis
    # functions), all paths
    w, pr, z, bv, ev, num = berm
    torch = torch_cuda
    E, P = ev.shape
    cash = (1e-3 * np.cos(np.arange(E * P) * 0.37)).astype(np.float32).reshape(E, P)
    ves = np.full(E, 2, np.int32)
    ves[-1] = 1
    coeffs = torch.zeros(E * 6, dtype=torch.float64, device="cuda")
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    koo.ls_first_pass_gpu(E, pr.bv, ves, 2, dev(torch, cash), pr.ev, pr.num, P, True, False,
                          pr.shift, pr.scale, pr.ws, coeffs, out)
    torch.cuda.synchronize()
    oprice, _ = O.ls_first_pass(_norm(w, bv), ves, cash.astype(np.float64),
                                ev.astype(np.float64), num.astype(np.float64))
    print(f"with flows: gpu={out.item():.10f} fp64={oprice:.10f}")
    assert abs(out.item() - oprice) < 1e-5


def test_bermudan_config3_price_vs_full_fp64_oracle(torch_cuda):
    # 2^18 paths, 10 exercise dates: GPU pipeline vs the oracle generating its
    # own FP64 normals from the same scrambled Sobol integers.
    w = configs.bermudan_swaption()
    pr = pipeline.BermudanPricer(w)
    p1, p2 = pr.price()
    torch_cuda.cuda.synchronize()
    ints = O.scramble(O.sobol_ints(w.paths, w.dims, 0), w.scramblers, w.dims)
    z = O.bridge_transform(O.invnorm_u32(ints), w.steps, w.factors)
    bv, ev, num = O.bermudan_paths(_oracle_model(w), z, w.exercise_steps)
    E = len(w.exercise_steps)
    oprice, _ = O.ls_first_pass(_norm(w, bv), np.full(E, 2, np.int32), np.zeros_like(ev), ev,
                                num, itm_only=True)
    print(f"config 3: first pass gpu={p1.item():.10f} fp64={oprice:.10f} "
          f"diff={p1.item() - oprice:.3e}; second pass gpu={p2.item():.10f}")
    assert abs(p1.item() - oprice) < 1e-5
    assert abs(p2.item() / p1.item() - 1) < 0.01


# ---------------------------------------------------------------- reference kernels

def test_basis_variables_and_exercise_values(torch_cuda):
    torch = torch_cuda
    rng = np.random.RandomState(8)
    E, batch, total, off = 3, 1000, 1500, 400
    r1, r2, r3 = (rng.uniform(0.01, 0.09, size=(E, batch)).astype(np.float32) for _ in range(3))
    out = torch.full((E * 3 * total,), -7.0, dtype=torch.float32, device="cuda")
    koo.adjoinBasisVariablesCaller_gpu(False, None, dev(torch, r1), dev(torch, r2), dev(torch, r3),
                                       batch, total, off, E, out)
    got = out.cpu().numpy().reshape(E, 3, total)
    assert np.array_equal(got[:, 0, off:off + batch], r1)
    assert np.array_equal(got[:, 2, off:off + batch], r3)
    assert np.all(got[:, :, :off] == -7.0) and np.all(got[:, :, off + batch:] == -7.0)
    shifts = np.array([0.01, 0.02, 0.03], np.float32)
    koo.adjoinBasisVariablesCaller_gpu(True, dev(torch, shifts), dev(torch, r1), dev(torch, r2),
                                       dev(torch, r3), batch, total, off, E, out)
    got = out.cpu().numpy().reshape(E, 3, total)
    assert np.allclose(got[:, 1, off:off + batch], np.log(r2.astype(np.float64) + 0.02), atol=2e-6)
    ev = torch.full((E * total,), -7.0, dtype=torch.float32, device="cuda")
    koo.adjoinExerciseValues_Bermudan_swaption_gpu(ev, dev(torch, np.array([0.05], np.float32)), E,
                                                   total, off, batch, dev(torch, r1), dev(torch, r2))
    got = ev.cpu().numpy().reshape(E, total)
    want = np.maximum((r1 - np.float32(0.05)) * r2, 0)
    assert np.allclose(got[:, off:off + batch], want, rtol=1e-6, atol=1e-9)
    assert np.all(got[:, :off] == -7.0)


@pytest.mark.parametrize("v,cross", [(1, True), (2, True), (3, True), (4, True), (3, False)])
def test_expand_basis_functions(torch_cuda, v, cross):
    torch = torch_cuda
    P, maxv, e = 777, 4, 1
    rng = np.random.RandomState(v)
    bv = rng.normal(size=(3, maxv, P)).astype(np.float32)
    k = O.basis_count(v, cross)
    out = torch.zeros(k * P, dtype=torch.float32, device="cuda")
    koo.expandBasisFunctions_quadratic_gpu(P, dev(torch, bv), e, maxv, v, cross, out)
    got = out.cpu().numpy().reshape(k, P)
    want = np.stack([O.basis_functions(bv[e, :v, p].astype(np.float64), cross) for p in range(P)],
                    axis=1)
    assert np.allclose(got, want, rtol=1e-6, atol=1e-6)


def test_continuation_update_and_deflation(torch_cuda):
    torch = torch_cuda
    rng = np.random.RandomState(9)
    P, k, E, depth, e = 2000, 6, 4, 2, 2
    bf = rng.normal(size=(k, P)).astype(np.float32)
    w = rng.normal(scale=0.1, size=(E, depth, k)).astype(np.float32)
    ev = np.maximum(rng.normal(scale=0.3, size=(E, P)), 0).astype(np.float32)
    net = torch.zeros(P, dtype=torch.float32, device="cuda")
    koo.continuation_net_value_evaluation_gpu(P, e, dev(torch, bf), k, dev(torch, w[e, 0]),
                                              dev(torch, ev[e]), 0.01, net)
    want_net = 0.01 + w[e, 0].astype(np.float64) @ bf.astype(np.float64) - ev[e]
    assert np.allclose(net.cpu().numpy(), want_net, atol=1e-5)
    lo, hi = np.array([-0.05, 0], np.float32), np.array([0.05, 0], np.float32)
    cont0 = rng.normal(size=P).astype(np.float32)
    cont = dev(torch, cont0)
    est = torch.zeros(P, dtype=torch.float32, device="cuda")
    exe = torch.zeros(P, dtype=torch.int32, device="cuda")
    koo.oneStepUpdateLS_multi_gpu(P, e, dev(torch, bf), k, k, depth, dev(torch, w), dev(torch, lo),
                                  dev(torch, hi), cont, None, dev(torch, ev), 0.0, est, exe)
    # depth 0, then depth 1 inside the band (LS_gpu.cu:360-392)
    c0 = w[e, 0].astype(np.float64) @ bf.astype(np.float64)
    c1 = w[e, 1].astype(np.float64) @ bf.astype(np.float64)
    n0 = c0 - ev[e]
    inside = (n0 >= lo[0]) & (n0 <= hi[0])
    c = np.where(inside, c1, c0)
    exercise = (c - ev[e]) < 0
    safe = np.abs(c - ev[e]) > 1e-5  # away from FP32 ties
    safe &= np.abs(n0 - lo[0]) > 1e-5
    safe &= np.abs(n0 - hi[0]) > 1e-5
    got_ex = exe.cpu().numpy()
    assert np.array_equal(got_ex[safe], np.where(exercise, 1, -1)[safe])
    assert np.allclose(est.cpu().numpy()[safe], c[safe], atol=1e-5)
    got_cont = cont.cpu().numpy()
    assert np.array_equal(got_cont[safe], np.where(exercise, ev[e], cont0)[safe])
    num = rng.uniform(1.0, 1.5, size=(E, P)).astype(np.float32)
    y = dev(torch, cont0)
    koo.updateDeflationOfContinuationValues_gpu(y, dev(torch, num), P, 1, 3)
    assert np.allclose(y.cpu().numpy(), cont0 * num[1] / num[3], rtol=1e-6)


# ---------------------------------------------------------------- frozen-strategy adjoint

@pytest.mark.parametrize("k", [4, 1, 20])
def test_frozen_bermudan_price_and_greeks_vs_fp64_oracle(torch_cuda, berm, k):
    w, pr, z, bv, ev, num = berm
    torch = torch_cuda
    pr.regress()
    price, greeks = pr.greeks(checkpoint_every=k)
    torch.cuda.synchronize()
    _check_frozen(torch, w, pr, z, price.item(), greeks, f"k={k}")


def test_frozen_bermudan_with_displaced_rates_vs_fp64_oracle(torch_cuda):
    """Displaced rates (d_r = 0.003 + 0.0002 r) through the Bermudan paths,
    first pass and frozen-strategy adjoint: the shared evolution step forms
    the drift coefficient from 1/tau_r - d_r, the exercise value and
    numeraire from f = exp(l) - d."""
    torch = torch_cuda
    w = configs.bermudan_swaption(paths=1 << 14)
    w.displacements = 0.003 + 0.0002 * np.arange(w.n_rates)
    pr = pipeline.BermudanPricer(w)
    pr.generate()
    pr.paths_kernel()
    pr.regress()
    price, greeks = pr.greeks(checkpoint_every=4)
    torch.cuda.synchronize()
    z = pr.z.cpu().numpy().astype(np.float64).reshape(w.steps, w.factors, w.paths)
    _check_frozen(torch, w, pr, z, price.item(), greeks, "displaced")


# ---------------------------------------------------------------- batched portfolio

def test_batched_portfolio_equals_per_product_calls(torch_cuda):
    """The batched launches run each product's blocks through the same
    per-block code as the single-product calls (products of 1-5 factors,
    10-40 rates, checkpoint intervals 1, 3 and 8 mixed in one batch): given
    the same coefficients the frozen forward/backward agree bit for bit; the
    batched LS uses fewer, fuller blocks per product, so its FP64 sums
    associate differently and its coefficients and first-pass prices agree
    with the single-product pass to rounding."""
    torch = torch_cuda
    prods = configs.bermudan_portfolio(n_products=9, paths=(1 << 12) + 77, seed=78)
    ks = lambda j: (1, 3, 8)[j % 3]  # noqa: E731
    batch = pipeline.PortfolioPricer(prods, checkpoint_every=ks, n_streams=2, batched=True,
                                     persistent=False)
    assert batch.batched and not batch.persistent
    print("batched portfolio, 9 products", flush=True)
    batch.run(capture=False)
    torch.cuda.synchronize()
    got = batch.results()
    for j, pr in enumerate(batch.pricers):
        p1, p2, g = got[j]
        coeffs = pr.coeffs.cpu().numpy()
        # frozen forward/backward on the batch's own paths and coefficients
        fp, fg = pr.greeks(checkpoint_every=ks(j))
        torch.cuda.synchronize()
        assert fp.item() == p2 and np.array_equal(fg.cpu().numpy(), g), j
        # LS first pass of the single-product call on the same path quantities
        pr.regress()
        torch.cuda.synchronize()
        c1 = pr.coeffs.cpu().numpy()
        assert np.allclose(c1, coeffs, rtol=1e-9, atol=1e-12), j
        assert abs(pr.price1.item() - p1) <= 1e-12 * abs(p1), j


@pytest.mark.parametrize("seed,paths", [(78, (1 << 12) + 77), (81, 1 << 14)])
def test_persistent_portfolio_equals_stage_calls(torch_cuda, seed, paths):
    """The whole portfolio in one persistent launch against the stage calls
    run on its own outputs: the generated variates equal sobol_bridge_gpu's
    to 2 ulp; the path quantities equal bermudan_paths_gpu's on those
    variates bit for bit; the first pass agrees with ls_first_pass_gpu to rounding (its
    partials are summed over the product's team, not the single call's
    blocks); the frozen price and Greeks equal bermudan_frozen_fwd/_bwd on
    the same variates and coefficients bit for bit (checkpoint intervals 1,
    3 and 8 mixed)."""
    torch = torch_cuda
    prods = configs.bermudan_portfolio(n_products=9, paths=paths, seed=seed)
    ks = lambda j: (1, 3, 8)[j % 3]  # noqa: E731
    pf = pipeline.PortfolioPricer(prods, checkpoint_every=ks, n_streams=2)
    assert pf.persistent
    print(f"persistent portfolio, 9 products x {paths} paths", flush=True)
    for pr in pf.pricers:
        for t in (pr.z, pr.bv, pr.ev, pr.num):
            t.fill_(float("nan"))
    pf.run(capture=False)
    torch.cuda.synchronize()
    got = pf.results()
    print(f"  grid {pf.pplan.grid} blocks ({pf.pplan.blocks_per_cu} per CU), "
          f"LDS {pf.pplan.lds_bytes} B", flush=True)
    for j, pr in enumerate(pf.pricers):
        p1, p2, g = got[j]
        assert np.isfinite(p1) and np.isfinite(p2) and np.all(np.isfinite(g)), j
        z, bv, ev, num = (t.clone() for t in (pr.z, pr.bv, pr.ev, pr.num))
        coeffs = pr.coeffs.clone()
        pr.generate()
        torch.cuda.synchronize()
        # the same Sobol points, quantiles and bridge; the LDS bridge and the
        # register bridge may round a bridge point (|W| up to about 25)
        # differently by an ulp or two, which the increments inherit
        d = (z - pr.z).abs()
        assert d.max().item() <= 4e-6, (j, d.max().item())
        if j == 0:
            print(f"  variates vs sobol_bridge_gpu: {int((d > 0).sum())} of {d.numel()} differ, "
                  f"max {d.max().item():.2e}", flush=True)
        pr.z.copy_(z)  # the stage calls below on the persistent kernel's variates
        pr.paths_kernel()
        torch.cuda.synchronize()
        assert torch.equal(bv, pr.bv) and torch.equal(ev, pr.ev) and torch.equal(num, pr.num), j
        fp, fg = pr.greeks(checkpoint_every=ks(j))
        torch.cuda.synchronize()
        assert fp.item() == p2 and np.array_equal(fg.cpu().numpy(), g), j
        pr.regress()
        torch.cuda.synchronize()
        # the FP64 normal equations summed in another order; a one-factor
        # product's are close to singular (condition numbers up to 1e12), so
        # its coefficients move by up to 1e-6 relative while the fitted
        # continuation values and the price do not
        c1, c0 = pr.coeffs.cpu().numpy(), coeffs.cpu().numpy()
        worst = np.max(np.abs(c1 - c0) / (np.abs(c0) + 1e-12))
        print(f"  product {j}: first pass {p1:.12f} vs {pr.price1.item():.12f}, coefficients "
              f"max rel diff {worst:.1e}", flush=True)
        assert np.allclose(c1, c0, rtol=1e-6, atol=1e-12), j
        assert abs(pr.price1.item() - p1) <= 1e-12 * abs(p1), j


def test_portfolio_persistent_ls_equals_per_round_launches(torch_cuda):
    """The portfolio's LS first passes as one persistent launch (resident
    blocks walking all products' blocks, the partials exchanged through
    sentinel-armed slots each round; KOO_LS_LAUNCH_PER_DATE=0) and as one
    launch per round (the default): same per-block code and partials, so
    every product's coefficients and first-pass price agree bit for bit."""
    import os
    torch = torch_cuda
    prods = configs.bermudan_portfolio(n_products=12, paths=(1 << 13) + 5, seed=79)
    pf = pipeline.PortfolioPricer(prods, checkpoint_every=4, n_streams=2, persistent=False)
    assert pf.batched and not pf.persistent
    print("portfolio LS, 12 products", flush=True)
    runs = []
    for mode in ("0", "1"):
        os.environ["KOO_LS_LAUNCH_PER_DATE"] = mode
        try:
            for pr in pf.pricers:
                pr.coeffs.fill_(float("nan"))
                pr.price1.fill_(float("nan"))
            pf.run(capture=False)
            torch.cuda.synchronize()
        finally:
            os.environ.pop("KOO_LS_LAUNCH_PER_DATE", None)
        runs.append([(pr.coeffs.clone(), pr.price1.item()) for pr in pf.pricers])
    for j, ((c0, p0), (c1, p1)) in enumerate(zip(*runs)):
        assert np.isfinite(p0) and p0 == p1, (j, p0, p1)
        assert torch.equal(c0, c1), j


def test_bermudan_job_graph_replays_equal_eager(torch_cuda):
    """The config-3 job recorded as one HIP graph: every replay, with eager
    work and clobbered outputs between replays, equals the eager job bit for
    bit (coefficients, first- and second-pass prices, path quantities)."""
    torch = torch_cuda
    w = configs.bermudan_swaption(paths=1 << 15)
    pr = pipeline.BermudanPricer(w)
    pr.job(capture=False)
    torch.cuda.synchronize()
    want = [t.clone() for t in (pr.coeffs, pr.price1, pr.price2, pr.bv, pr.ev, pr.num)]
    other = pipeline.BermudanPricer(configs.bermudan_swaption(paths=1 << 12, seed=9))
    for rep in range(4):
        for t in (pr.coeffs, pr.price1, pr.price2):
            t.fill_(float("nan"))
        pr.bv.zero_()
        pr.ws.fill_(7)
        other.job(capture=False)  # eager work of the program between replays
        torch.cuda.synchronize()
        pr.job(capture=True)
        torch.cuda.synchronize()
        got = (pr.coeffs, pr.price1, pr.price2, pr.bv, pr.ev, pr.num)
        for a, b in zip(want, got):
            assert torch.equal(a, b), rep
    print(f"graph replays equal eager: price {pr.price1.item():.8f} / {pr.price2.item():.8f}")


def test_portfolio_graph_replays_equal_eager_and_oracle(torch_cuda):
    torch = torch_cuda
    from useful_code_amd import autotune
    prods = configs.bermudan_portfolio(n_products=6, paths=1 << 13, seed=77)
    pf = pipeline.PortfolioPricer(prods, checkpoint_every=4, n_streams=3)
    assert pf.batched
    pf.run(capture=False)
    torch.cuda.synchronize()
    eager = pf.results()
    for rep in range(3):
        pf.run(capture=True)
        torch.cuda.synchronize()
        got = pf.results()
        for (a1, a2, ag), (b1, b2, bg) in zip(eager, got):
            assert a1 == b1 and a2 == b2 and np.array_equal(ag, bg), rep
        # eager work between replays
        pf.pricers[0].generate()
        torch.cuda.synchronize()
    tuning = autotune.tune_portfolio(pf)
    assert len(tuning["k"]) == 6 and set(tuning["stream"]) <= {0, 1, 2}
    pf.apply_tuning(tuning)
    pf.run(capture=True)
    torch.cuda.synchronize()
    for j, (w, pr, (p1, p2, g)) in enumerate(zip(prods, pf.pricers, pf.results())):
        z = pr.z.cpu().numpy().astype(np.float64).reshape(w.steps, w.factors, w.paths)
        _check_frozen(torch, w, pr, z, p2, g, f"product {j} {w.name} k={pf.k[j]}")


def test_config5_portfolio_at_full_size_vs_fp64_oracle(torch_cuda):
    """The 64-product portfolio at its own shapes (10-40 rates, 4-20
    exercise dates, 1-5 factors, 2^17 Sobol paths each), batched launches in
    one graph with the autotuner's choices: every product's first-pass price
    against the FP64 first pass on the GPU's path quantities, and its frozen
    price and 2N Greeks against the FP64 oracle on the same Sobol points."""
    torch = torch_cuda
    from useful_code_amd import autotune
    prods = configs.bermudan_portfolio()
    assert len(prods) == 64 and all(w.paths == 1 << 17 for w in prods)
    pf = pipeline.PortfolioPricer(prods, checkpoint_every=4, n_streams=4)
    assert pf.batched
    pf.apply_tuning(autotune.tune_portfolio(pf))
    print("config-5 portfolio: 64 products x 2^17 paths, graph replay", flush=True)
    pf.run(capture=True)
    pf.run(capture=True)
    torch.cuda.synchronize()
    worst = 0.0
    for j, (w, pr, (p1, p2, g)) in enumerate(zip(prods, pf.pricers, pf.results())):
        E, P = pr.E, w.paths
        bvn = _norm(w, pr.bv.cpu().numpy().reshape(E, 2, P))
        ev = pr.ev.cpu().numpy().astype(np.float64).reshape(E, P)
        num = pr.num.cpu().numpy().astype(np.float64).reshape(E, P)
        p64, _ = O.ls_first_pass(bvn, np.full(E, 2, np.int32), np.zeros_like(ev), ev, num,
                                 itm_only=True)
        assert abs(p1 - p64) < 1e-5, (j, p1, p64)
        z = pr.z.cpu().numpy().astype(np.float64).reshape(w.steps, w.factors, P)
        _, err = _check_frozen(torch, w, pr, z, p2, g,
                               f"product {j} {w.name} k={pf.k[j]} first pass {p1:.8f}/{p64:.8f}")
        worst = max(worst, err)
    print(f"config-5: worst Greek error / allowance {worst:.2e} over 64 products", flush=True)
