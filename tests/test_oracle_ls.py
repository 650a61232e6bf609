"""Pins the Longstaff-Schwartz part of the FP64 oracle: basis layouts, the
normal-equation solve, the backward induction against the reference's own
gold routine (oracle/_ref, built from the reference sources in this
container), the Bermudan path quantities against the staged curve routines
(themselves pinned against the gold), and limits with closed forms."""
import numpy as np
import pytest

import oracle as O

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference gold library not built")


def test_basis_function_layouts():
    x, y, z, w = 0.3, -1.7, 2.5, 0.9
    assert O.basis_count(0) == 1 and O.basis_count(1) == 3 and O.basis_count(2) == 6
    assert O.basis_count(3) == 10 and O.basis_count(4) == 13
    assert O.basis_count(3, use_cross=False) == 7
    assert np.allclose(O.basis_functions([x]), [1, x, x * x])
    assert np.allclose(O.basis_functions([x, y]), [1, x, y, x * x, x * y, y * y])
    assert np.allclose(O.basis_functions([x, y, z]),
                       [1, x, y, z, x * y, y * z, z * x, x * x, y * y, z * z])
    # four variables: neighbour cross terms only (x0x1, x1x2, x2x3, x3x0)
    assert np.allclose(O.basis_functions([x, y, z, w]),
                       [1, x, y, z, w, x * y, y * z, z * w, w * x, x * x, y * y, z * z, w * w])
    assert np.allclose(O.basis_functions([x, y], use_cross=False), [1, x, x * x, y, y * y])


def test_cholesky_solve_matches_numpy():
    rng = np.random.RandomState(5)
    for n in (1, 6, 13):
        G = rng.normal(size=(n + 3, n))
        A = G.T @ G + 0.1 * np.eye(n)
        b = rng.normal(size=n)
        x = O.cholesky_solve(A, b)
        assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)
    A = np.ones((3, 3))  # rank one: a zero pivot
    assert O.cholesky_solve(A, np.ones(3)) is None


def _small_bermudan(P=2048, N=10, F=3, seed=11, K=0.05):
    rt = 0.5 * np.arange(1, N + 2)
    vols = np.full(N, 0.2)
    A = O.flat_vol_pseudo_roots(rt, rt[:N], vols, F, 0.0, 0.1)
    rng = np.random.RandomState(seed)
    f0 = 0.05 + rng.uniform(-0.005, 0.005, N)
    z = rng.normal(size=(N, F, P))
    m = O.caplet_model(f0, np.full(N, 0.5), np.zeros(N), A, vols, K, P)
    ex = np.arange(0, N, 2, dtype=np.int32)
    return m, z, ex, f0, A


def test_bermudan_paths_equal_staged_curve_routines():
    m, z, ex, f0, A = _small_bermudan(P=96)
    N, P = m.n_rates, m.paths
    bv, ev, num = O.bermudan_paths(m, z, ex)
    taus = np.full(N, 0.5)
    rates, _, _ = O.lmm_pc_evolve(f0, taus, z, A, np.zeros(N), np.arange(N, dtype=np.int32))
    L = O.lib()
    alive = np.arange(N, dtype=np.int32)
    disc = np.zeros(N * (N + 1) * P)
    L.orc_discount_ratios(np.ascontiguousarray(rates).reshape(-1), taus, alive, P, N, N, disc)
    ann = np.zeros(N * N * P)
    L.orc_coterminal_annuities(disc, taus, alive, P, N, N, ann)
    sr = np.zeros(N * N * P)
    L.orc_coterminal_swap_rates(disc, ann, alive, P, N, N, sr)
    numer = np.zeros(N * P)
    L.orc_spot_numeraires(disc, P, N, numer)
    ann, sr, numer = ann.reshape(N, N, P), sr.reshape(N, N, P), numer.reshape(N, P)
    for e, s in enumerate(ex):
        assert np.allclose(bv[e, 0], sr[s, s], rtol=1e-12)
        assert np.allclose(bv[e, 1], rates[s, s], rtol=1e-12)
        assert np.allclose(num[e], numer[s], rtol=1e-12)
        want = np.maximum((sr[s, s] - m.strike) * ann[s, s], 0.0)
        assert np.allclose(ev[e], want, rtol=1e-11, atol=1e-16)


@needs_ref
@pytest.mark.parametrize("use_cross", [True, False])
@pytest.mark.parametrize("seed,P", [(11, 2048), (12, 8192)])
def test_ls_first_pass_matches_reference_gold(use_cross, seed, P):
    # Both solve by the truncated SVD (1e-8 cut) of the same mean-scaled
    # normal equations.  On raw rate-level variables (x ~ 0.05) the quadratic
    # basis is badly scaled and the cut drops real directions in both, so the
    # oracle on raw variables equals the gold without normalisation; on the
    # well-scaled (x - 0.05) * 100 the cut is inactive, the fit is the exact
    # least-squares one, and it equals the gold that standardises its
    # variables (normalise = true), where the cut is inactive too.  The two
    # prices differ from each other by a few percent (the raw cut's effect).
    m, z, ex, _, _ = _small_bermudan(P=P, seed=seed)
    bv, ev, num = O.bermudan_paths(m, z, ex)
    E, P = ev.shape
    ves = np.full(E, 2, np.int32)
    ves[-1] = 1  # the last date regresses on one variable, as near maturity in the reference
    cash = np.zeros((E, P))
    cash[1::2] = 1e-3 * np.cos(np.arange(P))  # exercise the cash-flow terms too
    price, _ = O.ls_first_pass(bv, ves, cash, ev, num, use_cross=use_cross)
    rprice, _ = O.ref_ls_first_pass(bv, ves, cash, ev, num, use_cross=use_cross, normalise=False)
    assert abs(price - rprice) <= 1e-12 * abs(rprice)
    price_n, _ = O.ls_first_pass((bv - 0.05) * 100.0, ves, cash, ev, num, use_cross=use_cross)
    rprice_n, _ = O.ref_ls_first_pass(bv, ves, cash, ev, num, use_cross=use_cross, normalise=True)
    assert abs(price_n - rprice_n) <= 1e-12 * abs(rprice_n)
    assert abs(price_n - price) > 1e-4 * price  # the raw cut binds: a real difference


def test_ls_first_pass_is_invariant_to_affine_rescaling_of_the_variables():
    # Where the 1e-8 cut drops nothing (well-scaled variables) the fit is the
    # exact least-squares one, which an affine change of each variable leaves
    # unchanged (the quadratic basis with cross terms spans the same functions).
    m, z, ex, _, _ = _small_bermudan(P=4096, seed=3)
    bv, ev, num = O.bermudan_paths(m, z, ex)
    E, P = ev.shape
    ves = np.full(E, 2, np.int32)
    cash = np.zeros((E, P))
    a, _ = O.ls_first_pass((bv - 0.05) * 100.0, ves, cash, ev, num, itm_only=True)
    b, _ = O.ls_first_pass((bv - 0.045) * 70.0 + 0.3, ves, cash, ev, num, itm_only=True)
    assert abs(a - b) <= 1e-12 * a


def test_ls_itm_only_equals_all_paths_when_every_path_is_in_the_money():
    rng = np.random.RandomState(2)
    E, P = 4, 500
    bv = rng.uniform(0.02, 0.08, size=(E, 2, P))
    ev = rng.uniform(0.001, 0.02, size=(E, P))  # strictly positive
    num = np.cumprod(1 + 0.5 * rng.uniform(0.03, 0.06, size=(E, P)), axis=0)
    cash = np.zeros((E, P))
    ves = np.full(E, 2, np.int32)
    a = O.ls_first_pass(bv, ves, cash, ev, num, itm_only=False)
    b = O.ls_first_pass(bv, ves, cash, ev, num, itm_only=True)
    assert a[0] == b[0] and np.array_equal(a[1], b[1])


def test_ls_itm_only_regresses_on_the_in_the_money_rows():
    # One date after the last: regression target is the deflated next value.
    m, z, ex, _, _ = _small_bermudan(P=4096)
    bv, ev, num = O.bermudan_paths(m, z, ex)
    E, P = ev.shape
    ves = np.full(E, 2, np.int32)
    cash = np.zeros((E, P))
    bvn = (bv - 0.05) * 100.0  # well scaled: the 1e-8 cut drops nothing
    _, coeffs = O.ls_first_pass(bvn, ves, cash, ev, num, itm_only=True)
    # Date E-2's coefficients are the least-squares fit of (ev_{E-1} B_{E-2}/B_{E-1})
    # on the ITM rows of date E-2 (the last date always exercises when ITM).
    e = E - 2
    y = np.where(ev[E - 1] > 0, ev[E - 1], 0.0) * num[e] / num[E - 1]
    itm = ev[e] > 0
    X = np.stack([O.basis_functions(bvn[e, :, p]) for p in np.nonzero(itm)[0]])
    beta, *_ = np.linalg.lstsq(X, y[itm], rcond=None)
    assert np.allclose(X @ coeffs[e], X @ beta, rtol=1e-8, atol=1e-12)


def test_ls_one_date_is_the_european_price():
    m, z, ex, _, _ = _small_bermudan(P=4096)
    bv, ev, num = O.bermudan_paths(m, z, ex)
    one = slice(2, 3)
    for itm in (False, True):
        price, coeffs = O.ls_first_pass(bv[one], [2], np.zeros((1, ev.shape[1])), ev[one],
                                        num[one], itm_only=itm)
        assert np.all(coeffs == 0.0)  # the regression target is identically zero
        assert np.isclose(price, np.mean(ev[2] / num[2]), rtol=1e-13)


def test_ls_second_pass_under_first_pass_strategy_is_close_and_bounded():
    m, z, ex, _, _ = _small_bermudan(P=8192)
    bv, ev, num = O.bermudan_paths(m, z, ex)
    E, P = ev.shape
    ves = np.full(E, 2, np.int32)
    cash = np.zeros((E, P))
    price1, coeffs = O.ls_first_pass(bv, ves, cash, ev, num, itm_only=True)
    price2 = O.ls_second_pass(bv, ves, cash, ev, num, coeffs, itm_only=True)
    # exercising never beats perfect foresight, and beats every fixed date
    foresight = np.mean(np.max(ev / num, axis=0))
    best_fixed = max(np.mean(ev[e] / num[e]) for e in range(E))
    assert best_fixed - 1e-12 <= price2 <= foresight
    assert abs(price1 - price2) < 0.02 * price2


def test_ls_second_pass_on_the_regression_paths_equals_the_first_pass():
    # With no cash flows the first pass's continuation value at date 0 is the
    # deflated exercise value at the first date its own decisions exercise,
    # which is exactly what the frozen-strategy pass collects on the same paths.
    m, z, ex, _, _ = _small_bermudan(P=4096, seed=7)
    bv, ev, num = O.bermudan_paths(m, z, ex)
    E, P = ev.shape
    ves = np.full(E, 2, np.int32)
    cash = np.zeros((E, P))
    for itm in (False, True):
        p1, coeffs = O.ls_first_pass(bv, ves, cash, ev, num, itm_only=itm)
        p2 = O.ls_second_pass(bv, ves, cash, ev, num, coeffs, itm_only=itm)
        assert abs(p1 - p2) <= 1e-14 * p1


def _frozen_setup(P=2048, seed=21):
    m, z, ex, f0, A = _small_bermudan(P=P, seed=seed)
    bv, ev, num = O.bermudan_paths(m, z, ex)
    E = len(ex)
    shift, scale = np.full((E, 2), 0.05), np.full((E, 2), 100.0)
    bvn = (bv - shift[:, :, None]) * scale[:, :, None]
    _, coeffs = O.ls_first_pass(bvn, np.full(E, 2, np.int32), np.zeros_like(ev), ev, num,
                                itm_only=True)
    return m, z, ex, f0, A, bv, ev, num, coeffs, shift, scale


def test_frozen_bermudan_value_equals_second_pass():
    m, z, ex, f0, A, bv, ev, num, coeffs, shift, scale = _frozen_setup()
    E, P = ev.shape
    fs = O.frozen_strategy(ex, coeffs, shift, scale)
    vals, exs, _, _ = O.bermudan_frozen_paths(m, fs, z, adjoint=False)
    bvn = (bv - shift[:, :, None]) * scale[:, :, None]
    want = O.ls_second_pass(bvn, np.full(E, 2, np.int32), np.zeros_like(ev), ev, num, coeffs,
                            itm_only=True)
    assert abs(vals.mean() - want) <= 1e-13 * want
    assert set(np.unique(exs)) <= set([-1] + list(ex))


def test_frozen_bermudan_adjoint_matches_central_fd_with_frozen_decisions():
    m, z, ex, f0, A, bv, ev, num, coeffs, shift, scale = _frozen_setup(P=512, seed=22)
    N, P = f0.size, z.shape[2]
    taus, disp, vols = np.full(N, 0.5), np.zeros(N), np.full(N, 0.2)
    fs = O.frozen_strategy(ex, coeffs, shift, scale)
    vals, ex0, d, v = O.bermudan_frozen_paths(m, fs, z)
    h = 1e-7
    for i in range(N):
        res = []
        for sgn in (1, -1):
            fb = f0.copy()
            fb[i] += sgn * h
            mb = O.caplet_model(fb, taus, disp, A, vols, m.strike, P)
            vb, exb, _, _ = O.bermudan_frozen_paths(mb, fs, z, adjoint=False)
            assert np.array_equal(exb, ex0)  # no decision moves at this bump size
            res.append(vb.sum())
        fd = (res[0] - res[1]) / (2 * h)
        assert abs(fd - d[i]) <= 1e-6 * np.max(np.abs(d)) + 1e-9
    # vegas: A(s, r, f) scales with sigma_r (flat-vol pseudo-roots are linear in the vols)
    for i in range(N):
        res = []
        for sgn in (1, -1):
            Ab = A.copy()
            Ab[:, i, :] *= (0.2 + sgn * h) / 0.2
            vb_ = vols.copy()
            vb_[i] += sgn * h
            mb = O.caplet_model(f0, taus, disp, Ab, vb_, m.strike, P)
            vb, exb, _, _ = O.bermudan_frozen_paths(mb, fs, z, adjoint=False)
            assert np.array_equal(exb, ex0)
            res.append(vb.sum())
        fv = (res[0] - res[1]) / (2 * h)
        assert abs(fv - v[i]) <= 1e-6 * np.max(np.abs(v)) + 1e-9


def test_bermudan_given_steps_replays_the_strategy_and_fixed_dates():
    """Exercise steps given per path: the strategy's own steps reproduce the
    frozen pass exactly (values and Greeks); 'never' prices zero; one date
    for every path prices that date's exercise value deflated by the
    numeraire, max(ev, 0)/B, path by path."""
    m, z, ex, f0, A, bv, ev, num, coeffs, shift, scale = _frozen_setup(P=1024, seed=23)
    E, P = ev.shape
    fs = O.frozen_strategy(ex, coeffs, shift, scale)
    vals, exs, d, v = O.bermudan_frozen_paths(m, fs, z)
    gv, gd, gvv = O.bermudan_given_paths(m, fs, z, exs)
    assert np.array_equal(gv, vals)
    # per-thread partial sums are combined in thread order: equal to rounding
    assert np.allclose(gd, d, rtol=1e-13, atol=0) and np.allclose(gvv, v, rtol=1e-13, atol=0)
    nv, nd, nvv = O.bermudan_given_paths(m, fs, z, np.full(P, -1, np.int32))
    assert not nv.any() and not nd.any() and not nvv.any()
    for e in (0, E // 2, E - 1):
        fv, _, _ = O.bermudan_given_paths(m, fs, z, np.full(P, ex[e], np.int32), adjoint=False)
        assert np.allclose(fv, np.maximum(ev[e], 0) / num[e], rtol=1e-12, atol=1e-15)
    # a moved decision changes exactly that path's value
    j = int(np.argmax(exs >= 0))
    moved = exs.copy()
    moved[j] = -1
    mv, _, _ = O.bermudan_given_paths(m, fs, z, moved, adjoint=False)
    assert mv[j] == 0.0 and vals[j] > 0.0
    assert np.array_equal(np.delete(mv, j), np.delete(vals, j))


def test_bermudan_given_exercise_at_the_payoff_kink_takes_the_in_the_money_derivative():
    """A path given as exercised where its FP64 exercise value is a hair below
    zero (within 1e-7, under the FP32 resolution of (SR - K) A) gets the
    in-the-money side's payoff derivative, like an FP32 caller that saw it in
    the money; clearly out of the money it gets none."""
    m, z, ex, f0, A = _small_bermudan(P=1, seed=31)
    N = f0.size
    taus, disp, vols = np.full(N, 0.5), np.zeros(N), np.full(N, 0.2)
    E = len(ex)
    fs = O.frozen_strategy(ex, np.zeros((E, 6)), np.full((E, 2), 0.05), np.full((E, 2), 100.0))
    low = 0.0
    ml = O.caplet_model(f0, taus, disp, A, vols, low, 1)
    bv, ev, _ = O.bermudan_paths(ml, z, ex)
    sr = bv[0, 0, 0]
    ann = ev[0, 0] / (sr - low)  # the annuity at the first date
    out = {}
    for tag, gap in (("itm", -5e-8), ("kink", 5e-8), ("otm", 1e-5)):
        mk = O.caplet_model(f0, taus, disp, A, vols, sr + gap / ann, 1)
        out[tag] = O.bermudan_given_paths(mk, fs, z, np.array([ex[0]], np.int32))
    v_itm, d_itm, g_itm = out["itm"]
    v_kink, d_kink, g_kink = out["kink"]
    assert v_itm[0] > 0.0 and v_kink[0] == 0.0
    assert np.any(d_itm != 0.0)
    assert np.allclose(d_kink, d_itm, rtol=1e-5, atol=0) and np.allclose(g_kink, g_itm, rtol=1e-5)
    v_otm, d_otm, g_otm = out["otm"]
    assert v_otm[0] == 0.0 and not d_otm.any() and not g_otm.any()
