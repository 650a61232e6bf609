"""The multi-regression LS first pass (ls_multi_first_pass_gpu) on the GPU
against the FP64 oracle (orc_ls_multi_first_pass, itself pinned to the
reference gold generateRegressionCoefficientsViaLSMultiQuadratic_gold by
test_oracle_golden.py) on the same FP32 inputs.

The GPU computes in FP64 from the FP32 inputs, so the selector's band search
sees the same counts as the oracle and takes the same steps; what differs is
the summation order of the FP64 sums.  Bars: price 1e-12 relative; means
1e-12 relative; coefficients 1e-9 of the largest and cuts 1e-12 relative for
the normalised basis.  Standard deviations 1e-9 relative: the gold's
one-pass sqrt(E[x^2] - mean^2) cancels by (mean / sd)^2, about 25 for the
whole sample and more inside the nested bands, which magnifies the
summation-order rounding.  The raw basis's normal equations are
conditioned near the solver's 1e-8 cut, so the dropped eigen-directions make
its coefficients (not the fitted values) sensitive to the last bits of the
sums: 1e-6 of the largest coefficient, and the cuts, proportional to the
standard deviation of the net values, 1e-8 relative."""
import importlib.util
import os

import numpy as np
import pytest

import oracle as O
from useful_code_amd import koo

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _make_golden():
    spec = importlib.util.spec_from_file_location("make_golden",
                                                  os.path.join(HERE, "golden", "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


M = _make_golden()


def dev(torch, x, dtype=None):
    dtype = dtype or torch.float32
    return torch.as_tensor(np.ascontiguousarray(x)).to(dtype).cuda()


def f32(c):
    """The case's inputs rounded to FP32 (what the GPU is given)."""
    return {k: (v.astype(np.float32).astype(np.float64) if v.dtype == np.float64 else v)
            for k, v in c.items()}


def run_gpu(torch, c, cross, depth, norm, cutoff=10, selector=M.SELECTOR, kind=0):
    E, mv, P = c["bv"].shape
    maxd = O.basis_count(mv, cross)
    ws = torch.empty(koo.ls_multi_workspace(P), dtype=torch.uint8, device="cuda")
    co = torch.full((E * depth * maxd,), np.nan, dtype=torch.float64, device="cuda")
    lo = torch.full((E * depth,), np.nan, dtype=torch.float64, device="cuda")
    up = torch.full_like(lo, np.nan)
    me = torch.full((E * depth * mv,), np.nan, dtype=torch.float64, device="cuda")
    sd = torch.full_like(me, np.nan)
    price = torch.zeros(1, dtype=torch.float64, device="cuda")
    print(f"ls_multi E={E} P={P} vars={mv} cross={cross} depth={depth} norm={norm} "
          f"cutoff={cutoff} selector={kind}", flush=True)
    koo.ls_multi_first_pass_gpu(E, dev(torch, c["bv"]), c["ves"], mv, dev(torch, c["cash"]),
                                dev(torch, c["ev"]), dev(torch, c["num"]), P, cross, norm, depth,
                                cutoff, *selector, ws, co, lo, up, me, sd, price, selector=kind)
    torch.cuda.synchronize()
    return (float(price.item()), co.cpu().numpy().reshape(E, depth, maxd),
            lo.cpu().numpy().reshape(E, depth), up.cpu().numpy().reshape(E, depth),
            me.cpu().numpy().reshape(E, depth, mv), sd.cpu().numpy().reshape(E, depth, mv))


def check(got, want, norm):
    price, co, lo, up, me, sd = got
    wprice, wco, wlo, wup, wme, wsd = want
    scale = max(1.0, np.abs(wco).max())
    rel = lambda a, b: float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))  # noqa: E731
    print(f"price rel {abs(price - wprice) / abs(wprice):.2e} coeff {np.abs(co - wco).max() / scale:.2e}"
          f" cuts rel {rel(up, wup) if np.any(wup) else 0.0:.2e} "
          f"sds rel {rel(sd[wsd != 0], wsd[wsd != 0]) if np.any(wsd) else 0.0:.2e}", flush=True)
    cut_tol = 1e-12 if norm else 1e-8
    assert price == pytest.approx(wprice, rel=1e-12, abs=1e-15)
    assert np.abs(co - wco).max() <= (1e-9 if norm else 1e-6) * scale
    assert np.allclose(lo, wlo, rtol=cut_tol, atol=1e-15)
    assert np.allclose(up, wup, rtol=cut_tol, atol=1e-15)
    assert np.allclose(me, wme, rtol=1e-12, atol=1e-15)
    assert np.allclose(sd, wsd, rtol=1e-9, atol=1e-15)


@pytest.mark.parametrize("norm", [0, 1])
@pytest.mark.parametrize("cross,depth", sorted(M.LS_MULTI_CASES))
def test_multi_first_pass_matches_oracle_on_the_gold_cases(torch_cuda, cross, depth, norm):
    """The twelve cases pinned against the reference gold (one to three
    regressions, two or three variables, with and without cross terms,
    normalised or raw)."""
    c = f32(M.ls_multi_case(*M.LS_MULTI_CASES[(cross, depth)]))
    got = run_gpu(torch_cuda, c, cross, depth, norm)
    want = O.ls_multi_first_pass(c["bv"], c["ves"], c["cash"], c["ev"], c["num"], cross, depth,
                                 10, *M.SELECTOR, normalise=norm)
    check(got, want, norm)
    if depth > 1:
        assert np.any(got[3][:, 1] != 0)  # nested regressions ran


def test_multi_first_pass_large_normalised(torch_cuda):
    """2^18 paths, 10 dates, three variables with cross terms, three
    regressions per date (band search over 65536 sorted keys and more)."""
    c = f32(M.ls_multi_case(31, E=10, P=2 ** 18, max_vars=3, last_vars=2))
    got = run_gpu(torch_cuda, c, 1, 3, 1)
    want = O.ls_multi_first_pass(c["bv"], c["ves"], c["cash"], c["ev"], c["num"], 1, 3, 10,
                                 *M.SELECTOR, normalise=1)
    check(got, want, 1)
    assert np.all(got[3][:, :2] > 0)  # two bands per date


def test_multi_first_pass_edge_cases(torch_cuda):
    """A date with no variables (constant basis), a path count off every
    block multiple, and a cutoff that stops the nesting on some dates."""
    c = f32(M.ls_multi_case(5, E=5, P=5003, max_vars=2, last_vars=1))
    c["ves"] = np.array([2, 0, 2, 1, 1], np.int32)
    for cutoff in (10, 1300):
        got = run_gpu(torch_cuda, c, 0, 3, 1, cutoff=cutoff)
        want = O.ls_multi_first_pass(c["bv"], c["ves"], c["cash"], c["ev"], c["num"], 0, 3,
                                     cutoff, *M.SELECTOR, normalise=1)
        check(got, want, 1)
    assert np.any(got[3][:, 1] == 0)  # the large cutoff stopped some nesting


def test_multi_first_pass_wide_window_keeps_every_path(torch_cuda):
    """A selector window above one keeps every member (the reference's
    10000 cut) and the nested regressions repeat the first one."""
    c = f32(M.ls_multi_case(9, E=4, P=4000, max_vars=2, last_vars=1))
    sel = (0.5, 1.5, 2.5, 0.8)
    got = run_gpu(torch_cuda, c, 1, 2, 1, selector=sel)
    want = O.ls_multi_first_pass(c["bv"], c["ves"], c["cash"], c["ev"], c["num"], 1, 2, 10,
                                 *sel, normalise=1)
    check(got, want, 1)
    assert np.all(got[3] == 10000.0)


@pytest.mark.parametrize("n_sds", [0.5, 1.5])
def test_multi_first_pass_sd_selector(torch_cuda, n_sds):
    """The standard-deviation selector (open band of n_sds standard
    deviations) at three depths, cross terms, normalised."""
    c = f32(M.ls_multi_case(13, E=6, P=30000, max_vars=3, last_vars=1))
    sel = (0.25, 0.26, n_sds, 0.8)
    got = run_gpu(torch_cuda, c, 1, 3, 1, selector=sel, kind=1)
    want = O.ls_multi_first_pass(c["bv"], c["ves"], c["cash"], c["ev"], c["num"], 1, 3, 10,
                                 *sel, normalise=1, selector=1)
    check(got, want, 1)
    assert np.all(got[3][:, :2] > 0)


def test_multi_first_pass_searches_past_the_ladder_and_trees(torch_cuda):
    """Net values that drive the band search past what its histograms hold,
    so the block-wide fallback counts decide: on one date 1% of the paths
    exercise for 1e6, so the standard deviation is about 1e5 while most net
    values are below 0.1; the shrink runs about 74 steps (past the 64-step
    ladder) and the bisection from [0.01, 2.5e5] about 31 (past the two
    ten-step trees).  (The growth side passes the ladder only when
    |mean| / sd > 4e6, where the gold's one-pass variance E[x^2] - mean^2
    has lost most of its digits -- at |mean| / sd = 1e8 the FP64 oracle's sd
    is NaN -- so no parity test can pin that branch.)"""
    c = f32(M.ls_multi_case(17, E=5, P=20000, max_vars=2, last_vars=1))
    rng = np.random.RandomState(5)
    c["ev"][1, rng.choice(20000, 200, replace=False)] = 1.0e6
    for cross in (0, 1):
        got = run_gpu(torch_cuda, c, cross, 2, 1)
        want = O.ls_multi_first_pass(c["bv"], c["ves"], c["cash"], c["ev"], c["num"], cross, 2,
                                     10, *M.SELECTOR, normalise=1)
        check(got, want, 1)
        assert 0 < got[3][1, 0] < 1.0  # the search ended far below its start


def test_multi_first_pass_many_exact_zero_net_values(torch_cuda):
    """The last date with no cash flow and 60% of the paths out of the money
    (exercise value 0): the regression target is 0, so 60% of the net values
    are exactly 0, more than the selector's upper fraction at every cut, and
    the gold's shrink loop runs all of its 1000 steps (the Bermudan
    swaption's own first pass does this)."""
    c = f32(M.ls_multi_case(19, E=4, P=20000, max_vars=2, last_vars=1))
    rng = np.random.RandomState(3)
    c["cash"][3, :] = 0.0
    c["ev"][3, rng.rand(20000) < 0.6] = 0.0
    got = run_gpu(torch_cuda, c, 1, 2, 1)
    want = O.ls_multi_first_pass(c["bv"], c["ves"], c["cash"], c["ev"], c["num"], 1, 2, 10,
                                 *M.SELECTOR, normalise=1)
    check(got, want, 1)


def test_multi_first_pass_is_deterministic(torch_cuda):
    c = f32(M.ls_multi_case(7, E=6, P=20000, max_vars=2, last_vars=1))
    a = run_gpu(torch_cuda, c, 1, 3, 1)
    b = run_gpu(torch_cuda, c, 1, 3, 1)
    assert a[0] == b[0]
    for x, y in zip(a[1:], b[1:]):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("P,sel", [(20000, 0), (1 << 18, 0), (20000, 1)])
def test_multi_persistent_pass_equals_per_phase_launches(torch_cuda, P, sel):
    """The pass as one persistent launch (KOO_LS_MULTI_PERSISTENT=1: grid
    barriers between the phases) and as one launch per phase (the default)
    run the same bodies over the same blocks, so
    every output agrees bit for bit; a case whose search bisects and whose
    deeper depths are skipped included (the persistent pass skips unreached
    depths and unneeded bisection stages)."""
    c = f32(M.ls_multi_case(23, E=8, P=P, max_vars=2, last_vars=1))
    selector = M.SELECTOR if sel == 0 else (0.25, 0.26, 1.5, 0.8)
    runs = []
    for mode in ("1", "0"):
        os.environ["KOO_LS_MULTI_PERSISTENT"] = mode
        try:
            runs.append(run_gpu(torch_cuda, c, 1, 3, 1, selector=selector, kind=sel))
        finally:
            os.environ.pop("KOO_LS_MULTI_PERSISTENT", None)
    a, b = runs
    print(f"persistent price {a[0]:.15f} per phase {b[0]:.15f}", flush=True)
    assert np.isfinite(a[0]) and a[0] == b[0]
    for x, y in zip(a[1:], b[1:]):
        assert np.array_equal(x, y)
