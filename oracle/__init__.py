"""FP64 CPU oracle for the LMM Monte Carlo pricing path (test infrastructure).

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may
import this package, and only as the checker: the product library
(useful_code_amd) never imports, links or calls it.

``lib()`` loads ``oracle/liboracle.so`` (the C restatement in koo_oracle.c);
``ref()`` loads ``oracle/_ref/libkoo_ref.so`` (the reference's own gold
routines, built from the reference sources by oracle/Makefile) when present.
The thin numpy wrappers below keep the reference's array layouts.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_c_int, _c_dbl, _c_u32, _c_i64 = ctypes.c_int, ctypes.c_double, ctypes.c_uint32, ctypes.c_int64


def build(ref=True):
    """Compile liboracle.so (and the reference gold library when the
    reference sources are present)."""
    targets = ["all"]
    if ref and os.path.isdir("/root/reference/kooderive-code-r256/LINUX/gold"):
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", _HERE] + targets, check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build(ref=False)
        _LIB = ctypes.CDLL(path)
        _declare(_LIB)
    return _LIB


def ref_available():
    return os.path.exists(os.path.join(_HERE, "_ref", "libkoo_ref.so"))


def ref():
    global _REF
    if _REF is None:
        _REF = ctypes.CDLL(os.path.join(_HERE, "_ref", "libkoo_ref.so"))
        _declare_ref(_REF)
    return _REF


class FrozenStrategy(ctypes.Structure):
    _fields_ = [("n_ex", _c_int), ("ex_steps", ctypes.c_void_p), ("coeffs", ctypes.c_void_p),
                ("shift", ctypes.c_void_p), ("scale", ctypes.c_void_p), ("itm_only", _c_int)]


class BasketModel(ctypes.Structure):
    _fields_ = [("assets", _c_int), ("steps", _c_int), ("paths", _c_int),
                ("x0", ctypes.c_void_p), ("vols", ctypes.c_void_p), ("chol", ctypes.c_void_p),
                ("weights", ctypes.c_void_p), ("rate", _c_dbl), ("dt", _c_dbl),
                ("strike", _c_dbl)]


class CapletModel(ctypes.Structure):
    _fields_ = [("n_rates", _c_int), ("factors", _c_int), ("paths", _c_int),
                ("f0", ctypes.c_void_p), ("taus", ctypes.c_void_p), ("disp", ctypes.c_void_p),
                ("A", ctypes.c_void_p), ("sigma", ctypes.c_void_p), ("strike", _c_dbl)]


def _declare(L):
    sig = {
        "orc_sobol_directions": (_c_int, [_c_int, _u32p, _u32p, _u32p, _c_int, _u32p]),
        "orc_sobol_ints": (None, [_c_int, _c_int, _c_u32, _u32p, _u32p]),
        "orc_scramble": (None, [_u32p, _u32p, _c_int, _c_int]),
        "orc_invnorm_sb_ref_f32": (ctypes.c_float, [_c_u32]),
        "orc_invnorm_u32": (_c_dbl, [_c_u32]),
        "orc_invnorm": (_c_dbl, [_c_dbl]),
        "orc_invnorm_u32_array": (None, [_u32p, _f64p, _c_i64]),
        "orc_bridge_allocate_diagonal": (None, [_c_int, _c_int, _i32p]),
        "orc_bridge_tables": (None, [_c_int, _i32p, _i32p, _i32p, _f64p, _f64p, _f64p]),
        "orc_bridge_transform": (None, [_c_int, _c_int, _c_int, _i32p, _f64p, _f64p]),
        "orc_bridge_pow2": (None, [_c_int, _c_int, _f64p, _f64p]),
        "orc_ls_multi_first_pass": (_c_dbl, [_c_int, _f64p, _i32p, _c_int, _f64p, _f64p, _f64p,
                                             _c_int, _c_int, _c_int, _c_int, _c_int, _c_int,
                                             _c_dbl, _c_dbl,
                                             _c_dbl, _c_dbl, _f64p, _f64p, _f64p, _f64p, _f64p]),
        "orc_sym_solve_truncated": (_c_int, [_c_int, _f64p, _f64p, _c_dbl, _f64p]),
        "orc_sym_solve_cut": (_c_int, [_c_int, _f64p, _f64p, _c_dbl, _f64p]),
        "orc_select_sds": (ctypes.c_long, [_f64p, _i32p, _i32p, _c_int, ctypes.c_long, _c_dbl,
                                           _f64p, _f64p]),
        "orc_select_fraction": (ctypes.c_long, [_f64p, _i32p, _i32p, _c_int, ctypes.c_long,
                                                _c_dbl, _c_dbl, _c_dbl, _c_dbl, _f64p, _f64p]),
        "orc_cashflows_ee": (None, [_c_int, _f64p, _f64p, _f64p, _f64p, _i32p, _c_int, _c_int,
                                    _c_int, ctypes.c_void_p, _f64p, _f64p, _c_int, _f64p, _i32p]),
        "orc_bridge_reorder": (None, [_f64p, _f64p, _c_int, _c_int, _c_int, _c_int]),
        "orc_correlate_drift": (None, [_f64p, _f64p, _f64p, _i32p, _f64p, _c_int, _c_int,
                                       _c_int, _c_int, _c_int]),
        "orc_spot_drifts": (None, [_f64p, _f64p, _f64p, _f64p, _c_int, _c_int, _c_int, _f64p]),
        "orc_lmm_pc_evolve": (None, [_f64p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p,
                                     _i32p, _c_int, _c_int, _c_int, _c_int, _c_int,
                                     _f64p, _f64p, _f64p]),
        "orc_lmm_steps": (None, [_f64p, _f64p, _f64p, _f64p, _f64p, _f64p, _i32p, _c_int, _c_int,
                                 _c_int, _c_int, _c_int, _f64p, _f64p]),
        "orc_discount_ratios": (None, [_f64p, _f64p, _i32p, _c_int, _c_int, _c_int, _f64p]),
        "orc_coterminal_annuities": (None, [_f64p, _f64p, _i32p, _c_int, _c_int, _c_int, _f64p]),
        "orc_coterminal_swap_rates": (None, [_f64p, _f64p, _i32p, _c_int, _c_int, _c_int, _f64p]),
        "orc_swap_rate": (None, [_f64p, _f64p, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int,
                                 _f64p]),
        "orc_spot_numeraires": (None, [_f64p, _c_int, _c_int, _f64p]),
        "orc_forward_extraction": (None, [_f64p, _i32p, _c_int, _c_int, _c_int, _f64p]),
        "orc_swap_flows": (None, [_f64p, _f64p, _f64p, _c_int, _c_int, _f64p]),
        "orc_tarn_flows": (None, [_f64p, _f64p, _f64p, _c_int, _c_int, _f64p, _f64p]),
        "orc_caplet_flows": (None, [_f64p, _f64p, _f64p, _c_int, _c_int, _f64p]),
        "orc_cash_flow_indices_weights": (None, [_f64p, _f64p, _c_int, _i32p, _i32p, _f64p]),
        "orc_cash_flow_discounting": (None, [_i32p, _i32p, _f64p, _f64p, _f64p, _f64p, _c_int,
                                             _c_int, _f64p, _f64p]),
        "orc_cash_flow_discounting_partial": (None, [_c_int, _i32p, _i32p, _i32p, _f64p, _i32p,
                                                     _f64p, _f64p, _f64p, _c_int, _c_int, _f64p]),
        "orc_sum_compensated": (_c_dbl, [_f64p, _c_i64]),
        "orc_sym_eigen": (_c_int, [_c_int, _f64p, _f64p, _f64p]),
        "orc_exp_corr": (None, [_f64p, _c_int, _c_dbl, _c_dbl, _f64p]),
        "orc_pseudo_sqrt": (_c_int, [_c_int, _f64p, _c_int, _c_int, _f64p]),
        "orc_flat_vol_pseudo_roots": (_c_int, [_f64p, _f64p, _f64p, _c_int, _c_int, _c_int,
                                               _c_dbl, _c_dbl, _f64p]),
        "orc_factor_pseudo_roots": (_c_int, [_f64p, _f64p, _c_int, _c_int, _c_dbl, _c_dbl,
                                             _f64p, ctypes.c_void_p]),
        "orc_caplet_price_paths": (None, [ctypes.POINTER(CapletModel), _f64p, _f64p]),
        "orc_caplet_adjoint_paths": (None, [ctypes.POINTER(CapletModel), _f64p, _f64p, _f64p,
                                            _f64p]),
        "orc_set_threads": (None, [_c_int]),
        "orc_get_threads": (_c_int, []),
        "orc_caplet_pipeline": (None, [ctypes.POINTER(CapletModel), _u32p, _c_int,
                                       ctypes.POINTER(_c_dbl), _f64p, _f64p]),
        "orc_basis_count": (_c_int, [_c_int, _c_int]),
        "orc_basis_functions": (None, [_f64p, _c_int, _c_int, _f64p]),
        "orc_cholesky_solve": (_c_int, [_c_int, _f64p, _f64p, _f64p]),
        "orc_ls_first_pass": (_c_dbl, [_c_int, _f64p, _i32p, _c_int, _f64p, _f64p, _f64p, _c_int,
                                       _c_int, _c_int, _f64p]),
        "orc_ls_second_pass": (_c_dbl, [_c_int, _f64p, _i32p, _c_int, _f64p, _f64p, _f64p,
                                        _c_int, _c_int, _c_int, _f64p]),
        "orc_bermudan_paths": (None, [ctypes.POINTER(CapletModel), _f64p, _c_int, _i32p, _f64p,
                                      _f64p, _f64p]),
        "orc_cholesky_lower": (_c_int, [_c_int, _f64p, _f64p]),
        "orc_bermudan_frozen_paths": (None, [ctypes.POINTER(CapletModel),
                                             ctypes.POINTER(FrozenStrategy), _f64p, _c_int, _f64p,
                                             ctypes.c_void_p, _f64p, _f64p]),
        "orc_bermudan_given_paths": (None, [ctypes.POINTER(CapletModel),
                                            ctypes.POINTER(FrozenStrategy), _f64p, _i32p, _c_int,
                                            _f64p, _f64p, _f64p]),
        "orc_correlated_paths": (None, [_f64p, _f64p, _f64p, _i32p, _c_int, _c_int, _c_int,
                                        _c_int, _c_int]),
        "orc_multid_bs_paths": (None, [_f64p, _f64p, ctypes.c_void_p, _f64p, _c_int, _c_int,
                                       _c_int]),
        "orc_basket_weightings": (None, [_f64p, _f64p, _f64p, _c_int, _c_int, _c_int]),
        "orc_basket_asian_paths": (None, [ctypes.POINTER(BasketModel), _f64p, _f64p, _f64p,
                                          _f64p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def _declare_ref(R):
    sig = {
        "ref_sobol_directions": (None, [_c_int, _u32p]),
        "ref_sobol_primitives": (None, [_c_int, _u32p, _u32p, _u32p]),
        "ref_sobol_ints_offset": (None, [_c_int, _c_int, _c_int, _u32p, _u32p]),
        "ref_scramble": (None, [_u32p, _u32p, _u32p, _c_int, _c_int]),
        "ref_invnorm_sb": (ctypes.c_float, [_c_u32]),
        "ref_bridge_allocate_diagonal": (None, [_c_int, _c_int, _i32p]),
        "ref_bridge_pow2": (None, [_c_int, _c_int, _f64p, _f64p]),
        "ref_ls_multi": (_c_dbl, [_c_int, _f64p, _i32p, _c_int, _f64p, _f64p, _f64p, _c_int,
                                  _c_int, _c_int, _c_int, _c_int, _c_dbl, _c_dbl, _c_dbl, _c_dbl, _f64p,
                                  _f64p, _f64p]),
        "ref_cashflows_ee": (None, [_c_int, _f64p, _f64p, _f64p, _c_int, _f64p, _i32p, _c_int,
                                    _c_int, _c_int, ctypes.c_void_p, _f64p, _f64p, _c_int]),
        "ref_bridge_multi": (None, [_c_int, _c_int, _c_int, _c_int, _f64p, _f64p]),
        "ref_correlate_drift": (None, [_f64p, _f64p, _f64p, _c_int, _f64p, _c_int, _c_int,
                                       _c_int, _c_int]),
        "ref_spot_drifts": (None, [_f64p, _f64p, _f64p, _f64p, _c_int, _c_int, _f64p]),
        "ref_lmm_pc": (None, [_f64p, _f64p, _f64p, _f64p, _f64p, _c_int, _c_int, _c_int, _c_int,
                              _f64p, _f64p]),
        "ref_lmm_euler": (None, [_f64p, _f64p, _f64p, _f64p, _f64p, _c_int, _c_int, _c_int,
                                 _c_int, _f64p, _f64p]),
        "ref_discount_ratios": (None, [_f64p, _f64p, _i32p, _c_int, _c_int, _c_int, _f64p]),
        "ref_annuities": (None, [_f64p, _f64p, _i32p, _c_int, _c_int, _c_int, _f64p]),
        "ref_cot_swap_rates": (None, [_f64p, _f64p, _i32p, _c_int, _c_int, _c_int, _f64p]),
        "ref_swap_rate": (None, [_f64p, _f64p, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int,
                                 _c_int, _f64p]),
        "ref_numeraires": (None, [_f64p, _c_int, _c_int, _f64p]),
        "ref_forward_extraction": (None, [_f64p, _i32p, _c_int, _c_int, _c_int, _f64p]),
        "ref_swap_flows": (None, [_f64p, _f64p, _f64p, _c_int, _c_int, _c_int, _f64p]),
        "ref_tarn_flows": (None, [_f64p, _f64p, _f64p, _c_int, _c_int, _c_int, _f64p, _f64p]),
        "ref_cash_flow_indices_weights": (None, [_f64p, _c_int, _f64p, _c_int, _i32p, _i32p,
                                                 _f64p]),
        "ref_cash_flow_discounting_partial": (None, [_c_int, _i32p, _i32p, _i32p, _f64p, _i32p,
                                                     _f64p, _f64p, _f64p, _c_int, _c_int, _f64p]),
        "ref_cash_flow_discounting": (None, [_i32p, _i32p, _f64p, _f64p, _f64p, _f64p, _c_int,
                                             _c_int, _f64p, _f64p]),
        "ref_flat_vol_pseudo_roots": (None, [_f64p, _f64p, _f64p, _c_int, _c_int, _c_int, _c_dbl,
                                             _c_dbl, _f64p]),
        "ref_ls_multi_depth0": (_c_dbl, [_c_int, _f64p, _i32p, _c_int, _f64p, _f64p, _f64p, _c_int,
                                         _c_int, _c_int, _f64p]),
        "ref_correlated_paths": (None, [_f64p, _f64p, _f64p, _c_int, _c_int, _c_int, _c_int]),
        "ref_multid_bs_paths": (None, [_f64p, _f64p, _f64p, _f64p, _c_int, _c_int, _c_int]),
        "ref_basket_weightings": (None, [_f64p, _f64p, _f64p, _c_int, _c_int, _c_int]),
        "ref_swap_pipeline": (_c_dbl, [_c_int, _c_int, _c_int, _c_int, _u32p, _f64p, _f64p, _f64p,
                                       _f64p, _f64p, _c_dbl]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(R, name)
        fn.restype = res
        fn.argtypes = args


# ---------------------------------------------------------------- wrappers

_PRIMITIVES = None


def sobol_primitives():
    """The joe-kuo-6.10200 (degree, a, m[17]) rows read out of the reference's
    compiled gold/sobol_primitives.cpp by tests/golden/make_golden.py."""
    global _PRIMITIVES
    if _PRIMITIVES is None:
        g = np.load(os.path.join(_HERE, "..", "tests", "golden", "sobol_primitives_ref.npz"))
        _PRIMITIVES = (np.ascontiguousarray(g["degree"], np.uint32),
                       np.ascontiguousarray(g["a"], np.uint32),
                       np.ascontiguousarray(g["m"], np.uint32))
    return _PRIMITIVES


def sobol_directions(n_dims, primitives=None):
    deg, a, m = sobol_primitives() if primitives is None else primitives
    if not 1 <= n_dims <= len(deg):
        raise ValueError("n_dims out of table range")
    d = np.zeros(n_dims * 32, np.uint32)
    if lib().orc_sobol_directions(n_dims, deg, a, np.ascontiguousarray(m.reshape(-1)), m.shape[1],
                                  d) != 0:
        raise ValueError("bad primitive rows")
    return d


def sobol_ints(n_vectors, n_dims, offset=0, dirs=None):
    dirs = sobol_directions(n_dims) if dirs is None else dirs
    out = np.zeros(n_vectors * n_dims, np.uint32)
    lib().orc_sobol_ints(n_vectors, n_dims, offset, dirs, out)
    return out.reshape(n_dims, n_vectors)


def scramble(data, scramblers, number_dim):
    data = np.ascontiguousarray(data, np.uint32).copy()
    lib().orc_scramble(data.reshape(-1), np.ascontiguousarray(scramblers, np.uint32),
                       number_dim, data.shape[-1])
    return data


def invnorm_u32(x):
    x = np.ascontiguousarray(x, np.uint32)
    out = np.zeros(x.size, np.float64)
    lib().orc_invnorm_u32_array(x.reshape(-1), out, x.size)
    return out.reshape(x.shape)


def bridge_allocate_diagonal(factors, steps):
    idx = np.zeros(factors * steps, np.int32)
    lib().orc_bridge_allocate_diagonal(factors, steps, idx)
    return idx


def bridge_tables(steps):
    bi, li, ri = (np.zeros(steps, np.int32) for _ in range(3))
    lw, rw, sd = (np.zeros(steps, np.float64) for _ in range(3))
    lib().orc_bridge_tables(steps, bi, li, ri, lw, rw, sd)
    return bi, li, ri, lw, rw, sd


def bridge_transform(normals, steps, factors, indices=None):
    """normals [D][P] -> increments (S,F,P)."""
    normals = np.ascontiguousarray(normals, np.float64)
    paths = normals.shape[-1]
    idx = bridge_allocate_diagonal(factors, steps) if indices is None else np.ascontiguousarray(indices, np.int32)
    out = np.zeros(steps * factors * paths, np.float64)
    lib().orc_bridge_transform(steps, factors, paths, idx, normals.reshape(-1), out)
    return out.reshape(steps, factors, paths)


def exp_corr(times, L, beta):
    t = np.ascontiguousarray(times, np.float64)
    out = np.zeros(t.size * t.size)
    lib().orc_exp_corr(t, t.size, L, beta, out)
    return out.reshape(t.size, t.size)


def pseudo_sqrt(cov, factors, normalise=True):
    cov = np.ascontiguousarray(cov, np.float64)
    n = cov.shape[0]
    out = np.zeros(n * factors)
    lib().orc_pseudo_sqrt(n, cov.reshape(-1), factors, int(normalise), out)
    return out.reshape(n, factors)


def flat_vol_pseudo_roots(rate_times, evol_times, vols, factors, L, beta):
    rt = np.ascontiguousarray(rate_times, np.float64)
    et = np.ascontiguousarray(evol_times, np.float64)
    v = np.ascontiguousarray(vols, np.float64)
    n, s = v.size, et.size
    A = np.zeros(s * n * factors)
    lib().orc_flat_vol_pseudo_roots(rt, et, v, n, s, factors, L, beta, A)
    return A.reshape(s, n, factors)


def factor_pseudo_roots(rate_times, vols, factors, L, beta):
    rt = np.ascontiguousarray(rate_times, np.float64)
    v = np.ascontiguousarray(vols, np.float64)
    n = v.size
    A = np.zeros(n * n * factors)
    B = np.zeros(n * factors)
    lib().orc_factor_pseudo_roots(rt, v, n, factors, L, beta, A, B.ctypes.data)
    return A.reshape(n, n, factors), B.reshape(n, factors)


def fixed_drifts(A):
    return -0.5 * np.sum(np.asarray(A, np.float64) ** 2, axis=2)


def spot_drifts(A_step, rates, taus, disp, alive=0):
    A_step = np.ascontiguousarray(A_step, np.float64)
    n, F = A_step.shape
    out = np.zeros(n)
    lib().orc_spot_drifts(A_step.reshape(-1), np.ascontiguousarray(rates, np.float64),
                          np.ascontiguousarray(taus, np.float64),
                          np.ascontiguousarray(disp, np.float64), F, n, alive, out)
    return out


def lmm_pc_evolve(f0, taus, variates, A, disp, alive, transposed=True, initial_drifts=None,
                  fixed=None):
    """Kernel-semantics PC evolver; variates (S,F,P) if transposed else (S,P,F)."""
    f0 = np.ascontiguousarray(f0, np.float64)
    taus = np.ascontiguousarray(taus, np.float64)
    disp = np.ascontiguousarray(disp, np.float64)
    A = np.ascontiguousarray(A, np.float64)
    S, N, F = A.shape
    v = np.ascontiguousarray(variates, np.float64)
    P = v.shape[2] if transposed else v.shape[1]
    alive = np.ascontiguousarray(alive, np.int32)
    logs0 = np.log(f0 + disp)
    mu0 = spot_drifts(A[0], f0, taus, disp, int(alive[0])) if initial_drifts is None else \
        np.ascontiguousarray(initial_drifts, np.float64)
    fx = fixed_drifts(A) if fixed is None else np.ascontiguousarray(fixed, np.float64)
    rates = np.zeros(S * N * P)
    logs = np.zeros(S * N * P)
    disc = np.zeros(S * (N + 1) * P)
    lib().orc_lmm_pc_evolve(f0, logs0, taus, v.reshape(-1), A.reshape(-1), mu0,
                            fx.reshape(-1), disp, alive, P, F, S, N, int(transposed),
                            rates, logs, disc)
    return rates.reshape(S, N, P), logs.reshape(S, N, P), disc.reshape(S, N + 1, P)


def ls_multi_first_pass(bv, vars_each_step, cash, ev, num, use_cross, max_regressions,
                        lower_path_cutoff, lower_frac, upper_frac, sd_guess, multiplier,
                        normalise=True, selector=0):
    """Multi-regression first pass with the fraction selector (selector 0)
    or the standard-deviation selector (1: sd_guess standard deviations),
    variables normalised per regression or raw.  bv (E, maxVars, P); returns price,
    coefficients (E, R, maxd), lower and upper cuts (E, R), means and sds
    (E, R, maxVars)."""
    bv = np.ascontiguousarray(bv, np.float64)
    E, mv, P = bv.shape
    R = int(max_regressions)
    maxd = basis_count(mv, use_cross)
    co = np.zeros(E * R * maxd)
    lo, up = np.zeros(E * R), np.zeros(E * R)
    me, sd = np.zeros(E * R * mv), np.zeros(E * R * mv)
    f = lambda a: np.ascontiguousarray(a, np.float64).reshape(-1)  # noqa: E731
    price = lib().orc_ls_multi_first_pass(
        E, bv.reshape(-1), np.ascontiguousarray(vars_each_step, np.int32), mv, f(cash), f(ev),
        f(num), P, int(use_cross), int(normalise), R, int(lower_path_cutoff), int(selector),
        float(lower_frac), float(upper_frac), float(sd_guess), float(multiplier), co, lo, up, me,
        sd)
    return (price, co.reshape(E, R, maxd), lo.reshape(E, R), up.reshape(E, R),
            me.reshape(E, R, mv), sd.reshape(E, R, mv))


def select_fraction(net, pre, lower_frac, upper_frac, sd_guess, multiplier):
    """The fraction selector on net values over the pre-selected paths;
    returns (selected mask, count, lower cut, upper cut)."""
    net = np.ascontiguousarray(net, np.float64)
    pre = np.ascontiguousarray(pre, np.int32)
    sel = np.zeros(len(net), np.int32)
    lo, up = np.zeros(1), np.zeros(1)
    c = lib().orc_select_fraction(net, pre, sel, len(net), int(pre.sum()), float(lower_frac),
                                  float(upper_frac), float(sd_guess), float(multiplier), lo, up)
    return sel, int(c), float(lo[0]), float(up[0])


def select_sds(net, pre, n_sds):
    """The standard-deviation selector; returns (mask, count, lower, upper)."""
    net = np.ascontiguousarray(net, np.float64)
    pre = np.ascontiguousarray(pre, np.int32)
    sel = np.zeros(len(net), np.int32)
    lo, up = np.zeros(1), np.zeros(1)
    c = lib().orc_select_sds(net, pre, sel, len(net), int(pre.sum()), float(n_sds), lo, up)
    return sel, int(c), float(lo[0]), float(up[0])


def ls_multi_strategy_tables(shifts, var_shift, var_div, lower, upper, weights, vars_per_date,
                             cross):
    """The reference's packed LS multi-regression strategy arrays
    (LSAMultiExerciseStrategyQuadratic(Cross)_gpu::outputDataVectors,
    cashFlowGen_earlyExProduct_gpu.h:775-830): float data = Andersen shifts
    [E], variable shifts [E][depth][maxv], divisors [E][depth][maxv], lower
    cuts [E][depth], upper cuts [E][depth], weights [E][depth][row]; int data
    = depth, maxv, row, variables per date [E]."""
    E, depth, maxv = np.shape(var_shift)
    v = int(max(vars_per_date))
    assert maxv == v, "the variable stride of the packed tables is the largest variable count"
    row = (3 * v + 1 if v > 2 else (6 if v == 2 else 3)) if cross else 2 * v + 1
    fd = np.concatenate([np.ravel(shifts), np.ravel(var_shift), np.ravel(var_div),
                         np.ravel(lower), np.ravel(upper), np.ravel(weights)]).astype(np.float64)
    assert np.shape(weights) == (E, depth, row)
    iv = np.concatenate([[depth, maxv, row], vars_per_date]).astype(np.int32)
    return fd, iv


def cashflows_ee(cross, aux, fd, iv, is_ex, r1, sv):
    """Cancellable swap second pass; r1 (S,P), sv (E, maxSize, P); returns
    flows1, flows2 (S,P), continuation estimates (S,P) (exercise dates only)
    and the steps processed per path."""
    r1 = np.ascontiguousarray(r1, np.float64)
    S, P = r1.shape
    sv = np.ascontiguousarray(sv, np.float64)
    E, smax, _ = sv.shape
    f1, f2, est = np.zeros((S, P)), np.zeros((S, P)), np.zeros((S, P))
    idx = np.zeros(P, np.int32)
    ex = np.ascontiguousarray(is_ex, np.uint8)
    lib().orc_cashflows_ee(int(cross), f1.reshape(-1), f2.reshape(-1),
                           np.ascontiguousarray(aux, np.float64), np.ascontiguousarray(fd),
                           np.ascontiguousarray(iv, np.int32), P, S, E, ex.ctypes.data,
                           r1.reshape(-1), sv.reshape(-1), smax, est.reshape(-1), idx)
    return f1, f2, est, idx


def bridge_pow2(log2, z):
    """Legacy power-of-two bridge of z [2^log2][n] (differenced)."""
    z = np.ascontiguousarray(z, np.float64)
    out = np.zeros_like(z)
    lib().orc_bridge_pow2(int(log2), z.shape[1], z.reshape(-1), out.reshape(-1))
    return out


def bridge_reorder(z, log2, factors, ordering):
    """[D][P] -> (S, P, F) arrangement of the multi-factor power-of-two bridge."""
    z = np.ascontiguousarray(z, np.float64)
    P = z.shape[1]
    out = np.zeros(((1 << log2), P, factors))
    lib().orc_bridge_reorder(z.reshape(-1), out.reshape(-1), int(log2), int(factors), P,
                             int(ordering))
    return out


def lmm_steps(f0, taus, incr, A, disp, alive, euler=False, initial_drifts=None):
    """The multi-kernel PC (or Euler) evolver on correlated increments
    incr (S,N,P) = A z + fixed drift; returns rates and log-rates (S,N,P)
    (entries of dead rates, r < alive[s] for s >= 1, stay zero)."""
    f0 = np.ascontiguousarray(f0, np.float64)
    taus = np.ascontiguousarray(taus, np.float64)
    disp = np.ascontiguousarray(disp, np.float64)
    A = np.ascontiguousarray(A, np.float64)
    S, N, F = A.shape
    incr = np.ascontiguousarray(incr, np.float64)
    P = incr.shape[2]
    alive = np.ascontiguousarray(alive, np.int32)
    logs0 = np.log(f0 + disp)
    mu0 = spot_drifts(A[0], f0, taus, disp, 0) if initial_drifts is None else \
        np.ascontiguousarray(initial_drifts, np.float64)
    rates = np.zeros(S * N * P)
    logs = np.zeros(S * N * P)
    lib().orc_lmm_steps(logs0, taus, incr.reshape(-1), A.reshape(-1), mu0, disp, alive, P, F, S, N,
                        int(euler), rates, logs)
    return rates.reshape(S, N, P), logs.reshape(S, N, P)


def caplet_model(f0, taus, disp, A, sigma, strike, paths):
    arrs = [np.ascontiguousarray(x, np.float64) for x in (f0, taus, disp, A, sigma)]
    m = CapletModel(len(arrs[0]), arrs[3].shape[2], paths, arrs[0].ctypes.data,
                    arrs[1].ctypes.data, arrs[2].ctypes.data, arrs[3].ctypes.data,
                    arrs[4].ctypes.data, float(strike))
    m._keep = arrs
    return m


def caplet_price_paths(model, z):
    z = np.ascontiguousarray(z, np.float64)
    out = np.zeros(model.paths)
    lib().orc_caplet_price_paths(ctypes.byref(model), z.reshape(-1), out)
    return out


def caplet_adjoint_paths(model, z):
    z = np.ascontiguousarray(z, np.float64)
    pp = np.zeros(model.paths)
    d = np.zeros(model.n_rates)
    v = np.zeros(model.n_rates)
    lib().orc_caplet_adjoint_paths(ctypes.byref(model), z.reshape(-1), pp, d, v)
    return pp, d, v


def set_threads(n):
    """OpenMP threads of the oracle's parallel loops."""
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


def caplet_pipeline(model, ints, adjoint=False):
    ints = np.ascontiguousarray(ints, np.uint32)
    price = ctypes.c_double(0.0)
    d = np.zeros(model.n_rates)
    v = np.zeros(model.n_rates)
    lib().orc_caplet_pipeline(ctypes.byref(model), ints.reshape(-1), int(adjoint),
                              ctypes.byref(price), d, v)
    return price.value, d, v


# ---------------------------------------------------------------- stage (d)

def basis_count(n_vars, use_cross=True):
    return lib().orc_basis_count(int(n_vars), int(use_cross))


def basis_functions(x, use_cross=True):
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros(basis_count(len(x), use_cross))
    lib().orc_basis_functions(x, len(x), int(use_cross), out)
    return out


def cholesky_solve(A, b):
    A = np.ascontiguousarray(A, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(len(b))
    st = lib().orc_cholesky_solve(len(b), A.reshape(-1), b, x)
    return x if st == 0 else None


def sym_solve_cut(A, b, tol):
    """Cholesky when a trace bound shows no eigenvalue below tol times the
    largest, else the truncated eigen-solve (the LS passes' solve)."""
    A = np.ascontiguousarray(A, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(len(b))
    lib().orc_sym_solve_cut(len(b), A.reshape(-1), b, float(tol), x)
    return x


def sym_solve_truncated(A, b, tol):
    """Symmetric solve with eigenvalues below tol * the largest dropped."""
    A = np.ascontiguousarray(A, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(len(b))
    lib().orc_sym_solve_truncated(len(b), A.reshape(-1), b, float(tol), x)
    return x


def _ls_args(bv, vars_each_step, cash, ev, num):
    bv = np.ascontiguousarray(bv, np.float64)
    E, V, P = bv.shape
    ves = np.ascontiguousarray(vars_each_step, np.int32)
    arrs = [np.ascontiguousarray(a, np.float64).reshape(E, P) for a in (cash, ev, num)]
    return bv, E, V, P, ves, arrs


def ls_first_pass(bv, vars_each_step, cash, ev, num, use_cross=True, itm_only=False):
    """Backward induction; returns (price, coefficients (E, maxData))."""
    bv, E, V, P, ves, (c, e, n) = _ls_args(bv, vars_each_step, cash, ev, num)
    coeffs = np.zeros((E, basis_count(V, use_cross)))
    price = lib().orc_ls_first_pass(E, bv.reshape(-1), ves, V, c.reshape(-1), e.reshape(-1),
                                    n.reshape(-1), P, int(use_cross), int(itm_only),
                                    coeffs.reshape(-1))
    return price, coeffs


def ls_second_pass(bv, vars_each_step, cash, ev, num, coeffs, use_cross=True, itm_only=False):
    bv, E, V, P, ves, (c, e, n) = _ls_args(bv, vars_each_step, cash, ev, num)
    coeffs = np.ascontiguousarray(coeffs, np.float64)
    return lib().orc_ls_second_pass(E, bv.reshape(-1), ves, V, c.reshape(-1), e.reshape(-1),
                                    n.reshape(-1), P, int(use_cross), int(itm_only),
                                    coeffs.reshape(-1))


def ref_ls_first_pass(bv, vars_each_step, cash, ev, num, use_cross=True, normalise=True):
    """The reference's gold LS (one regression per date).  The coefficients
    returned with ``normalise`` apply to the gold's standardised variables."""
    bv, E, V, P, ves, (c, e, n) = _ls_args(bv, vars_each_step, cash, ev, num)
    coeffs = np.zeros((E, basis_count(V, use_cross)))
    price = ref().ref_ls_multi_depth0(E, bv.reshape(-1), ves, V, c.reshape(-1), e.reshape(-1),
                                      n.reshape(-1), P, int(use_cross), int(normalise),
                                      coeffs.reshape(-1))
    return price, coeffs


def bermudan_paths(model, z, ex_steps):
    """Per exercise date: basis variables (E,2,P) = (co-terminal swap rate,
    f_{t_e}), exercise values (E,P) and spot numeraires (E,P)."""
    z = np.ascontiguousarray(z, np.float64)
    ex = np.ascontiguousarray(ex_steps, np.int32)
    E, P = len(ex), model.paths
    bv = np.zeros((E, 2, P))
    ev = np.zeros((E, P))
    num = np.zeros((E, P))
    lib().orc_bermudan_paths(ctypes.byref(model), z.reshape(-1), E, ex, bv.reshape(-1),
                             ev.reshape(-1), num.reshape(-1))
    return bv, ev, num


def normals_from_ints(ints, steps, factors):
    """Inverse normal (FP64) then the diagonal Brownian bridge: (S,F,P)."""
    ints = np.ascontiguousarray(ints, np.uint32)
    return bridge_transform(invnorm_u32(ints), steps, factors)


# ---------------------------------------------------------------- multi-d Black-Scholes

def cholesky_lower(C):
    C = np.ascontiguousarray(C, np.float64)
    L = np.zeros_like(C)
    st = lib().orc_cholesky_lower(C.shape[0], C.reshape(-1), L.reshape(-1))
    return L if st == 0 else None


def correlated_paths(z_spf, A, alive=None):
    """z (S,P,F), A (N,F) for every step -> (S,N,P)."""
    z = np.ascontiguousarray(z_spf, np.float64)
    S, P, F = z.shape
    A = np.ascontiguousarray(A, np.float64)
    N = A.shape[0]
    al = np.zeros(S, np.int32) if alive is None else np.ascontiguousarray(alive, np.int32)
    out = np.zeros((S, N, P))
    lib().orc_correlated_paths(z.reshape(-1), out.reshape(-1), A.reshape(-1), al, 0, F, N, P, S)
    return out


def multid_bs_paths(log0, incr, drifts=None):
    incr = np.ascontiguousarray(incr, np.float64)
    S, N, P = incr.shape
    d = None if drifts is None else np.ascontiguousarray(drifts, np.float64)
    out = np.zeros_like(incr)
    lib().orc_multid_bs_paths(np.ascontiguousarray(log0, np.float64), incr.reshape(-1),
                              None if d is None else d.ctypes.data, out.reshape(-1), P, N, S)
    return out


def basket_weightings(paths_snp, w):
    x = np.ascontiguousarray(paths_snp, np.float64)
    S, N, P = x.shape
    out = np.zeros((S, P))
    lib().orc_basket_weightings(x.reshape(-1), out.reshape(-1), np.ascontiguousarray(w, np.float64),
                                P, N, S)
    return out


def basket_model(x0, vols, chol, weights, rate, dt, strike, steps, paths):
    arrs = [np.ascontiguousarray(a, np.float64) for a in (x0, vols, chol, weights)]
    m = BasketModel(len(arrs[0]), steps, paths, arrs[0].ctypes.data, arrs[1].ctypes.data,
                    arrs[2].ctypes.data, arrs[3].ctypes.data, rate, dt, strike)
    m._keep = arrs
    return m


def basket_asian_paths(model, z):
    """z (S,A,P) -> (price per path, delta sums (A), vega sums (A))."""
    z = np.ascontiguousarray(z, np.float64)
    pp = np.zeros(model.paths)
    d = np.zeros(model.assets)
    v = np.zeros(model.assets)
    lib().orc_basket_asian_paths(ctypes.byref(model), z.reshape(-1), pp, d, v)
    return pp, d, v


# ---------------------------------------------------------------- frozen Bermudan

def frozen_strategy(ex_steps, coeffs, shift, scale, itm_only=True):
    arrs = [np.ascontiguousarray(ex_steps, np.int32)] + \
        [np.ascontiguousarray(a, np.float64) for a in (coeffs, shift, scale)]
    fs = FrozenStrategy(len(arrs[0]), arrs[0].ctypes.data, arrs[1].ctypes.data,
                        arrs[2].ctypes.data, arrs[3].ctypes.data, int(itm_only))
    fs._keep = arrs
    return fs


def bermudan_frozen_paths(model, strategy, z, adjoint=True):
    """-> (value per path, exercise step per path (-1: never), delta sums, vega sums)."""
    z = np.ascontiguousarray(z, np.float64)
    vals = np.zeros(model.paths)
    ex = np.zeros(model.paths, np.int32)
    d = np.zeros(model.n_rates)
    v = np.zeros(model.n_rates)
    lib().orc_bermudan_frozen_paths(ctypes.byref(model), ctypes.byref(strategy), z.reshape(-1),
                                    int(adjoint), vals, ex.ctypes.data, d, v)
    return vals, ex, d, v


def bermudan_given_paths(model, strategy, z, given, adjoint=True):
    """Frozen-strategy value and adjoint with each path's exercise step given
    (-1: never) -> (value per path, delta sums, vega sums)."""
    z = np.ascontiguousarray(z, np.float64)
    given = np.ascontiguousarray(given, np.int32)
    assert given.shape == (model.paths,)
    vals = np.zeros(model.paths)
    d = np.zeros(model.n_rates)
    v = np.zeros(model.n_rates)
    lib().orc_bermudan_given_paths(ctypes.byref(model), ctypes.byref(strategy), z.reshape(-1),
                                   given, int(adjoint), vals, d, v)
    return vals, d, v


def ref_swap_pipeline(paths, f0, taus, A, fixed_drifts, strike, scramblers, power_of_two,
                      disp=None):
    """The reference's own CPU pipeline for the plain-LMM payer swap
    (LMMEvolutionRoutineGold + gold curve, flow and discounting routines);
    returns its price.  scramblers: 2^power_of_two * factors words."""
    A = np.ascontiguousarray(A, np.float64)
    S, N, F = A.shape
    d = np.zeros(N) if disp is None else np.ascontiguousarray(disp, np.float64)
    return ref().ref_swap_pipeline(paths, N, F, power_of_two,
                                   np.ascontiguousarray(scramblers, np.uint32), A.reshape(-1),
                                   np.ascontiguousarray(fixed_drifts, np.float64).reshape(-1), d,
                                   np.ascontiguousarray(f0, np.float64),
                                   np.ascontiguousarray(taus, np.float64), float(strike))
