"""ctypes binding of libkoo_amd.so and torch-tensor wrappers.

Every wrapper keeps the reference launcher's name and argument order
(include/koo_c_abi.h cites the reference header of each), takes torch tensors
on the GPU for device buffers, launches on torch's current stream and raises
``KooError`` (with the library's message) on a non-zero status.  Nothing here
computes on the CPU: if the library is missing, ``lib()`` raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkoo_amd.so")
_LIB = None

_p = ctypes.c_void_p
_i = ctypes.c_int
_u = ctypes.c_uint
_ll = ctypes.c_longlong
_d = ctypes.c_double
_sz = ctypes.c_size_t
_f = ctypes.c_float


class KooError(RuntimeError):
    pass


class BermProduct(ctypes.Structure):
    """koo_berm_product (include/koo_c_abi.h)."""
    _fields_ = [("paths", _i), ("rates", _i), ("factors", _i), ("checkpoint_every", _i),
                ("variates", _p), ("initial_rates", _p), ("taus", _p), ("displacements", _p),
                ("pseudo_roots", _p), ("vols", _p), ("strike", _f), ("n_ex", _i),
                ("exercise_steps", _p), ("coefficients", _p), ("var_shift", _p),
                ("var_scale", _p), ("itm_only", _i), ("checkpoints", _p), ("path_buffers", _p),
                ("segment_scratch", _p), ("partials", _p), ("price_out", _p), ("greeks_out", _p)]


class PortfolioPlan(ctypes.Structure):
    """koo_portfolio_plan (include/koo_c_abi.h)."""
    _fields_ = [("products", _i), ("blocks", _i), ("max_rates", _i), ("ls_blocks", _i),
                ("max_ex", _i), ("classes", _i), ("class_first_block", _i * 4),
                ("class_blocks", _i * 4), ("class_max_rates", _i * 4)]


class LsProduct(ctypes.Structure):
    """koo_ls_product (include/koo_c_abi.h)."""
    _fields_ = [("paths", _i), ("rates", _i), ("factors", _i), ("variates", _p),
                ("initial_rates", _p), ("taus", _p), ("displacements", _p), ("pseudo_roots", _p),
                ("strike", _f), ("n_ex", _i), ("exercise_steps", _p), ("var_shift", _p),
                ("var_scale", _p), ("itm_only", _i), ("basis_variables", _p),
                ("exercise_values", _p), ("numeraires", _p), ("workspace", _p),
                ("coefficients", _p), ("price_out", _p), ("ls_blocks", _i)]


class GenProduct(ctypes.Structure):
    """koo_gen_product (include/koo_c_abi.h)."""
    _fields_ = [("path_offset", _u), ("number_scrambled", _i), ("directions", _p),
                ("scramblers", _p), ("allocation_indices", _p)]


class PersistentPlan(ctypes.Structure):
    """koo_persistent_plan (include/koo_c_abi.h)."""
    _fields_ = [("products", _i), ("grid", _i), ("lds_bytes", _i), ("tab_offset", _i),
                ("blocks_per_cu", _i), ("err_offset", _sz), ("ticks_offset", _sz),
                ("team_first", _i * 512), ("team_blocks", _i * 512)]


# name -> (restype, argtypes); pointers are passed as integers (c_void_p).
_SIGNATURES = {
    "koo_last_error": (ctypes.c_char_p, []),
    "koo_version": (ctypes.c_char_p, []),
    # stage (a)
    "initSobolDirectionVectors": (_i, [_i, _p]),
    "sobolGPUInts": (_i, [_i, _i, _u, _p, _p, _p]),
    "scramble_gpu": (_i, [_p, _p, _i, _i, _p]),
    "inverse_normal_u32_gpu": (_i, [_p, _p, _ll, _p]),
    "sobol_scramble_normal_gpu": (_i, [_i, _i, _u, _p, _p, _i, _p, _p]),
    # host set-up
    "koo_exp_corr": (_i, [_p, _i, _d, _d, _p]),
    "koo_pseudo_sqrt": (_i, [_i, _p, _i, _i, _p]),
    "FlatVolPseudoRootsOfCovariances": (_i, [_p, _p, _p, _i, _i, _i, _d, _d, _p]),
    "koo_factor_pseudo_roots": (_i, [_p, _p, _i, _i, _d, _d, _p, _p]),
    "koo_fixed_drifts": (_i, [_p, _i, _i, _i, _p]),
    "spotDriftComputer_getDrifts": (_i, [_p, _p, _p, _p, _i, _i, _i, _p]),
    # stage (b)
    "bridge_allocate_diagonal": (_i, [_i, _i, _p]),
    "brownian_bridge_gpu": (_i, [_i, _p, _p, _p, _i, _i, _p]),
    "sobol_bridge_gpu": (_i, [_i, _u, _i, _i, _p, _p, _i, _p, _p, _p]),
    "brownianBridgeGPU_constant_memory": (_i, [_i, _i, _p, _p, _p]),
    "brownianBridgeMultiGPUReorder": (_i, [_p, _p, _i, _i, _i, _i, _i, _p]),
    "correlated_drift_paths_GPU": (_i, [_p, _p, _p, _p, _p, _i, _i, _i, _i, _i, _p]),
    "LMM_evolver_pc_all_gpu": (_i, [_p] * 9 + [_i] * 4 + [_p] * 5 + [_i, _i, _p, _i, _i]),
    "LMM_evolver_pc_gpu": (_i, [_p] * 8 + [_i] * 4 + [_p] * 4 + [_i, _p]),
    "LMM_evolver_euler_gpu": (_i, [_p] * 8 + [_i] * 4 + [_p] * 4),
    "discount_ratios_computation_gpu": (_i, [_p, _p, _p, _p, _i, _i, _i, _p, _i, _p]),
    "coterminal_annuity_ratios_computation_gpu": (_i, [_p, _p, _p, _i, _i, _i, _p, _p]),
    "coterminal_swap_rate_computation_gpu": (_i, [_p, _p, _p, _i, _i, _i, _p, _p]),
    "swap_rate_computation_gpu": (_i, [_p, _p, _i, _i, _i, _i, _i, _i, _p, _p]),
    "spot_measure_numeraires_computation_gpu": (_i, [_p, _p, _i, _i, _p]),
    "spot_measure_numeraires_computation_offset_gpu": (_i, [_p, _p, _i, _i, _i, _i, _p]),
    "forward_rate_extraction_gpu": (_i, [_p, _p, _i, _i, _i, _p, _p]),
    "forward_rate_extraction_selecting_step_gpu": (_i, [_p, _p, _p, _i, _i, _i, _p, _p]),
    "spot_measure_numeraires_extraction_gpu": (_i, [_p, _p, _i, _i, _i, _i, _i, _p, _p]),
    # stage (c)
    "cashFlowGeneratorCallerSwap": (_i, [_p] * 3 + [_i, _i] + [_p] * 5 + [_p]),
    "cashFlowGeneratorCallerTARN": (_i, [_p] * 3 + [_i, _i] + [_p] * 5 + [_p]),
    "cashFlowGeneratorCallerCaplet": (_i, [_p] * 3 + [_i, _i] + [_p] * 5 + [_p]),
    "cashFlowDiscounting_gpu": (_i, [_p] * 6 + [_i, _i, _i, _i, _p, _p, _p]),
    "cashFlowDiscounting_partial_gpu": (_i, [_i] + [_p] * 8 + [_i, _i, _p, _p]),
    "AggregateFlows_gpu": (_i, [_p, _i, _i, _p, _i, _i, _i, _p, _i, _p]),
    "PartialReduce_using8s_gpu": (_i, [_p, _p, _i, _i, _i, _p, _p, _p]),
    "PartialReduce_using16s_gpu": (_i, [_p, _p, _i, _i, _i, _p, _p, _p]),
    "reduceSinglePass": (_i, [_i, _i, _i, _p, _p, _p, _p]),
    "koo_sum_workspace_doubles": (_ll, [_ll]),
    "koo_sum_f32_to_f64": (_i, [_p, _ll, _p, _p, _p]),
    "koo_sum_f64": (_i, [_p, _ll, _p, _p, _p]),
    # fused pricing + adjoint
    "lmm_price_workspace": (_i, [_i, _i, _i, _i, _p, _p, _p]),
    "lmm_price_fwd": (_i, [_i, _i, _i, _i] + [_p] * 5 + [_i] + [_p] * 5),
    "lmm_price_bwd": (_i, [_i, _i, _i, _i] + [_p] * 6 + [_i] + [_p] * 6),
    # stage (d)
    "adjoinBasisVariablesCaller_gpu": (_i, [_i] + [_p] * 7 + [_i] * 5 + [_p] * 4),
    "adjoinExerciseValues_Bermudan_swaption_gpu": (_i, [_p] * 5 + [_i] * 5 + [_p] * 6),
    "expandBasisFunctions_quadratic_gpu": (_i, [_i, _p, _i, _i, _i, _i, _p, _p]),
    "continuation_net_value_evaluation_gpu": (_i, [_i, _i, _p, _i, _p, _p, _f, _p, _p]),
    "oneStepUpdateLS_multi_gpu": (_i, [_i, _i, _p, _i, _i, _i] + [_p] * 6 + [_f, _p, _p, _p]),
    "updateDeflationOfContinuationValues_gpu": (_i, [_p, _p, _i, _i, _i, _p]),
    "reduceOuterProduct_gpu": (_i, [_i, _i, _i, _i, _i, _i] + [_p] * 5),
    "reducedOuterProduct_gpu": (_i, [_i, _i, _i, _i, _i] + [_p] * 5),
    "reduceOuterProductSymmetric_gpu": (_i, [_i, _i, _i, _i] + [_p] * 4),
    "outer_product_workspace_floats": (_sz, [_i]),
    "reduceOuterProductSymmetric_cublas": (_i, [_i, _i, _f, _f] + [_p] * 4),
    "reduceOuterProduct_cublas": (_i, [_i, _i, _i, _f, _f] + [_p] * 5),
    "PointwiseProduct_gpu": (_i, [_i, _i, _i, _p, _p, _p, _p]),
    "PointwiseProductSymmetricUpperTriangular_gpu": (_i, [_i, _i, _i, _i, _p, _p, _p]),
    "PointwiseProductMultipleFirst_gpu": (_i, [_i, _i, _i, _p, _p, _p, _i, _p]),
    "bermudan_paths_gpu": (_i, [_i, _i, _i] + [_p] * 5 + [_i, _p, _f] + [_p] * 4),
    "cashFlowGeneratorEE_cancellable_swap_gpu": (_i, [_i] + [_p] * 6 + [_i, _p, _i, _i, _i, _i]
                                                 + [_p] * 5 + [_i] + [_p] * 5),
    "bermudan_portfolio_descriptor_bytes": (_sz, [_i]),
    "bermudan_ls_portfolio_descriptor_bytes": (_sz, [_i]),
    "bermudan_ls_portfolio_prepare": (_i, [_i, _p, _p, _p]),
    "bermudan_ls_portfolio_run": (_i, [_p, _p, _p]),
    "bermudan_portfolio_prepare": (_i, [_i, _p, _p, _p]),
    "bermudan_portfolio_fwd": (_i, [_p, _p, _p]),
    "bermudan_portfolio_bwd": (_i, [_p, _p, _p]),
    "bermudan_portfolio_persistent_bytes": (_sz, [_i, _p]),
    "bermudan_portfolio_persistent_prepare": (_i, [_i, _p, _p, _p, _p, _p, _p]),
    "bermudan_portfolio_persistent_run": (_i, [_p, _p, _p]),
    "ls_workspace": (_i, [_i, _p]),
    "ls_first_pass_gpu": (_i, [_i, _p, _p, _i, _p, _p, _p, _i, _i, _i] + [_p] * 6),
    "ls_first_pass_noalloc_gpu": (_i, [_i, _p, _p, _i, _p, _p, _p, _i, _i, _i] + [_p] * 5),
    "ls_multi_workspace": (_i, [_i, _p]),
    "ls_multi_first_pass_gpu": (_i, [_i, _p, _p, _i, _p, _p, _p, _i, _i, _i, _i, _i, _i, _d, _d,
                                     _d, _d] + [_p] * 8),
    "ls_second_pass_gpu": (_i, [_i, _p, _p, _i, _p, _p, _p, _i, _i, _i] + [_p] * 6),
    "ls_normal_eq_gpu": (_i, [_i, _i] + [_p] * 6),
    "cholesky_solve_f64_gpu": (_i, [_i] + [_p] * 5),
    # multi-asset Black-Scholes
    "koo_cholesky_lower": (_i, [_i, _p, _p]),
    "correlated_paths_multistep_GPU": (_i, [_p] * 4 + [_i] * 5 + [_p]),
    "multi_dim_BS_path_generator_kernel": (_i, [_p] * 3 + [_i] * 3 + [_p]),
    "basketWeightings_gpu": (_i, [_p] * 3 + [_i] * 3 + [_p]),
    "basket_asian_workspace": (_i, [_i, _i, _p]),
    "basket_asian_gpu": (_i, [_i, _i, _i] + [_p] * 5 + [_f, _f, _f, _p, _p, _p]),
    # frozen-strategy Bermudan
    "bermudan_adjoint_workspace": (_i, [_i, _i, _i, _i, _p, _p, _p, _p]),
    "bermudan_frozen_fwd": (_i, [_i, _i, _i, _i] + [_p] * 5 + [_f, _i] + [_p] * 4 + [_i] +
                            [_p] * 5),
    "bermudan_frozen_bwd": (_i, [_i, _i, _i, _i] + [_p] * 6 + [_f] + [_p] * 6),
}


def lib():
    """The loaded libkoo_amd.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise KooError(f"{LIB_PATH} is missing: build it with `python -m useful_code_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def exported_symbols():
    return list(_SIGNATURES)


def check(status, name):
    if status != 0:
        msg = lib().koo_last_error().decode(errors="replace")
        raise KooError(f"{name} failed ({status}): {msg}")


def call(name, *args):
    check(getattr(lib(), name)(*args), name)


def ptr(t):
    """Device (or host numpy) address of a tensor/array; None -> NULL."""
    if t is None:
        return None
    if isinstance(t, np.ndarray):
        return t.ctypes.data
    return t.data_ptr()


def stream_handle(stream=None):
    import torch
    s = torch.cuda.current_stream() if stream is None else stream
    return s.cuda_stream


# ------------------------------------------------------------ host set-up

def _f64(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def sobol_directions(n_dims):
    out = np.zeros(n_dims * 32, np.uint32)
    call("initSobolDirectionVectors", n_dims, out.ctypes.data)
    return out


def bridge_allocate_diagonal(factors, steps):
    out = np.zeros(factors * steps, np.int32)
    call("bridge_allocate_diagonal", factors, steps, out.ctypes.data)
    return out


def exp_corr(times, L, beta):
    t = _f64(times)
    out = np.zeros((t.size, t.size))
    call("koo_exp_corr", t.ctypes.data, t.size, L, beta, out.ctypes.data)
    return out


def pseudo_sqrt(cov, factors, normalise=True):
    c = _f64(cov)
    out = np.zeros((c.shape[0], factors))
    call("koo_pseudo_sqrt", c.shape[0], c.ctypes.data, factors, int(normalise), out.ctypes.data)
    return out


def flat_vol_pseudo_roots(rate_times, evol_times, vols, factors, L, beta):
    rt, et, v = _f64(rate_times), _f64(evol_times), _f64(vols)
    A = np.zeros((et.size, v.size, factors))
    call("FlatVolPseudoRootsOfCovariances", rt.ctypes.data, et.ctypes.data, v.ctypes.data, v.size,
         et.size, factors, L, beta, A.ctypes.data)
    return A


def factor_pseudo_roots(rate_times, vols, factors, L, beta):
    rt, v = _f64(rate_times), _f64(vols)
    A = np.zeros((v.size, v.size, factors))
    B = np.zeros((v.size, factors))
    call("koo_factor_pseudo_roots", rt.ctypes.data, v.ctypes.data, v.size, factors, L, beta,
         A.ctypes.data, B.ctypes.data)
    return A, B


def fixed_drifts(A):
    A = _f64(A)
    S, N, F = A.shape
    out = np.zeros((S, N))
    call("koo_fixed_drifts", A.ctypes.data, S, N, F, out.ctypes.data)
    return out


def spot_drifts(A_step, rates, taus, disp, alive=0):
    A_step = _f64(A_step)
    N, F = A_step.shape
    r, t, d = _f64(rates), _f64(taus), _f64(disp)
    out = np.zeros(N)
    call("spotDriftComputer_getDrifts", A_step.ctypes.data, r.ctypes.data, t.ctypes.data,
         d.ctypes.data, F, N, alive, out.ctypes.data)
    return out


# ------------------------------------------------------------ device ops

def sobolGPUInts(n_vectors, n_dimensions, path_offset, d_directions, d_output, stream=None):
    call("sobolGPUInts", n_vectors, n_dimensions, path_offset, ptr(d_directions), ptr(d_output),
         stream_handle(stream))


def scramble_gpu(device_data, scramblers, number_dim, paths, stream=None):
    call("scramble_gpu", ptr(device_data), ptr(scramblers), number_dim, paths,
         stream_handle(stream))


def inverse_normal_u32_gpu(x, out, stream=None):
    call("inverse_normal_u32_gpu", ptr(x), ptr(out), x.numel(), stream_handle(stream))


def sobol_scramble_normal_gpu(n_vectors, n_dimensions, path_offset, d_directions, scramblers,
                              number_scrambled, out, stream=None):
    call("sobol_scramble_normal_gpu", n_vectors, n_dimensions, path_offset, ptr(d_directions),
         ptr(scramblers), number_scrambled, ptr(out), stream_handle(stream))


def brownian_bridge_gpu(steps, allocation_indices, d_in, d_out, paths, factors, stream=None):
    call("brownian_bridge_gpu", steps, ptr(allocation_indices), ptr(d_in), ptr(d_out), paths,
         factors, stream_handle(stream))


def sobol_bridge_gpu(paths, path_offset, steps, factors, d_directions, scramblers,
                     number_scrambled, allocation_indices, out, stream=None):
    call("sobol_bridge_gpu", paths, path_offset, steps, factors, ptr(d_directions),
         ptr(scramblers), number_scrambled, ptr(allocation_indices), ptr(out),
         stream_handle(stream))


def brownianBridgeGPU_constant_memory(n_vectors, n_poweroftwo, inp, out, stream=None):
    call("brownianBridgeGPU_constant_memory", n_vectors, n_poweroftwo, ptr(inp), ptr(out),
         stream_handle(stream))


def brownianBridgeMultiGPUReorder(inp, out, n_poweroftwo, factors, number_paths, allocator,
                                  stream=None):
    call("brownianBridgeMultiGPUReorder", ptr(inp), ptr(out), n_poweroftwo, factors,
         number_paths, int(allocator), 0, stream_handle(stream))


def correlated_drift_paths_GPU(d_in, d_out, A, alive, drift_fixed, A_offset_per_step, factors,
                               out_dimensions, number_paths, steps, stream=None):
    call("correlated_drift_paths_GPU", ptr(d_in), ptr(d_out), ptr(A), ptr(alive),
         ptr(drift_fixed), A_offset_per_step, factors, out_dimensions, number_paths, steps,
         stream_handle(stream))


def LMM_evolver_pc_all_gpu(initial_rates, initial_log_rates, taus, variates, pseudo_roots,
                           initial_drifts, fixed_drifts_dev, displacements, alive, paths,
                           factors, steps, rates, evolved_rates, evolved_log_rates, discounts,
                           transposeData=True, doDiscounts=True, stream=None):
    call("LMM_evolver_pc_all_gpu", ptr(initial_rates), ptr(initial_log_rates), ptr(taus),
         ptr(variates), ptr(pseudo_roots), ptr(initial_drifts), ptr(fixed_drifts_dev),
         ptr(displacements), ptr(alive), paths, factors, steps, rates, None, None,
         ptr(evolved_rates), ptr(evolved_log_rates), ptr(discounts), 0, int(transposeData),
         stream_handle(stream), 0, int(doDiscounts))


def _host_ints(v):
    return np.ascontiguousarray(v, dtype=np.int32)


def LMM_evolver_pc_gpu(initial_rates, initial_log_rates, taus, increments, pseudo_roots,
                       initial_drifts, displacements, alive_host, paths, factors, steps, rates,
                       evolved_rates, evolved_log_rates, stream=None):
    a = _host_ints(alive_host)
    call("LMM_evolver_pc_gpu", ptr(initial_rates), ptr(initial_log_rates), ptr(taus),
         ptr(increments), ptr(pseudo_roots), ptr(initial_drifts), ptr(displacements),
         a.ctypes.data, paths, factors, steps, rates, None, None, ptr(evolved_rates),
         ptr(evolved_log_rates), 0, stream_handle(stream))


def LMM_evolver_euler_gpu(initial_rates, initial_log_rates, taus, increments, pseudo_roots,
                          initial_drifts, displacements, alive_host, paths, factors, steps, rates,
                          evolved_rates, evolved_log_rates, stream=None):
    a = _host_ints(alive_host)
    call("LMM_evolver_euler_gpu", ptr(initial_rates), ptr(initial_log_rates), ptr(taus),
         ptr(increments), ptr(pseudo_roots), ptr(initial_drifts), ptr(displacements),
         a.ctypes.data, paths, factors, steps, rates, None, ptr(evolved_rates),
         ptr(evolved_log_rates), stream_handle(stream))


def discount_ratios_computation_gpu(rates_dev, taus, alive_host, paths, steps, n_rates, disc,
                                    stream=None):
    a = _host_ints(alive_host)
    call("discount_ratios_computation_gpu", ptr(rates_dev), ptr(taus), a.ctypes.data, None, paths,
         steps, n_rates, ptr(disc), 1, stream_handle(stream))


def coterminal_annuity_ratios_computation_gpu(disc, taus, alive_host, paths, steps, n_rates, ann,
                                              stream=None):
    a = _host_ints(alive_host)
    call("coterminal_annuity_ratios_computation_gpu", ptr(disc), ptr(taus), a.ctypes.data, paths,
         steps, n_rates, ptr(ann), stream_handle(stream))


def coterminal_swap_rate_computation_gpu(disc, ann, alive_host, paths, steps, n_rates, sr,
                                         stream=None):
    a = _host_ints(alive_host)
    call("coterminal_swap_rate_computation_gpu", ptr(disc), ptr(ann), a.ctypes.data, paths, steps,
         n_rates, ptr(sr), stream_handle(stream))


def swap_rate_computation_gpu(disc, taus, start, end, paths, step_in, step_out, n_rates, out,
                              stream=None):
    call("swap_rate_computation_gpu", ptr(disc), ptr(taus), start, end, paths, step_in, step_out,
         n_rates, ptr(out), stream_handle(stream))


def spot_measure_numeraires_computation_gpu(disc, num, paths, n_rates, stream=None):
    call("spot_measure_numeraires_computation_gpu", ptr(disc), ptr(num), paths, n_rates,
         stream_handle(stream))


def spot_measure_numeraires_computation_offset_gpu(disc, num, paths, out_paths, offset, n_rates,
                                                   stream=None):
    call("spot_measure_numeraires_computation_offset_gpu", ptr(disc), ptr(num), paths, out_paths,
         offset, n_rates, stream_handle(stream))


def forward_rate_extraction_gpu(fwd, indices_host, paths, steps, n_rates, out, stream=None):
    ix = _host_ints(indices_host)
    call("forward_rate_extraction_gpu", ptr(fwd), ix.ctypes.data, paths, steps, n_rates, ptr(out),
         stream_handle(stream))


def forward_rate_extraction_selecting_step_gpu(fwd, indices_host, steps_host, paths, n_rates, out,
                                               stream=None):
    ix, st = _host_ints(indices_host), _host_ints(steps_host)
    call("forward_rate_extraction_selecting_step_gpu", ptr(fwd), ix.ctypes.data, st.ctypes.data,
         ix.size, paths, n_rates, ptr(out), stream_handle(stream))


def spot_measure_numeraires_extraction_gpu(all_num, some_num, paths_in, paths_out, offset,
                                           n_rates, n_ex, ex_indices_dev, stream=None):
    call("spot_measure_numeraires_extraction_gpu", ptr(all_num), ptr(some_num), paths_in,
         paths_out, offset, n_rates, n_ex, ptr(ex_indices_dev), stream_handle(stream))


def cash_flow_generator(product, f1, f2, aux, paths, steps, r1, r2=None, r3=None, stream=None):
    name = {"swap": "cashFlowGeneratorCallerSwap", "tarn": "cashFlowGeneratorCallerTARN",
            "caplet": "cashFlowGeneratorCallerCaplet"}[product]
    call(name, ptr(f1), ptr(f2), ptr(aux), paths, steps, ptr(r1), ptr(r2), ptr(r3), None, None,
         stream_handle(stream))


def cashFlowDiscounting_gpu(first, second, theta, disc, flows, num, paths, steps, out, summed,
                            stream=None):
    call("cashFlowDiscounting_gpu", ptr(first), ptr(second), ptr(theta), ptr(disc), ptr(flows),
         ptr(num), paths, steps, 0, 0, ptr(out), ptr(summed), stream_handle(stream))


def cashFlowDiscounting_partial_gpu(n_flows, gen, first, second, theta, target, disc, flows, num,
                                    paths, steps, out, stream=None):
    call("cashFlowDiscounting_partial_gpu", n_flows, ptr(gen), ptr(first), ptr(second), ptr(theta),
         ptr(target), ptr(disc), ptr(flows), ptr(num), paths, steps, ptr(out),
         stream_handle(stream))


def AggregateFlows_gpu(agg, total_paths, n_ex, flows, batch_paths, offset, steps, preceding,
                       start=0, stream=None):
    call("AggregateFlows_gpu", ptr(agg), total_paths, n_ex, ptr(flows), batch_paths, offset, steps,
         ptr(preceding), start, stream_handle(stream))


def partial_reduce(data, workspace, points, stop_point=1, width=8, stream=None):
    """Returns (device address of the remaining values, their count)."""
    res = ctypes.c_void_p()
    rem = ctypes.c_int()
    name = "PartialReduce_using8s_gpu" if width == 8 else "PartialReduce_using16s_gpu"
    call(name, ptr(data), ptr(workspace), points, 0, stop_point, ctypes.addressof(res),
         ctypes.addressof(rem), stream_handle(stream))
    return res.value, rem.value


def reduceSinglePass(size, d_in, wsp, final, stream=None):
    call("reduceSinglePass", size, 0, 0, ptr(d_in), ptr(wsp), ptr(final), stream_handle(stream))


def sum_workspace_doubles(n):
    return int(lib().koo_sum_workspace_doubles(n))


def sum_f32_to_f64(x, workspace, out, stream=None):
    call("koo_sum_f32_to_f64", ptr(x), x.numel(), ptr(workspace), ptr(out), stream_handle(stream))


def sum_f64(x, workspace, out, stream=None):
    call("koo_sum_f64", ptr(x), x.numel(), ptr(workspace), ptr(out), stream_handle(stream))


def lmm_price_workspace(paths, rates, factors, checkpoint_every):
    a, b, c = _sz(), _sz(), _sz()
    call("lmm_price_workspace", paths, rates, factors, checkpoint_every, ctypes.addressof(a),
         ctypes.addressof(b), ctypes.addressof(c))
    return a.value, b.value, c.value


KOO_PRODUCT_CAPLETS = 1


def lmm_price_fwd(paths, rates, factors, k, variates, f0, taus, disp, A, product, aux_host,
                  checkpoints, partials, price_out, stream=None):
    aux = np.ascontiguousarray(aux_host, np.float32)
    call("lmm_price_fwd", paths, rates, factors, k, ptr(variates), ptr(f0), ptr(taus), ptr(disp),
         ptr(A), product, aux.ctypes.data, ptr(checkpoints), ptr(partials), ptr(price_out),
         stream_handle(stream))


def lmm_price_bwd(paths, rates, factors, k, variates, f0, taus, disp, A, vols, product, aux_host,
                  checkpoints, scratch, partials, greeks_out, stream=None):
    aux = np.ascontiguousarray(aux_host, np.float32)
    call("lmm_price_bwd", paths, rates, factors, k, ptr(variates), ptr(f0), ptr(taus), ptr(disp),
         ptr(A), ptr(vols), product, aux.ctypes.data, ptr(checkpoints), ptr(scratch),
         ptr(partials), ptr(greeks_out), stream_handle(stream))


# ---------------------------------------------------------------- stage (d)

def _host_f32(x):
    return None if x is None else np.ascontiguousarray(x, np.float32)


def bermudan_paths_gpu(paths, rates, factors, variates, f0, taus, disp, A, exercise_steps, strike,
                       basis_variables, exercise_values, numeraires, stream=None):
    ex = _host_ints(exercise_steps)
    call("bermudan_paths_gpu", paths, rates, factors, ptr(variates), ptr(f0), ptr(taus),
         ptr(disp), ptr(A), len(ex), ex.ctypes.data, float(strike), ptr(basis_variables),
         ptr(exercise_values), ptr(numeraires), stream_handle(stream))


def ls_workspace(paths):
    n = _sz()
    call("ls_workspace", paths, ctypes.addressof(n))
    return n.value


def ls_first_pass_gpu(n_ex, basis_variables, vars_each_step, max_vars, cash, ev, num, paths,
                      use_cross, itm_only, var_shift, var_scale, workspace, coefficients,
                      price_out, stream=None):
    ves = _host_ints(vars_each_step)
    sh, sc = _host_f32(var_shift), _host_f32(var_scale)
    call("ls_first_pass_gpu", n_ex, ptr(basis_variables), ves.ctypes.data, max_vars, ptr(cash),
         ptr(ev), ptr(num), paths, int(use_cross), int(itm_only), ptr(sh), ptr(sc),
         ptr(workspace), ptr(coefficients), ptr(price_out), stream_handle(stream))


def ls_first_pass_noalloc_gpu(n_ex, basis_variables, vars_each_step, max_vars, cash, ev, num,
                              paths, use_cross, itm_only, var_shift, var_scale, coefficients,
                              price_out, stream=None):
    """ls_first_pass_gpu with the library's own per-GPU scratch array as the
    workspace (no allocation; calls on one GPU must not overlap)."""
    ves = _host_ints(vars_each_step)
    sh, sc = _host_f32(var_shift), _host_f32(var_scale)
    call("ls_first_pass_noalloc_gpu", n_ex, ptr(basis_variables), ves.ctypes.data, max_vars,
         ptr(cash), ptr(ev), ptr(num), paths, int(use_cross), int(itm_only), ptr(sh), ptr(sc),
         ptr(coefficients), ptr(price_out), stream_handle(stream))


def ls_multi_workspace(paths):
    n = ctypes.c_size_t(0)
    call("ls_multi_workspace", paths, ctypes.addressof(n))
    return n.value


def ls_multi_first_pass_gpu(n_ex, basis_variables, vars_each_step, max_vars, cash, ev, num, paths,
                            use_cross, normalise, max_regressions, lower_path_cutoff, lower_frac,
                            upper_frac, sd_guess, multiplier, workspace, coefficients, lower_cuts,
                            upper_cuts, means, sds, price_out, stream=None, selector=0):
    ves = _host_ints(vars_each_step)
    call("ls_multi_first_pass_gpu", n_ex, ptr(basis_variables), ves.ctypes.data, max_vars,
         ptr(cash), ptr(ev), ptr(num), paths, int(use_cross), int(normalise),
         int(max_regressions), int(lower_path_cutoff), int(selector), float(lower_frac),
         float(upper_frac), float(sd_guess), float(multiplier), ptr(workspace),
         ptr(coefficients), ptr(lower_cuts), ptr(upper_cuts), ptr(means), ptr(sds),
         ptr(price_out), stream_handle(stream))


def ls_second_pass_gpu(n_ex, basis_variables, vars_each_step, max_vars, cash, ev, num, paths,
                       use_cross, itm_only, var_shift, var_scale, coefficients, workspace,
                       price_out, stream=None):
    ves = _host_ints(vars_each_step)
    sh, sc = _host_f32(var_shift), _host_f32(var_scale)
    call("ls_second_pass_gpu", n_ex, ptr(basis_variables), ves.ctypes.data, max_vars, ptr(cash),
         ptr(ev), ptr(num), paths, int(use_cross), int(itm_only), ptr(sh), ptr(sc),
         ptr(coefficients), ptr(workspace), ptr(price_out), stream_handle(stream))


def ls_normal_eq_gpu(paths, n_basis, basis_functions, target, mask, workspace, gram_out,
                     stream=None):
    call("ls_normal_eq_gpu", paths, n_basis, ptr(basis_functions), ptr(target), ptr(mask),
         ptr(workspace), ptr(gram_out), stream_handle(stream))


def cholesky_solve_f64_gpu(n, A, b, x, info=None, stream=None):
    call("cholesky_solve_f64_gpu", n, ptr(A), ptr(b), ptr(x), ptr(info), stream_handle(stream))


def adjoinBasisVariablesCaller_gpu(use_logs, float_data, rate1, rate2, rate3, paths_per_batch,
                                   total_paths, paths_previously_done, n_ex, basis_variables,
                                   stream=None):
    call("adjoinBasisVariablesCaller_gpu", int(use_logs), None, ptr(float_data), None, None,
         ptr(rate1), ptr(rate2), ptr(rate3), paths_per_batch, total_paths, paths_previously_done,
         0, n_ex, None, None, ptr(basis_variables), stream_handle(stream))


def adjoinExerciseValues_Bermudan_swaption_gpu(exercise_values, aux_float_data, n_ex,
                                               paths_for_output, output_offset, batch_paths,
                                               rates1, rates2, stream=None):
    call("adjoinExerciseValues_Bermudan_swaption_gpu", ptr(exercise_values), ptr(aux_float_data),
         None, None, None, n_ex, paths_for_output, output_offset, batch_paths, 0, ptr(rates1),
         ptr(rates2), None, None, None, stream_handle(stream))


def expandBasisFunctions_quadratic_gpu(total_paths, basis_variables, exercise_number, max_size,
                                       n_vars, use_cross, basis_functions, stream=None):
    call("expandBasisFunctions_quadratic_gpu", total_paths, ptr(basis_variables), exercise_number,
         max_size, n_vars, int(use_cross), ptr(basis_functions), stream_handle(stream))


def continuation_net_value_evaluation_gpu(paths, exercise_number, basis_functions, n_basis,
                                          weights, exercise_values, shift, out, stream=None):
    call("continuation_net_value_evaluation_gpu", paths, exercise_number, ptr(basis_functions),
         n_basis, ptr(weights), ptr(exercise_values), float(shift), ptr(out),
         stream_handle(stream))


def oneStepUpdateLS_multi_gpu(paths, exercise_number, basis_functions, n_basis, row_size,
                              max_depth, weights, lower_cuts, upper_cuts, continuation_values,
                              cash, exercise_values, shift, est, exercises, stream=None):
    call("oneStepUpdateLS_multi_gpu", paths, exercise_number, ptr(basis_functions), n_basis,
         row_size, max_depth, ptr(weights), ptr(lower_cuts), ptr(upper_cuts),
         ptr(continuation_values), ptr(cash), ptr(exercise_values), float(shift), ptr(est),
         ptr(exercises), stream_handle(stream))


def updateDeflationOfContinuationValues_gpu(values, numeraires, paths, new_step, old_step,
                                            stream=None):
    call("updateDeflationOfContinuationValues_gpu", ptr(values), ptr(numeraires), paths, new_step,
         old_step, stream_handle(stream))


def reduceOuterProduct_gpu(paths, row_size, col_size, symmetric, threads, blocks, in1, in2,
                           answer, workspace, stream=None):
    call("reduceOuterProduct_gpu", paths, row_size, col_size, int(symmetric), threads, blocks,
         ptr(in1), ptr(in2), ptr(answer), ptr(workspace), stream_handle(stream))


def reduceOuterProductSymmetric_cublas(paths, row_size, alpha, beta, inp, answer, workspace,
                                       stream=None):
    call("reduceOuterProductSymmetric_cublas", paths, row_size, float(alpha), float(beta),
         ptr(inp), ptr(answer), ptr(workspace), stream_handle(stream))


def outer_product_workspace_floats(paths):
    return lib().outer_product_workspace_floats(paths)


def PointwiseProduct_gpu(blocks, threads, paths, in1, in2, out, stream=None):
    call("PointwiseProduct_gpu", blocks, threads, paths, ptr(in1), ptr(in2), ptr(out),
         stream_handle(stream))


def PointwiseProductSymmetricUpperTriangular_gpu(blocks, threads, paths, rows, inp, out,
                                                 stream=None):
    call("PointwiseProductSymmetricUpperTriangular_gpu", blocks, threads, paths, rows, ptr(inp),
         ptr(out), stream_handle(stream))


def PointwiseProductMultipleFirst_gpu(blocks, threads, paths, in1, in2, out, rows, stream=None):
    call("PointwiseProductMultipleFirst_gpu", blocks, threads, paths, ptr(in1), ptr(in2), ptr(out),
         rows, stream_handle(stream))


# ---------------------------------------------------------------- multi-asset Black-Scholes

def cholesky_lower(corr):
    c = _f64(corr)
    L = np.zeros_like(c)
    call("koo_cholesky_lower", c.shape[0], c.ctypes.data, L.ctypes.data)
    return L


def correlated_paths_multistep_GPU(d_in, d_out, A, alive, A_offset_per_step, factors,
                                   out_dimensions, number_paths, steps, stream=None):
    call("correlated_paths_multistep_GPU", ptr(d_in), ptr(d_out), ptr(A), ptr(alive),
         A_offset_per_step, factors, out_dimensions, number_paths, steps, stream_handle(stream))


def multi_dim_BS_path_generator_kernel(log0, incr, out, paths, rates, steps, stream=None):
    call("multi_dim_BS_path_generator_kernel", ptr(log0), ptr(incr), ptr(out), paths, rates, steps,
         stream_handle(stream))


def basketWeightings_gpu(inp, out, weights, paths, stocks, steps, stream=None):
    call("basketWeightings_gpu", ptr(inp), ptr(out), ptr(weights), paths, stocks, steps,
         stream_handle(stream))


def basket_asian_workspace(paths, assets):
    n = _sz()
    call("basket_asian_workspace", paths, assets, ctypes.addressof(n))
    return n.value


def basket_asian_gpu(paths, assets, steps, variates, log_spots, vols, scaled_chol, weights, rate,
                     dt, strike, partials, out, stream=None):
    call("basket_asian_gpu", paths, assets, steps, ptr(variates), ptr(log_spots), ptr(vols),
         ptr(scaled_chol), ptr(weights), float(rate), float(dt), float(strike), ptr(partials),
         ptr(out), stream_handle(stream))


# ---------------------------------------------------------------- frozen Bermudan

def bermudan_adjoint_workspace(paths, rates, factors, checkpoint_every):
    a, b, c, d = _sz(), _sz(), _sz(), _sz()
    call("bermudan_adjoint_workspace", paths, rates, factors, checkpoint_every,
         ctypes.addressof(a), ctypes.addressof(b), ctypes.addressof(c), ctypes.addressof(d))
    return a.value, b.value, c.value, d.value


def bermudan_frozen_fwd(paths, rates, factors, k, variates, f0, taus, disp, A, strike,
                        exercise_steps, coeffs_dev, shift_host, scale_host, itm_only, checkpoints,
                        path_buffers, partials, price_out, stream=None):
    ex = _host_ints(exercise_steps)
    sh = None if shift_host is None else _f64(shift_host)
    sc = None if scale_host is None else _f64(scale_host)
    call("bermudan_frozen_fwd", paths, rates, factors, k, ptr(variates), ptr(f0), ptr(taus),
         ptr(disp), ptr(A), float(strike), len(ex), ex.ctypes.data, ptr(coeffs_dev), ptr(sh),
         ptr(sc), int(itm_only), ptr(checkpoints), ptr(path_buffers), ptr(partials),
         ptr(price_out), stream_handle(stream))


def cashFlowGeneratorEE_cancellable_swap_gpu(cross, f1, f2, aux, strat_float, strat_int, paths,
                                             steps, n_ex, is_ex, r1, r2, r3, strat_vars,
                                             strat_max_size, est, ex_index, stream=None):
    call("cashFlowGeneratorEE_cancellable_swap_gpu", int(cross), ptr(f1), ptr(f2), ptr(aux), None,
         None, ptr(strat_float), strat_float.numel(), ptr(strat_int), strat_int.numel(), paths,
         steps, n_ex, ptr(is_ex), ptr(r1), ptr(r2), ptr(r3), ptr(strat_vars), strat_max_size,
         None, None, ptr(est), ptr(ex_index), stream_handle(stream))


def bermudan_portfolio_prepare(products, descriptors):
    """products: list of BermProduct; descriptors: a device buffer of
    bermudan_portfolio_descriptor_bytes(len(products)) bytes.  Returns the
    plan (host)."""
    arr = (BermProduct * len(products))(*products)
    plan = PortfolioPlan()
    call("bermudan_portfolio_prepare", len(products), ctypes.addressof(arr), ptr(descriptors),
         ctypes.addressof(plan))
    return plan


def bermudan_ls_portfolio_prepare(products, descriptors):
    """products: list of LsProduct; descriptors: a device buffer of
    bermudan_ls_portfolio_descriptor_bytes(len(products)) bytes."""
    arr = (LsProduct * len(products))(*products)
    plan = PortfolioPlan()
    call("bermudan_ls_portfolio_prepare", len(products), ctypes.addressof(arr), ptr(descriptors),
         ctypes.addressof(plan))
    return plan


def bermudan_ls_portfolio_descriptor_bytes(n):
    return lib().bermudan_ls_portfolio_descriptor_bytes(n)


def bermudan_ls_portfolio_run(plan, descriptors, stream=None):
    call("bermudan_ls_portfolio_run", ctypes.addressof(plan), ptr(descriptors),
         stream_handle(stream))


def bermudan_portfolio_descriptor_bytes(n):
    return lib().bermudan_portfolio_descriptor_bytes(n)


def bermudan_portfolio_fwd(plan, descriptors, stream=None):
    call("bermudan_portfolio_fwd", ctypes.addressof(plan), ptr(descriptors), stream_handle(stream))


def bermudan_portfolio_bwd(plan, descriptors, stream=None):
    call("bermudan_portfolio_bwd", ctypes.addressof(plan), ptr(descriptors), stream_handle(stream))


def bermudan_portfolio_persistent_bytes(paths):
    """Device buffer bytes of the persistent portfolio for products of the
    given path counts."""
    arr = np.ascontiguousarray(paths, np.int32)
    return lib().bermudan_portfolio_persistent_bytes(len(arr), arr.ctypes.data)


def bermudan_portfolio_persistent_prepare(gen, ls, adj, buffer, work=None):
    """gen: list of GenProduct or None (variates given); ls: list of
    LsProduct; adj: list of BermProduct (the same products in the same
    order); buffer: a device buffer of bermudan_portfolio_persistent_bytes;
    work: per-product work weights for the team sizes (None: the library's
    model).  Returns the plan (host)."""
    n = len(adj)
    garr = (GenProduct * n)(*gen) if gen is not None else None
    larr = (LsProduct * n)(*ls)
    aarr = (BermProduct * n)(*adj)
    w = None if work is None else np.ascontiguousarray(work, np.float64)
    plan = PersistentPlan()
    call("bermudan_portfolio_persistent_prepare", n,
         ctypes.addressof(garr) if garr is not None else None, ctypes.addressof(larr),
         ctypes.addressof(aarr), None if w is None else w.ctypes.data, ptr(buffer),
         ctypes.addressof(plan))
    return plan


def persistent_block_ticks(plan, buffer):
    """Per-block run times (100 MHz ticks) of the last persistent run, as a
    host array of plan.grid entries (synchronises)."""
    import torch
    t = buffer[plan.ticks_offset:plan.ticks_offset + 8 * plan.grid].cpu().numpy()
    return t.view(np.uint64).copy()


def bermudan_portfolio_persistent_run(plan, buffer, stream=None):
    call("bermudan_portfolio_persistent_run", ctypes.addressof(plan), ptr(buffer),
         stream_handle(stream))


def bermudan_frozen_bwd(paths, rates, factors, k, variates, f0, taus, disp, A, vols, strike,
                        checkpoints, path_buffers, scratch, partials, greeks_out, stream=None):
    call("bermudan_frozen_bwd", paths, rates, factors, k, ptr(variates), ptr(f0), ptr(taus),
         ptr(disp), ptr(A), ptr(vols), float(strike), ptr(checkpoints), ptr(path_buffers),
         ptr(scratch), ptr(partials), ptr(greeks_out), stream_handle(stream))
