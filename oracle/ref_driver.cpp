// ref_driver.cpp -- C entry points onto the reference's own gold CPU routines.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile.ref against the
// reference sources where they lie (kooderive-code-r256/LINUX/gold, the
// Linux-portable copy of gold/) into oracle/_ref/libkoo_ref.so.  The tests
// use it to pin the FP64 restatement in koo_oracle.c against the reference's
// actual code; the product never touches it.  Every function below only
// marshals plain arrays into the gold routine named in its comment.
#include <cstdint>
#include <cstring>
#include <vector>

#include <gold/Bridge_allocator_gold.h>
#include <gold/Bridge_gold.h>
#include <gold/Correlation_gold.h>
#include <gold/LMM_evolver_full_gold.h>
#include <gold/LMM_evolver_gold.h>
#include <gold/MultiD_Path_Gen_BS_gold.h>
#include <gold/MultiLS_regression.h>
#include <gold/Regression_Selector_concrete_gold.h>
#include <gold/cashFlowDiscounting_gold.h>
#include <gold/cashFlowGeneration_earlyEx_product.h>
#include <gold/early_exercise_value_generator_gold.h>
#include <gold/LS_regression.h>
#include <gold/cashFlowGeneration_product_gold.h>
#include <gold/correlate_drift_gold.h>
#include <gold/math/Normals_gold.h>
#include <gold/scramble_gold.h>
#include <gold/sobol_gold.h>
#include <gold/sobol_primitives.h>
#include <gold/volstructs_gold.h>
#include <gold/weighted_average_gold.h>

template <class T>
static std::vector<T> vec(const T* p, size_t n) { return std::vector<T>(p, p + n); }

extern "C" {

// gold/sobol_gold.cpp:118 initSobolDirectionVectors
void ref_sobol_directions(int n_dims, unsigned* dirs) { initSobolDirectionVectors(n_dims, dirs); }

// gold/sobol_primitives.cpp:71 sobol_primitives[] (the compiled joe-kuo-6.10200
// table): degree, a and the max_m m values of the first n rows.
void ref_sobol_primitives(int n, unsigned* degree, unsigned* a, unsigned* m) {
  for (int d = 0; d < n; ++d) {
    degree[d] = sobol_primitives[d].degree;
    a[d] = sobol_primitives[d].a;
    for (int i = 0; i < max_m; ++i) m[(size_t)d * max_m + i] = sobol_primitives[d].m[i];
  }
}

// gold/sobol_gold.cpp:191 sobolCPUintsOffset
void ref_sobol_ints_offset(int n_vectors, int n_dims, int offset, unsigned* dirs, unsigned* out) {
  sobolCPUintsOffset(n_vectors, n_dims, offset, dirs, out);
}

// gold/scramble_gold.cpp:11 scramble_gold
void ref_scramble(const unsigned* in, unsigned* out, const unsigned* scr, int dims, int paths) {
  std::vector<unsigned> i(in, in + (size_t)dims * paths), o((size_t)dims * paths), s(scr, scr + dims);
  scramble_gold(i, o, s, dims, paths);
  std::memcpy(out, o.data(), o.size() * sizeof(unsigned));
}

// gold/math/normals_gold.h:191 inverseCumulativeShawBrickmanUnsignedIntGold<float>
float ref_invnorm_sb(unsigned x) { return inverseCumulativeShawBrickmanUnsignedIntGold<float>()(x); }

// gold/Bridge_allocator_gold.cpp:38 bridge_allocate_diagonal
void ref_bridge_allocate_diagonal(int factors, int steps, int* idx) {
  std::vector<int> v(factors * steps);
  bridge_allocate_diagonal(v, factors, steps);
  std::memcpy(idx, v.data(), v.size() * sizeof(int));
}

// gold/Bridge_gold.h:168-252 BrownianBridge<double>::GenerateBridge on each
// vector of [S][n]
void ref_bridge_pow2(int log2, int vectors, const double* in, double* out) {
  BrownianBridge<double> bb(log2);
  const int S = bb.dimensions();
  std::vector<double> v(S), o(S);
  for (int p = 0; p < vectors; ++p) {
    for (int k = 0; k < S; ++k) v[k] = in[(size_t)k * vectors + p];
    bb.GenerateBridge(&v[0], &o[0]);
    for (int k = 0; k < S; ++k) out[(size_t)k * vectors + p] = o[k];
  }
}

// gold/Bridge_gold.h:291-400 BrownianBridgeMultiDim<double>::GenerateBridge
// per path ([F][S] per path in the gold), written as (S, P, F)
void ref_bridge_multi(int log2, int factors, int ordering, int paths, const double* in,
                      double* out) {
  BrownianBridgeMultiDim<double> bb(log2, factors,
                                    static_cast<BrownianBridgeMultiDim<double>::ordering>(ordering));
  const int S = 1 << log2, N = S * factors;
  std::vector<double> v(N), o(N);
  for (int p = 0; p < paths; ++p) {
    for (int i = 0; i < N; ++i) v[i] = in[(size_t)i * paths + p];
    bb.GenerateBridge(&v[0], &o[0]);
    for (int f = 0; f < factors; ++f)
      for (int s = 0; s < S; ++s) out[((size_t)s * paths + p) * factors + f] = o[f * S + s];
  }
}

// gold/cashFlowGeneration_earlyEx_product.h:26-86 cashFlowGeneratorEE_gold with
// Swap_gold, earlyExerciseNullGold and LSAMultiExerciseStrategy over the
// quadratic (cross) generators, built from the GPU strategy's packed arrays.
}  // extern "C"
template <class Gen>
static void ref_ee_impl(double* f1, double* f2, const double* aux, int aux_n, const double* fd,
                        const int* id, int paths, int steps, int E, const unsigned char* is_ex,
                        const double* r1, const double* sv, int sv_max) {
  const int depth = id[0], maxv = id[1], row = id[2];
  std::vector<int> vps(id + 3, id + 3 + E);
  std::vector<double> shifts(fd, fd + E);
  const double* msh = fd + E;
  const double* mdv = msh + (size_t)E * depth * maxv;
  const double* lo = mdv + (size_t)E * depth * maxv;
  const double* up = lo + (size_t)E * depth;
  const double* wts = up + (size_t)E * depth;
  std::vector<double> weights(wts, wts + (size_t)E * depth * row);
  std::vector<double> means(msh, msh + (size_t)E * depth * maxv), sds(mdv, mdv + (size_t)E * depth * maxv);
  std::vector<double> lower(lo, lo + (size_t)E * depth), upper(up, up + (size_t)E * depth);
  MatrixConstFacade<double> L(&lower[0], E, depth), U(&upper[0], E, depth);
  CubeConstFacade<double> M(&means[0], E, depth, maxv), D(&sds[0], E, depth, maxv);
  std::vector<Gen> producers;
  for (int e = 0; e < E; ++e) producers.push_back(Gen(vps[e]));
  LSAMultiExerciseStrategy<Gen, double> strategy(E, vps, shifts, weights, depth, L, U, M, D,
                                                 producers, maxv);
  std::vector<double> a(aux, aux + aux_n);
  Swap_gold<double> swap(a, steps);
  earlyExerciseNullGold<double> null_value;
  std::vector<bool> ex(steps);
  for (int i = 0; i < steps; ++i) ex[i] = is_ex[i] != 0;
  std::vector<double> rates(r1, r1 + (size_t)steps * paths), dummy(1, 0.0);
  std::vector<double> svv(sv, sv + (size_t)E * sv_max * paths);
  CubeConstFacade<double> svc(&svv[0], E, sv_max, paths);
  std::vector<double> g1, g2;
  cashFlowGeneratorEE_gold(g1, g2, swap, null_value, strategy, paths, steps, ex, rates, rates,
                           rates, svc, dummy, dummy);
  std::memcpy(f1, g1.data(), g1.size() * sizeof(double));
  std::memcpy(f2, g2.data(), g2.size() * sizeof(double));
}
extern "C" {

void ref_cashflows_ee(int cross, double* f1, double* f2, const double* aux, int aux_n,
                      const double* fd, const int* id, int paths, int steps, int E,
                      const unsigned char* is_ex, const double* r1, const double* sv, int sv_max) {
  if (cross)
    ref_ee_impl<quadraticPolynomialCrossGenerator>(f1, f2, aux, aux_n, fd, id, paths, steps, E,
                                                   is_ex, r1, sv, sv_max);
  else
    ref_ee_impl<quadraticPolynomialGenerator>(f1, f2, aux, aux_n, fd, id, paths, steps, E, is_ex,
                                              r1, sv, sv_max);
}

// gold/correlate_drift_gold.h:22 correlate_drift_paths_gold<double>
void ref_correlate_drift(const double* in, double* out, const double* A, int a_off, const double* drift,
                         int factors, int out_dims, int paths, int steps) {
  std::vector<double> i = vec(in, (size_t)steps * paths * factors);
  std::vector<double> o((size_t)steps * out_dims * paths);
  std::vector<double> a = vec(A, (size_t)steps * a_off);
  std::vector<double> d = vec(drift, (size_t)steps * out_dims);
  correlate_drift_paths_gold<double>(i, o, a, a_off, d, factors, out_dims, paths, steps);
  std::memcpy(out, o.data(), o.size() * sizeof(double));
}

// gold/LMM_evolver_gold.h:178 spotDriftComputer<double>::getDrifts
void ref_spot_drifts(const double* A_step, const double* rates, const double* taus, const double* disp,
                     int factors, int n_rates, double* drifts) {
  std::vector<double> t = vec(taus, n_rates), d = vec(disp, n_rates), r = vec(rates, n_rates);
  std::vector<double> out(n_rates);
  spotDriftComputer<double> dc(t, factors, d);
  MatrixConstFacade<double> A(A_step, n_rates, factors);
  dc.getDrifts(A, r, out);
  std::memcpy(drifts, out.data(), n_rates * sizeof(double));
}

// gold/LMM_evolver_gold.h:391 LMM_evolver_pc_gold<double>
void ref_lmm_pc(const double* f0, const double* taus, const double* incr, const double* A, const double* disp,
                int paths, int factors, int steps, int rates, double* rates_out, double* logs_out) {
  std::vector<double> i = vec(f0, rates), t = vec(taus, rates), d = vec(disp, rates);
  std::vector<double> c = vec(incr, (size_t)steps * rates * paths);
  std::vector<double> a = vec(A, (size_t)steps * rates * factors);
  std::vector<double> ro, lo;
  LMM_evolver_pc_gold<double>(i, t, c, a, d, paths, factors, steps, rates, ro, lo, 100.0);
  std::memcpy(rates_out, ro.data(), ro.size() * sizeof(double));
  std::memcpy(logs_out, lo.data(), lo.size() * sizeof(double));
}

// gold/LMM_evolver_gold.h:236 LMM_evolver_Euler_gold<double>
void ref_lmm_euler(const double* f0, const double* taus, const double* incr, const double* A,
                   const double* disp, int paths, int factors, int steps, int rates,
                   double* rates_out, double* logs_out) {
  std::vector<double> i = vec(f0, rates), t = vec(taus, rates), d = vec(disp, rates);
  std::vector<double> c = vec(incr, (size_t)steps * rates * paths);
  std::vector<double> a = vec(A, (size_t)steps * rates * factors);
  std::vector<double> ro, lo;
  LMM_evolver_Euler_gold<double>(i, t, c, a, d, paths, factors, steps, rates, ro, lo);
  std::memcpy(rates_out, ro.data(), ro.size() * sizeof(double));
  std::memcpy(logs_out, lo.data(), lo.size() * sizeof(double));
}

// gold/LMM_evolver_gold.h:460 discount_ratios_computation_gold<double>
void ref_discount_ratios(const double* rates, const double* taus, const int* alive, int paths, int steps,
                         int n_rates, double* discounts) {
  std::vector<double> r = vec(rates, (size_t)steps * n_rates * paths), t = vec(taus, n_rates);
  std::vector<int> al = vec(alive, steps);
  std::vector<double> out((size_t)steps * (n_rates + 1) * paths, 0.0);
  discount_ratios_computation_gold<double>(r, t, al, paths, steps, n_rates, out, 0);
  std::memcpy(discounts, out.data(), out.size() * sizeof(double));
}

// gold/LMM_evolver_gold.h:494 coterminal_annuity_ratios_computation_gold<double>
void ref_annuities(const double* disc, const double* taus, const int* alive, int paths, int steps, int n_rates,
                   double* ann) {
  std::vector<double> d = vec(disc, (size_t)steps * (n_rates + 1) * paths), t = vec(taus, n_rates);
  std::vector<int> al = vec(alive, steps);
  std::vector<double> out((size_t)steps * n_rates * paths, 0.0);
  coterminal_annuity_ratios_computation_gold<double>(d, t, al, paths, steps, n_rates, out, 0);
  std::memcpy(ann, out.data(), out.size() * sizeof(double));
}

// gold/LMM_evolver_gold.h:525 coterminal_swap_rates_computation_gold<double>
void ref_cot_swap_rates(const double* disc, const double* ann, const int* alive, int paths, int steps,
                        int n_rates, double* sr) {
  std::vector<double> d = vec(disc, (size_t)steps * (n_rates + 1) * paths);
  std::vector<double> a = vec(ann, (size_t)steps * n_rates * paths);
  std::vector<int> al = vec(alive, steps);
  std::vector<double> out((size_t)steps * n_rates * paths, 0.0);
  coterminal_swap_rates_computation_gold<double>(d, a, al, paths, steps, n_rates, out, 0);
  std::memcpy(sr, out.data(), out.size() * sizeof(double));
}

// gold/LMM_evolver_gold.h:560 swap_rate_computation_gold<double>
void ref_swap_rate(const double* disc, const double* taus, int start, int end, int paths, int step_in,
                   int step_out, int steps, int n_rates, double* out) {
  std::vector<double> d = vec(disc, (size_t)steps * (n_rates + 1) * paths), t = vec(taus, n_rates);
  std::vector<double> o((size_t)steps * paths, 0.0);
  swap_rate_computation_gold<double>(d, t, start, end, paths, step_in, step_out, steps, n_rates, o);
  std::memcpy(out, o.data(), o.size() * sizeof(double));
}

// gold/LMM_evolver_gold.h:614 spot_measure_numeraires_computation_offset_gold<double>
void ref_numeraires(const double* disc, int paths, int n_rates, double* out) {
  std::vector<double> d = vec(disc, (size_t)n_rates * (n_rates + 1) * paths);
  std::vector<double> o((size_t)n_rates * paths, 0.0);
  spot_measure_numeraires_computation_offset_gold<double>(d, paths, paths, 0, n_rates, o);
  std::memcpy(out, o.data(), o.size() * sizeof(double));
}

// gold/LMM_evolver_gold.h:642 forward_rate_extraction_gold<double>
void ref_forward_extraction(const double* fwd, const int* idx, int paths, int steps, int n_rates, double* out) {
  std::vector<double> f = vec(fwd, (size_t)steps * n_rates * paths);
  std::vector<int> ix = vec(idx, steps);
  std::vector<double> o((size_t)steps * paths, 0.0);
  forward_rate_extraction_gold<double>(f, ix, paths, steps, n_rates, o, 0);
  std::memcpy(out, o.data(), o.size() * sizeof(double));
}

// gold/cashFlowGeneration_product_gold.h cashFlowGeneratorCallerSwap_gold<double>
void ref_swap_flows(double* f1, double* f2, const double* aux, int aux_n, int paths, int steps, const double* r1) {
  std::vector<double> o1((size_t)steps * paths), o2((size_t)steps * paths), a = vec(aux, aux_n);
  std::vector<double> rr = vec(r1, (size_t)steps * paths), dummy(1, 0.0);
  cashFlowGeneratorCallerSwap_gold<double>(o1, o2, a, paths, steps, rr, rr, rr, dummy, dummy);
  std::memcpy(f1, o1.data(), o1.size() * sizeof(double));
  std::memcpy(f2, o2.data(), o2.size() * sizeof(double));
}

// gold/cashFlowGeneration_product_gold.h cashFlowGeneratorCallerTARN_gold<double>
void ref_tarn_flows(double* f1, double* f2, const double* aux, int aux_n, int paths, int steps, const double* r1,
                    const double* r2) {
  std::vector<double> o1((size_t)steps * paths), o2((size_t)steps * paths), a = vec(aux, aux_n);
  std::vector<double> x1 = vec(r1, (size_t)steps * paths), x2 = vec(r2, (size_t)steps * paths), dummy(1, 0.0);
  cashFlowGeneratorCallerTARN_gold<double>(o1, o2, a, paths, steps, x1, x2, x2, dummy, dummy);
  std::memcpy(f1, o1.data(), o1.size() * sizeof(double));
  std::memcpy(f2, o2.data(), o2.size() * sizeof(double));
}

// gold/cashFlowDiscounting_gold.h:210 generateCashFlowIndicesAndWeights<double>
void ref_cash_flow_indices_weights(const double* rate_times, int n_rate_times, const double* pay, int n,
                                   int* first, int* second, double* thetas) {
  std::vector<int> a, b;
  std::vector<double> th, rt = vec(rate_times, n_rate_times), pt = vec(pay, n);
  generateCashFlowIndicesAndWeights<double>(a, b, th, rt, pt);
  std::memcpy(first, a.data(), n * sizeof(int));
  std::memcpy(second, b.data(), n * sizeof(int));
  std::memcpy(thetas, th.data(), n * sizeof(double));
}

// gold/cashFlowDiscounting_gold.h:20 cashFlowDiscounting_gold<double>
void ref_cash_flow_discounting(const int* first, const int* second, const double* thetas, const double* disc,
                               const double* flows, const double* num, int paths, int steps, double* out,
                               double* summed) {
  std::vector<int> a = vec(first, steps), b = vec(second, steps);
  std::vector<double> th = vec(thetas, steps), d = vec(disc, (size_t)steps * (steps + 1) * paths);
  std::vector<double> f = vec(flows, (size_t)steps * paths), nv = vec(num, (size_t)steps * paths), o, s;
  cashFlowDiscounting_gold<double>(a, b, th, d, f, nv, paths, steps, o, s, 0);
  std::memcpy(out, o.data(), o.size() * sizeof(double));
  std::memcpy(summed, s.data(), s.size() * sizeof(double));
}

// gold/cashFlowDiscounting_gold.h:97-205 cashFlowDiscountingPartial_gold<double>:
// flows generated at step gen[k], discounted log-linearly to the payment time
// and by the numeraire ratio to step target[k]; n_flows <= steps.
void ref_cash_flow_discounting_partial(int n_flows, const int* gen, const int* first,
                                       const int* second, const double* thetas, const int* target,
                                       const double* disc, const double* flows, const double* num,
                                       int paths, int steps, double* out) {
  std::vector<int> g = vec(gen, n_flows), a = vec(first, n_flows), b = vec(second, n_flows);
  std::vector<int> t = vec(target, n_flows);
  std::vector<double> th = vec(thetas, n_flows);
  std::vector<double> d = vec(disc, (size_t)steps * (steps + 1) * paths);
  // the gold sizes the flow matrices by steps (n_flows <= steps rows used)
  std::vector<double> f((size_t)steps * paths, 0.0), nv = vec(num, (size_t)steps * paths);
  std::memcpy(f.data(), flows, sizeof(double) * (size_t)n_flows * paths);
  std::vector<double> o((size_t)steps * paths, 0.0);
  CubeConstFacade<double> dc(&d[0], steps, steps + 1, paths);
  MatrixConstFacade<double> fm(&f[0], steps, paths), nm(&nv[0], steps, paths);
  MatrixFacade<double> om(&o[0], steps, paths);
  cashFlowDiscountingPartial_gold<double>(g, a, b, th, t, dc, fm, nm, paths, steps, om);
  std::memcpy(out, o.data(), sizeof(double) * (size_t)n_flows * paths);
}

// gold/volstructs_gold.cpp:236 FlatVolPseudoRootsOfCovariances (exp-long correlation)
void ref_flat_vol_pseudo_roots(const double* rate_times, const double* evol_times, const double* vols,
                               int n_rates, int steps, int factors, double L, double beta, double* A) {
  std::vector<double> rt = vec(rate_times, n_rates + 1), et = vec(evol_times, steps), v = vec(vols, n_rates);
  Cube_gold<double> c = FlatVolPseudoRootsOfCovariances(rt, et, v, factors, L, beta);
  std::memcpy(A, &c(0, 0, 0), sizeof(double) * steps * n_rates * factors);
}

// gold/MultiLS_regression.cpp:523 generateRegressionCoefficientsViaLSMultiQuadratic_gold,
// one regression per date (maxNumberOfExtraRegressions = 0); with `normalise` the gold
// standardises each basis variable over the regression paths first.
double ref_ls_multi_depth0(int n_ex, const double* basis_vars, const int* vars_each_step, int max_vars,
                           const double* cash, const double* ex_values, const double* numeraires,
                           int paths, int use_cross, int normalise, double* coeffs_out) {
  std::vector<double> products, targets;
  const int maxd = use_cross ? (max_vars == 0 ? 1 : max_vars == 1 ? 3 : max_vars == 2 ? 6 : 1 + 3 * max_vars)
                             : 1 + 2 * max_vars;
  std::vector<double> coeffs((size_t)n_ex * 1 * maxd, 0.0), lower(n_ex, 0.0), upper(n_ex, 0.0);
  std::vector<double> means((size_t)n_ex * max_vars, 0.0), sds((size_t)n_ex * max_vars, 1.0);
  CubeFacade<double> coeff_cube(&coeffs[0], n_ex, 1, maxd);
  MatrixFacade<double> lower_mat(&lower[0], n_ex, 1), upper_mat(&upper[0], n_ex, 1);
  CubeFacade<double> means_cube(&means[0], n_ex, 1, max_vars), sds_cube(&sds[0], n_ex, 1, max_vars);
  std::vector<double> bv(basis_vars, basis_vars + (size_t)n_ex * max_vars * paths);
  std::vector<double> cf(cash, cash + (size_t)n_ex * paths), ev(ex_values, ex_values + (size_t)n_ex * paths);
  std::vector<double> nm(numeraires, numeraires + (size_t)n_ex * paths);
  CubeConstFacade<double> bv_cube(&bv[0], n_ex, max_vars, paths);
  MatrixConstFacade<double> cf_mat(&cf[0], n_ex, paths), ev_mat(&ev[0], n_ex, paths);
  MatrixFacade<double> nm_mat(&nm[0], n_ex, paths);
  std::vector<int> ves(vars_each_step, vars_each_step + n_ex);
  // Fractions that make the selector keep every path without its cut search
  // (Regression_Selector_concrete_gold.cpp: the early return when the
  // pre-selected count lies between the two targets): the cuts are unused at
  // depth 0, and the search can loop forever when many net values are exactly 0.
  RegressionSelector_gold_Fraction selector(0.5, 1.5, 2.0, 0.8);
  double price = generateRegressionCoefficientsViaLSMultiQuadratic_gold(
      n_ex, products, targets, coeff_cube, lower_mat, upper_mat, means_cube, sds_cube, normalise != 0, bv_cube, ves,
      max_vars, cf_mat, ev_mat, nm_mat, paths, 0, 10, selector, use_cross != 0);
  std::memcpy(coeffs_out, &coeffs[0], sizeof(double) * coeffs.size());
  return price;
}

// The same with extra regressions and the fraction selector
// (gold/MultiLS_regression.cpp:20-420), variables normalised or not.
double ref_ls_multi(int n_ex, const double* basis_vars, const int* vars_each_step, int max_vars,
                    const double* cash, const double* ex_values, const double* numeraires, int paths,
                    int use_cross, int normalise, int max_regressions, int lower_path_cutoff, double lower_frac,
                    double upper_frac, double sd_guess, double multiplier, double* coeffs_out,
                    double* lower_out, double* upper_out) {
  std::vector<double> products, targets;
  const int maxd = use_cross ? (max_vars == 0 ? 1 : max_vars == 1 ? 3 : max_vars == 2 ? 6 : 1 + 3 * max_vars)
                             : 1 + 2 * max_vars;
  const int R = max_regressions;
  std::vector<double> coeffs((size_t)n_ex * R * maxd, 0.0), lower((size_t)n_ex * R, 0.0),
      upper((size_t)n_ex * R, 0.0);
  std::vector<double> means((size_t)n_ex * R * max_vars, 0.0), sds((size_t)n_ex * R * max_vars, 1.0);
  CubeFacade<double> coeff_cube(&coeffs[0], n_ex, R, maxd);
  MatrixFacade<double> lower_mat(&lower[0], n_ex, R), upper_mat(&upper[0], n_ex, R);
  CubeFacade<double> means_cube(&means[0], n_ex, R, max_vars), sds_cube(&sds[0], n_ex, R, max_vars);
  std::vector<double> bv(basis_vars, basis_vars + (size_t)n_ex * max_vars * paths);
  std::vector<double> cf(cash, cash + (size_t)n_ex * paths), ev(ex_values, ex_values + (size_t)n_ex * paths);
  std::vector<double> nm(numeraires, numeraires + (size_t)n_ex * paths);
  CubeConstFacade<double> bv_cube(&bv[0], n_ex, max_vars, paths);
  MatrixConstFacade<double> cf_mat(&cf[0], n_ex, paths), ev_mat(&ev[0], n_ex, paths);
  MatrixFacade<double> nm_mat(&nm[0], n_ex, paths);
  std::vector<int> ves(vars_each_step, vars_each_step + n_ex);
  RegressionSelector_gold_Fraction selector(lower_frac, upper_frac, sd_guess, multiplier);
  double price = generateRegressionCoefficientsViaLSMultiQuadratic_gold(
      n_ex, products, targets, coeff_cube, lower_mat, upper_mat, means_cube, sds_cube, normalise != 0, bv_cube,
      ves, max_vars, cf_mat, ev_mat, nm_mat, paths, R - 1, lower_path_cutoff, selector, use_cross != 0);
  std::memcpy(coeffs_out, &coeffs[0], sizeof(double) * coeffs.size());
  std::memcpy(lower_out, &lower[0], sizeof(double) * lower.size());
  std::memcpy(upper_out, &upper[0], sizeof(double) * upper.size());
  return price;
}

// gold/Correlation_gold.h:22-43 correlated_paths_gold<double>, every step
// (the gold loop starts at path 1; path 0 of the output is left at zero).
void ref_correlated_paths(const double* in, double* out, const double* A, int factors, int out_dims,
                          int paths, int steps) {
  std::vector<double> i = vec(in, (size_t)steps * paths * factors);
  std::vector<double> o((size_t)steps * out_dims * paths, 0.0);
  std::vector<double> a = vec(A, (size_t)out_dims * factors);
  for (int s = 0; s < steps; ++s) correlated_paths_gold<double>(i, o, a, factors, out_dims, paths, s);
  std::memcpy(out, o.data(), o.size() * sizeof(double));
}

// gold/MultiD_Path_Gen_BS_gold.h (LINUX copy):85-101 multi_dim_BS_path_generator_nonlog_gold<double>.
void ref_multid_bs_paths(const double* log0, const double* incr, const double* drifts, double* out,
                         int paths, int rates, int steps) {
  std::vector<double> l = vec(log0, rates), c = vec(incr, (size_t)steps * rates * paths), o;
  std::vector<double> d = vec(drifts, (size_t)steps * rates);
  multi_dim_BS_path_generator_nonlog_gold<double>(l, c, d, o, paths, rates, steps);
  std::memcpy(out, o.data(), o.size() * sizeof(double));
}

// gold/weighted_average_gold.h:19-51 basketWeightings_gold<double>.
void ref_basket_weightings(const double* in, double* out, const double* w, int paths, int stocks,
                           int steps) {
  std::vector<double> i = vec(in, (size_t)steps * stocks * paths), o((size_t)steps * paths);
  std::vector<double> wv = vec(w, stocks);
  CubeConstFacade<double> ic(&i[0], steps, stocks, paths);
  MatrixFacade<double> om(&o[0], steps, paths);
  basketWeightings_gold<double>(ic, om, wv, paths, stocks, steps);
  std::memcpy(out, o.data(), o.size() * sizeof(double));
}

// The reference's own CPU pipeline for the plain-LMM payer swap:
// gold/LMM_evolver_full_gold.h:140-230 LMMEvolutionRoutineGold<double> (Sobol,
// scramble, inverse normal, power-of-two bridge, correlate + drift, PC
// evolver), then discount ratios, spot numeraires, extraction of the rate
// fixing at each step, swap flows and their discounting (the gold routines
// wrapped above).  Returns the mean discounted value; the timing of this call
// is the reference CPU baseline of the workload.
double ref_swap_pipeline(int paths, int rates, int factors, int power_of_two,
                         const unsigned* scramblers, const double* A, const double* fixed_drifts,
                         const double* disp, const double* f0, const double* taus, double strike) {
  const int S = rates;
  std::vector<unsigned> scr(scramblers, scramblers + (1 << power_of_two) * factors);
  std::vector<double> a = vec(A, (size_t)S * rates * factors), fd = vec(fixed_drifts, (size_t)S * rates);
  std::vector<double> d = vec(disp, rates), f = vec(f0, rates), t = vec(taus, rates), lf(rates);
  for (int r = 0; r < rates; ++r) lf[r] = std::log(f[r] + d[r]);
  std::vector<double> er, el;
  LMMEvolutionRoutineGold<double>(paths, 0, rates, factors, S, power_of_two, scr, a, fd, d, f, lf, t,
                                  er, el);
  std::vector<int> alive(S);
  for (int s = 0; s < S; ++s) alive[s] = s;
  std::vector<double> disc((size_t)S * (rates + 1) * paths, 0.0), num((size_t)S * paths, 0.0);
  discount_ratios_computation_gold<double>(er, t, alive, paths, S, rates, disc, 0);
  spot_measure_numeraires_computation_offset_gold<double>(disc, paths, paths, 0, rates, num);
  std::vector<int> idx(S);
  for (int s = 0; s < S; ++s) idx[s] = s;
  std::vector<double> r1((size_t)S * paths, 0.0);
  forward_rate_extraction_gold<double>(er, idx, paths, S, rates, r1, 0);
  std::vector<double> aux(2 + 2 * S);
  aux[0] = strike;
  aux[1] = 1.0;
  for (int s = 0; s < S; ++s) aux[2 + s] = aux[2 + S + s] = t[s];
  std::vector<double> f1((size_t)S * paths), f2((size_t)S * paths), dummy(1, 0.0);
  cashFlowGeneratorCallerSwap_gold<double>(f1, f2, aux, paths, S, r1, r1, r1, dummy, dummy);
  std::vector<int> first(S), second(S);
  std::vector<double> th(S, 0.0), o, s1, s2;
  for (int s = 0; s < S; ++s) first[s] = second[s] = s + 1;
  cashFlowDiscounting_gold<double>(first, second, th, disc, f1, num, paths, S, o, s1, 0);
  cashFlowDiscounting_gold<double>(first, second, th, disc, f2, num, paths, S, o, s2, 0);
  double tot = 0.0;
  for (int p = 0; p < paths; ++p) tot += s1[p] + s2[p];
  return tot / paths;
}

}  // extern "C"
