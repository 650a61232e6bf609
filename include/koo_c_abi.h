/*
 * koo_c_abi.h -- C ABI of the MI355X (gfx950) LMM Monte Carlo pricing library
 * (libkoo_amd.so).
 *
 * Every entry point keeps the name and the parameter order of the reference
 * host-callable launcher it replaces (cited per function, paths relative to
 * kooderive-code-r256/), with three uniform changes:
 *   - it returns int status (KOO_OK = 0) instead of void / exit() on error;
 *   - it takes a trailing `void* stream` (a hipStream_t; NULL = null stream)
 *     and never synchronises, so calls are stream-ordered and capturable;
 *   - C++-typed parameters (std::vector<int>&, cudaStream_t, bool) become
 *     plain `const int*` + count, `void*` and `int`.
 * All pointers are device pointers unless stated; the caller owns every
 * buffer.  Layouts are the reference's path-minor ones: [D][P] for Sobol
 * draws, (S,F,P) for bridged increments, (S,N,P) for rates, (S,N+1,P) for
 * discount ratios, (S,N,F) for pseudo-roots.
 */
#ifndef KOO_C_ABI_H
#define KOO_C_ABI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KOO_OK 0
#define KOO_ERR_INVALID (-1) /* bad argument: nothing launched */
#define KOO_ERR_HIP (-2)     /* HIP launch / runtime failure */

/* Message of the last failing call on this host thread ("" if none). */
const char* koo_last_error(void);
/* Library version string. */
const char* koo_version(void);

/* =====================================================================
 * Stage (a): Sobol integers, scrambling, inverse normal
 * =================================================================== */

/* Host: 32 Bratley-Fox direction numbers per dimension (Joe-Kuo table),
 * written to host memory [n_dimensions][32].
 * Replaces initSobolDirectionVectors, gold/sobol_gold.cpp:118. */
int initSobolDirectionVectors(int n_dimensions, unsigned int* directions_host);

/* d_output[d][i] = Sobol integer of Gray-code point i + path_offset + 1 in
 * dimension d, for i < n_vectors, d < n_dimensions; d_directions is
 * [n_dimensions][32].  Replaces sobolGPUInts, sobol_int_gpu.h:68. */
int sobolGPUInts(int n_vectors, int n_dimensions, unsigned path_offset,
                 unsigned int* d_directions, unsigned int* d_output, void* stream);

/* In place: device_data[d][p] ^= scramblers[d] for d < number_dim, p < paths.
 * Replaces scramble_gpu, scramble_gpu.h:8. */
int scramble_gpu(unsigned int* device_data, unsigned int* scramblers, int number_dim, int paths,
                 void* stream);

/* out[i] = normal quantile of in[i] (tail-exact Shaw-Brickman, FP32).
 * Replaces the thrust::transform of inverseCumulativeShawBrickmanUnsignedInt
 * in LMM_evolver_full.cu:330-334 (no named launcher in the reference). */
int inverse_normal_u32_gpu(const unsigned int* in, float* out, long long n, void* stream);

/* Fused generator: out[d][i] = invnorm(sobol(d, i + path_offset + 1) ^ s_d)
 * with s_d = scramblers[d] for d < number_scrambled, else 0.  One write of
 * FP32 normals; no integer buffer.  New (no reference counterpart): fuses
 * sobolGPUInts + scramble_gpu + the inverse-normal transform. */
int sobol_scramble_normal_gpu(int n_vectors, int n_dimensions, unsigned path_offset,
                              const unsigned int* d_directions, const unsigned int* scramblers,
                              int number_scrambled, float* out, void* stream);

/* =====================================================================
 * Host set-up (FP64, host memory): correlation, PCA pseudo-roots, drifts
 * =================================================================== */

/* corr[i][j] = L + (1-L) exp(-beta |t_i - t_j|), corr[i][i] = 1.
 * Replaces GenerateExpLongCorrMatrix, gold/volstructs_gold.cpp:157. */
int koo_exp_corr(const double* times, int n, double L, double beta, double* corr);

/* Lower Cholesky factor (row-major, L L^T = corr) of a positive-definite
 * n x n matrix.  New: the full-rank correlation root of the multi-asset
 * Black-Scholes path (the reference feeds an arbitrary A to
 * correlated_paths_multistep_GPU, Correlate_gpu.h:10-19). */
int koo_cholesky_lower(int n, const double* corr, double* L);

/* Rank-`factors` PCA root of a symmetric n x n matrix (largest eigenvalues),
 * rows rescaled to keep the diagonal when normalise != 0; out is n x factors.
 * Replaces pseudoSqrtSchur, gold/math/pseudoSquareRoot.cpp:76. */
int koo_pseudo_sqrt(int n, const double* cov, int factors, int normalise, double* out);

/* A (steps, n_rates, factors): per-step pseudo-roots of the flat-vol
 * covariance with exponential correlation on rate_times[0..n_rates).
 * rate_times has n_rates+1 entries.  Replaces FlatVolPseudoRootsOfCovariances,
 * gold/volstructs_gold.cpp:236. */
int FlatVolPseudoRootsOfCovariances(const double* rate_times, const double* evol_times,
                                    const double* vols, int n_rates, int steps, int factors,
                                    double L, double beta, double* A);

/* Factor model A(s,r,f) = sigma_r sqrt(T_s - T_{s-1}) B(r,f) for r >= s, 0
 * otherwise; B (n_rates, factors) is the renormalised PCA root of the rate
 * correlation (optional output).  New: makes dA/dsigma_r = A/sigma_r exact. */
int koo_factor_pseudo_roots(const double* rate_times, const double* vols, int n_rates,
                            int factors, double L, double beta, double* A, double* B);

/* out(s,r) = -0.5 sum_f A(s,r,f)^2.  Replaces the fixed-drift loop of
 * gold/LMM_evolver_classes_gold.h:193-205. */
int koo_fixed_drifts(const double* A, int steps, int n_rates, int factors, double* out);

/* Spot-measure drifts of one step.  Replaces spotDriftComputer::getDrifts,
 * gold/LMM_evolver_gold.h:178-231. */
int spotDriftComputer_getDrifts(const double* A_step, const double* rates, const double* taus,
                                const double* displacements, int factors, int n_rates, int alive,
                                double* drifts);

/* =====================================================================
 * Stage (b): Brownian bridge, correlation + drift, LMM evolution
 * =================================================================== */

/* Host: diagonal allocation of the F*S Sobol dimensions to (factor, step),
 * indices [F][S].  Replaces bridge_allocate_diagonal,
 * gold/Bridge_allocator_gold.cpp:38. */
int bridge_allocate_diagonal(int factors, int steps, int* indices_host);

/* Jaeckel bridge on t_i = i+1 plus successive differencing, per factor:
 * in [D][P] (dimension indices(f,i) feeds bridge variate i of factor f),
 * out (S,F,P).  Replaces brownian_bridge::transform, brownian_bridge.h:49
 * (a C++ class in the reference; this is its C shim). */
int brownian_bridge_gpu(int steps, const int* allocation_indices, const float* in, float* out,
                        int paths, int factors, void* stream);

/* Fused stage (a)+(b): Sobol point -> scramble -> normal -> bridge, written
 * once as (S,F,P) increments; D = steps*factors dimensions, dimension
 * indices(f,i) scrambled with scramblers[indices(f,i)] when it is below
 * number_scrambled.  New (no reference counterpart): replaces the chain
 * SobDeviceInts -> scramble_main -> transform -> bb.transform of
 * LMM_evolver_full.cu:312-346. */
int sobol_bridge_gpu(int paths, unsigned path_offset, int steps, int factors,
                     const unsigned int* d_directions, const unsigned int* scramblers,
                     int number_scrambled, const int* allocation_indices, float* out,
                     void* stream);

/* Bridge_gpu.h:11 (brownianBridgeGPU_constant_memory): the legacy
 * power-of-two bridge (BrownianBridge<float>, gold/Bridge_gold.h:168-252) on
 * n_vectors vectors of 2^n_poweroftwo points, [points][vectors] in and out
 * (differenced).  n_poweroftwo <= 8; the construction scalars travel in the
 * kernel arguments (no constant-memory state, so calls are reentrant). */
int brownianBridgeGPU_constant_memory(int n_vectors, int n_poweroftwo, const float* input,
                                      float* output, void* stream);

/* Bridge_multi_gpu.h:12-20 (brownianBridgeMultiGPUReorder): Sobol dimensions
 * [2^n F][P] to the (points, paths, factors) arrangement the power-of-two
 * bridge then transforms as points x (paths * factors) vectors (the
 * reference's MultiDBridge, Bridge_main.cu:111-173).  allocator: 0 row,
 * 1 column, 2 triangular (BrownianBridgeMultiDim<float>::ordering); useTextures
 * is ignored. */
int brownianBridgeMultiGPUReorder(const float* input, float* output, int n_poweroftwo,
                                  int factors, int number_paths, int allocator, int useTextures,
                                  void* stream);

/* out(s,r,p) = sum_f A(s,r,f) in(s,p,f) + drift_fixed(s,r) for r >= alive[s].
 * Replaces correlated_drift_paths_GPU, Correlate_drift_gpu.h:12. */
int correlated_drift_paths_GPU(float* input_data_device, float* output_data_device,
                               float* A_device, int* alive_device, float* drift_fixed_device,
                               int A_offset_per_step, int factors, int out_dimensions,
                               int number_paths, int steps, void* stream);

/* Predictor-corrector log-Euler over all steps (spot measure, displaced
 * diffusion) with fused discount ratios.  variates (S,F,P) when
 * transposeData != 0, else (S,P,F).  Writes rates and log-rates (S,N,P) for
 * alive rates and, when doDiscounts != 0, discount ratios (S,N+1,P).
 * e_buffer_device / e_buffer_pred_device / sharedMemForLogs / numberThreads
 * are accepted for signature compatibility and ignored (the running drift
 * sums live in registers).  Replaces LMM_evolver_pc_all_gpu,
 * LMM_evolver_all.h:22-46. */
int LMM_evolver_pc_all_gpu(float* initial_rates_device, float* initial_log_rates_device,
                           float* taus_device, float* variates_device, float* pseudo_roots_device,
                           float* initial_drifts_device, float* fixed_drifts_device,
                           float* displacements_device, int* alive_device, int paths,
                           int factors, int steps, int rates, float* e_buffer_device,
                           float* e_buffer_pred_device, float* evolved_rates_device,
                           float* evolved_log_rates_device, float* discounts_device,
                           int sharedMemForLogs, int transposeData, void* stream,
                           int numberThreads, int doDiscounts);

/* The per-step evolvers on correlated increments (S,N,P) = A z + fixed
 * drift (the output of correlated_drift_paths_GPU): predictor-corrector
 * (replaces LMM_evolver_pc_gpu, LMM_evolver_gpu.h:16-32) and log-Euler with
 * the log of the displaced rate capped at log 10 (replaces
 * LMM_evolver_euler_gpu, LMM_evolver_gpu.h:35-50), both in one launch over
 * all steps.  alive_indices_host (host, `steps` entries) replaces the
 * reference's std::vector aliveIndices; step 0 evolves every rate from the
 * initial drifts, later steps the rates r >= alive[s] (the other entries of
 * the output cubes are not written).  e buffers and useShared are accepted
 * and ignored. */
int LMM_evolver_pc_gpu(float* initial_rates_device, float* initial_log_rates_device,
                       float* taus_device, float* correlated_increments_device,
                       float* pseudo_roots_device, float* initial_drifts_device,
                       float* displacements_device, const int* alive_indices_host, int paths,
                       int factors, int steps, int rates, float* e_buffer_device,
                       float* e_buffer_pred_device, float* evolved_rates_device,
                       float* evolved_log_rates_device, int useShared, void* stream);
int LMM_evolver_euler_gpu(float* initial_rates_device, float* initial_log_rates_device,
                          float* taus_device, float* correlated_increments_device,
                          float* pseudo_roots_device, float* initial_drifts_device,
                          float* displacements_device, const int* alive_indices_host, int paths,
                          int factors, int steps, int rates, float* e_buffer_device,
                          float* evolved_rates_device, float* evolved_log_rates_device,
                          void* stream);

/* P(s,alive) = 1, P(s,r+1) = P(s,r)/(1 + tau_r f(s,r)).  alive_host has
 * `steps` entries (host copy of the reference's aliveIndices vector),
 * alive_device the same on the device.  Replaces
 * discount_ratios_computation_gpu, LMM_evolver_gpu.h:55-64. */
int discount_ratios_computation_gpu(float* evolved_rates_device, float* taus_device,
                                    const int* alive_host, int* alive_device, int paths,
                                    int steps, int rates, float* discounts_device,
                                    int allStepsAtOnce, void* stream);

/* Ann(s,r) = sum_{j>=r} tau_j P(s,j+1).  Replaces
 * coterminal_annuity_ratios_computation_gpu, LMM_evolver_gpu.h:68-75. */
int coterminal_annuity_ratios_computation_gpu(float* discounts_device, float* taus_device,
                                              const int* alive_host, int paths, int steps,
                                              int rates, float* annuities_device, void* stream);

/* SR(s,r) = (P(s,r) - P(s,N)) / Ann(s,r).  Replaces
 * coterminal_swap_rate_computation_gpu, LMM_evolver_gpu.h:78-85. */
int coterminal_swap_rate_computation_gpu(float* discounts_device, float* annuities_device,
                                         const int* alive_host, int paths, int steps, int rates,
                                         float* cot_swap_rates_device, void* stream);

/* Swap rate from startIndex to endIndex on the curve of step_for_offset_in,
 * written to row step_for_offset_out of a (.,P) matrix.  Replaces
 * swap_rate_computation_gpu, LMM_evolver_gpu.h:88-97. */
int swap_rate_computation_gpu(float* discounts_device, float* taus_device, int startIndex,
                              int endIndex, int paths, int step_for_offset_in,
                              int step_for_offset_out, int rates, float* swap_rate_device,
                              void* stream);

/* Discretely compounded money-market numeraire, B(0) = 1,
 * B(i) = B(i-1) P(i-1,i-1)/P(i-1,i); out (N,P).  Replaces
 * spot_measure_numeraires_computation_gpu, LMM_evolver_gpu.h:100-103. */
int spot_measure_numeraires_computation_gpu(float* discount_factors_global,
                                            float* numeraire_values_global, int paths,
                                            int n_rates, void* stream);

/* As above, written at column offset pathsOffsetForOutput of a
 * (N, pathsForOutput) matrix.  Replaces
 * spot_measure_numeraires_computation_offset_gpu, LMM_evolver_gpu.h:104-111. */
int spot_measure_numeraires_computation_offset_gpu(float* discount_factors_global,
                                                   float* numeraire_values_global, int paths,
                                                   int pathsForOutput, int pathsOffsetForOutput,
                                                   int n_rates, void* stream);

/* out(s,p) = f(s, forwardIndices[s], p).  forwardIndices_host has `steps`
 * entries.  Replaces forward_rate_extraction_gpu, LMM_evolver_gpu.h:115-121. */
int forward_rate_extraction_gpu(float* all_forwards_device, const int* forwardIndices_host,
                                int paths, int steps, int rates, float* select_forwards_device,
                                void* stream);

/* out(i,p) = f(stepIndices[i], forwardIndices[i], p) for i < n_indices.
 * Replaces forward_rate_extraction_selecting_step_gpu,
 * LMM_evolver_gpu.h:123-130. */
int forward_rate_extraction_selecting_step_gpu(float* all_forwards_device,
                                               const int* forwardIndices_host,
                                               const int* stepIndices_host, int n_indices,
                                               int paths, int rates,
                                               float* select_forwards_device, void* stream);

/* out[offsetForOutput + p + pathsOutput*i] = all[p + pathsInput*exerciseIndices[i]].
 * exerciseIndices is a device array.  Replaces
 * spot_measure_numeraires_extraction_gpu, LMM_evolver_gpu.h:132-140. */
int spot_measure_numeraires_extraction_gpu(float* all_numeraire_values_global,
                                           float* some_numeraire_values_global, int pathsInput,
                                           int pathsOutput, int offsetForOutput,
                                           int numberRates, int numberExerciseTimes,
                                           int* exerciseIndices_vec, void* stream);

/* =====================================================================
 * Stage (c): cash flows, discounting, aggregation, reductions
 * =================================================================== */

/* Swap of rate1 against a fixed coupon; aux_data = [K, payReceive,
 * accruals1[numberSteps], accruals2[numberSteps]] (device).
 * Replaces cashFlowGeneratorCallerSwap, cashFlowGeneration_product_gpu.h:190. */
int cashFlowGeneratorCallerSwap(float* genFlows1, float* genFlows2, float* aux_data, int paths,
                                int numberSteps, float* rates1, float* rates2, float* rates3,
                                float* forwards, float* discRatios, void* stream);

/* Target redemption note; aux_data = [totalCoupon, K, m, accruals1[S],
 * accruals2[S]].  Replaces cashFlowGeneratorCallerTARN,
 * cashFlowGeneration_product_gpu.h:111. */
int cashFlowGeneratorCallerTARN(float* genFlows1, float* genFlows2, float* aux_data, int paths,
                                int numberSteps, float* rates1, float* rates2, float* rates3,
                                float* forwards, float* discRatios, void* stream);

/* Caplet strip on rate1 (new product, same interface as Swap):
 * aux_data = [K, accruals[numberSteps]]; genFlows1 = acc max(rate1 - K, 0),
 * genFlows2 = 0. */
int cashFlowGeneratorCallerCaplet(float* genFlows1, float* genFlows2, float* aux_data, int paths,
                                  int numberSteps, float* rates1, float* rates2, float* rates3,
                                  float* forwards, float* discRatios, void* stream);

/* cashFlowGeneratorEE_gpu<Swap, earlyExerciseNull,
 * LSAMultiExerciseStrategyQuadratic(Cross)_gpu> (cashFlowGen_earlyExProduct_gpu.h:
 * 28-172, 323-497, 756-1049; the instantiations of LSmulti_test.cu:2185-2236 and
 * bermudanIRDPricer.cu:1391-1440): the cancellable swap's second pass under a
 * Longstaff-Schwartz multi-regression strategy.  use_cross_terms picks the
 * cross strategy.  The strategy arrays are the reference's packed layout
 * (outputDataVectors): int data = depth, max variables, row size, variables
 * per date [E]; float data = Andersen shifts [E], variable shifts and
 * divisors [E][depth][maxVars], lower and upper cuts [E][depth], weights
 * [E][depth][row].  isExerciseDate is one byte per step (the reference's
 * bool array); rates (steps, paths); strategyVariables (E, maxSize, paths);
 * estContinuationValues (steps, paths) is written at exercise dates;
 * exerciseIndices[p] = steps processed.  The exercise-value tables, forwards
 * and discount ratios are unused (earlyExerciseNull), as in the reference. */
int cashFlowGeneratorEE_cancellable_swap_gpu(
    int use_cross_terms, float* genFlows1, float* genFlows2, const float* productData,
    const float* exerciseValueDataFloat, const int* exerciseValueDataInt,
    const float* exerciseStrategyDataFloat, int exerciseStrategyDataFloat_size,
    const int* exerciseStrategyDataInt, int exerciseStrategyDataInt_size, int paths,
    int numberSteps, int numberExerciseDates, const unsigned char* isExerciseDate,
    const float* rates1, const float* rates2, const float* rates3, const float* strategyVariables,
    int strategyVariables_maxSize, const float* forwards, const float* discRatios,
    float* estContinuationValues, int* exerciseIndices, void* stream);

/* Discount flows generated at step s, paid between rate times
 * firstIndex[s] and secondIndex[s] (log-linear weight theta[s]), to the
 * numeraire: discountedFlows(s,p) and summedDiscountedFlows(p).  The two
 * flags are accepted and ignored.  Replaces cashFlowDiscounting_gpu,
 * cashFlowDiscounting_gpu.h:14-27. */
int cashFlowDiscounting_gpu(int* firstIndex, int* secondIndex, float* theta,
                            float* discountRatios, float* genFlows, float* numeraireValues,
                            int paths, int numberSteps, int useTextures, int useShared,
                            float* discountedFlows, float* summedDiscountedFlows, void* stream);

/* Discount each of numberFlowsPerPath flows (generated at step
 * generationIndices[i]) to rate time targetIndices[i] via the numeraire.
 * Replaces cashFlowDiscounting_partial_gpu, cashFlowDiscounting_gpu.h:29-42. */
int cashFlowDiscounting_partial_gpu(int numberFlowsPerPath, int* generationIndices_dev,
                                    int* firstIndex_dev, int* secondIndex_dev, float* theta_dev,
                                    int* targetIndices_dev, float* discountRatios_dev,
                                    float* genFlows_dev, float* numeraireValues_dev, int paths,
                                    int numberSteps, float* discountedFlows_dev, void* stream);

/* aggregatedFlows(precedingIndex[s], offSetForOutput + p) += genFlows(s,p)
 * for start <= s < numberSteps (accumulates, never overwrites).
 * Replaces AggregateFlows_gpu, cashFlowAggregation_gpu.h:18-27. */
int AggregateFlows_gpu(float* aggregatedFlows, int totalPaths, int numberExerciseDates,
                       float* genFlows, int batchPaths, int offSetForOutput, int numberSteps,
                       int* precedingIndex, int start, void* stream);

/* Partial sums: out[i] = sum of data[8i .. 8i+7], repeated until at most
 * stopPoint values remain; *result receives the pointer into workspace
 * (>= points floats) that holds them and *remaining their count.  As in the
 * reference, a level whose length is not a multiple of 8 drops its last
 * (length % 8) values.
 * Replaces PartialReduce_using8s_gpu, reductions_gpu.h:9-14. */
int PartialReduce_using8s_gpu(const float* data, float* workspace, int points, int threads,
                              int stopPoint, float** result, int* remaining, void* stream);
/* Same with 16-way sums.  Replaces PartialReduce_using16s_gpu,
 * reductions_gpu.h:15-19. */
int PartialReduce_using16s_gpu(const float* data, float* workspace, int points, int threads,
                               int stopPoint, float** result, int* remaining, void* stream);

/* *final = sum of in[0..size); wsp holds >= 1024 floats.  threads/blocks are
 * hints and may be 0.  Replaces reduceSinglePass, reduction_thread_fence.h:13. */
int reduceSinglePass(int size, int threads, int blocks, float* in, float* wsp, float* final,
                     void* stream);

/* FP64 compensated sum of n floats or doubles into *out (device double);
 * workspace holds at least koo_sum_workspace_doubles(n) doubles.  New. */
long long koo_sum_workspace_doubles(long long n);
int koo_sum_f32_to_f64(const float* in, long long n, double* workspace, double* out, void* stream);
int koo_sum_f64(const double* in, long long n, double* workspace, double* out, void* stream);

/* =====================================================================
 * Fused LMM pricing with adjoint Greeks (new; the reference lists adjoint
 * Greeks as an open task, doc/todo.txt:10).
 *
 * Model: N rates, steps == N, alive(s) = s, pseudo-roots A (S,N,F) with
 * fixed drift -0.5 sum_f A^2, step-0 drift from the initial rates, spot
 * measure.  Product KOO_PRODUCT_CAPLETS: caplet i fixes f_i at step i and pays
 * tau_i max(f_i - K, 0) at T_{i+1} (product_aux = [K]).  Prices are in units
 * of the numeraire at T_0 (B_0 = 1).
 *
 * Forward: price (device double) and, when checkpoints != NULL, the
 * log-rate state at the start of every checkpoint_every-th step plus the
 * numeraire there (with checkpoint_every == 1 also each step's final drift
 * sums, so the backward sweep recomputes nothing), then the model's step
 * table (fixed drifts and drift constants, rates*(rates+1) floats) that the
 * backward reads.  Backward: one reverse sweep per path from those
 * checkpoints (recomputing the steps in between into segment_scratch), giving
 * greeks_out[0..N) = dV/df_i(0) and greeks_out[N..2N) = dV/dsigma_i with
 * A(s,i,f) proportional to sigma_i.  Sizes of all device buffers come from
 * lmm_price_workspace().  rates <= 90, factors <= 8.  product_aux is a HOST
 * array (read at call time);
 * every other pointer is a device pointer.  variates are (S,F,P); price_out
 * and greeks_out receive means over paths (FP64).
 * =================================================================== */
#define KOO_PRODUCT_CAPLETS 1

int lmm_price_workspace(int paths, int rates, int factors, int checkpoint_every,
                        size_t* checkpoint_bytes, size_t* scratch_bytes, size_t* partial_bytes);

int lmm_price_fwd(int paths, int rates, int factors, int checkpoint_every, const float* variates,
                  const float* initial_rates, const float* taus, const float* displacements,
                  const float* pseudo_roots, int product, const float* product_aux,
                  float* checkpoints, double* partials, double* price_out, void* stream);

int lmm_price_bwd(int paths, int rates, int factors, int checkpoint_every, const float* variates,
                  const float* initial_rates, const float* taus, const float* displacements,
                  const float* pseudo_roots, const float* vols, int product,
                  const float* product_aux, const float* checkpoints, float* segment_scratch,
                  double* partials, double* greeks_out, void* stream);

/* =====================================================================
 * Stage (d): Longstaff-Schwartz.
 *
 * Layouts are the reference's: basis variables (E, maxVars, P), basis
 * functions (k, P), exercise values / cash flows / numeraires (E, P), all
 * path-minor FP32 device arrays.  Host arrays are marked.
 * =================================================================== */

/* LS_Basis_gpu.h:168-184 (basisVariableExample / basisVariableLog_gpu):
 * three variables per exercise date from rate1..3 (E, batch) into
 * (E, 3, total_paths) at column offset paths_previously_done; with use_logs,
 * log(rate_i + float_data[i]) (float_data is a device array of 3 shifts).
 * integer_data, forwards, discount_ratios, steps_and_rates and the exercise
 * index arrays are accepted for signature parity and unused, as in the
 * reference kernels. */
int adjoinBasisVariablesCaller_gpu(int use_logs, const int* integer_data, const float* float_data,
                                   const float* forwards, const float* discount_ratios,
                                   const float* rate1, const float* rate2, const float* rate3,
                                   int paths_per_batch, int total_paths, int paths_previously_done,
                                   int steps_and_rates, int number_exercise_times,
                                   const int* exercise_indices, const int* exercise_indicators,
                                   float* basis_variables, void* stream);

/* early_exercise_value_generator_gpu.h:195-210 (earlyExerciseBermudanPayerSwaption,
 * exercise_values_examples.h:49-78): max((rate1 - K) * rate2, 0) with rate1 the
 * co-terminal swap rate and rate2 its annuity, K = aux_float_data[0] (device). */
int adjoinExerciseValues_Bermudan_swaption_gpu(
    float* exercise_values, const float* aux_float_data, const int* aux_int_data,
    const int* exercise_indices, const int* exercise_indicators, int number_exercise_times,
    int paths_for_output, int output_offset, int batch_paths, int number_steps_and_rates,
    const float* rates1, const float* rates2, const float* rates3, const float* forwards,
    const float* disc_ratios, void* stream);

/* LS_gpu.h:21-29: quadratic basis functions of exercise date
 * exercise_number's variables (cross terms: 1, x_i, neighbour products,
 * squares; plain: 1, x_0, x_0^2, ...).  At most 4 variables. */
int expandBasisFunctions_quadratic_gpu(int total_paths, const float* basis_variables,
                                       int exercise_number, int max_basis_variable_size,
                                       int variable_size_this_step, int use_cross_terms,
                                       float* basis_functions, void* stream);

/* LS_gpu.h:32-39: out = shift + sum_i w_i basis_i - exercise_values
 * (exercise_values already points at this date, as in the reference). */
int continuation_net_value_evaluation_gpu(int paths, int exercise_number,
                                          const float* basis_functions, int number_basis_functions,
                                          const float* basis_weights, const float* exercise_values,
                                          float andersen_shift, float* output, void* stream);

/* LS_gpu.h:58-72: one date of the multi-regression first pass: nested
 * regressions inside the cut bands, exercise when the net value is negative
 * (continuation_values := exercise value), est/exercises per path. */
int oneStepUpdateLS_multi_gpu(int paths, int exercise_number, const float* basis_functions,
                              int number_basis_functions, int basis_functions_row_size,
                              int max_regression_depth, const float* basis_weights,
                              const float* lower_cuts, const float* upper_cuts,
                              float* continuation_values, const float* deflated_cash_flows,
                              const float* exercise_values, float andersen_shift,
                              float* est_continuation_values, int* exercises, void* stream);

/* LS_gpu.h:77-81: values *= numeraires[new] / numeraires[old]. */
int updateDeflationOfContinuationValues_gpu(float* deflated_next_step_values,
                                            const float* numeraire_values, int paths,
                                            int new_step_number, int old_step_number,
                                            void* stream);

/* outerProduct_gpu.h:12-40: answer[r][c] = sum_p in1[r][p] in2[c][p]
 * (row-major), on MFMA with FP32 accumulation and FP64 block partials.  The
 * workspace keeps the reference's contract (>= blocks*row_size*col_size
 * floats, 8-byte aligned); threads is ignored. */
int reduceOuterProduct_gpu(int paths, int row_size, int col_size, int symmetric, int threads,
                           int blocks, const float* input_data_global1,
                           const float* input_data_global2, float* answer_global,
                           float* workspace_data_global, void* stream);
int reducedOuterProduct_gpu(int paths, int row_size, int col_size, int threads, int blocks,
                            const float* input_data_global1, const float* input_data_global2,
                            float* answer_global, float* workspace_data_global, void* stream);
int reduceOuterProductSymmetric_gpu(int paths, int row_size, int threads, int blocks,
                                    const float* input_data_global1, float* answer_global,
                                    float* workspace_data_global, void* stream);

/* outerProduct_gpu.h:42-67 (the cuBLAS variants): answer = alpha * in1 in2^T
 * + beta * answer, column-major with leading dimension row_size, computed by
 * the same MFMA kernel; the workspace (new argument) holds
 * outer_product_workspace_floats(paths) floats. */
size_t outer_product_workspace_floats(int paths);
int reduceOuterProductSymmetric_cublas(int paths, int row_size, float alpha, float beta,
                                       const float* input_data_global, float* answer_global,
                                       float* workspace_data_global, void* stream);
int reduceOuterProduct_cublas(int paths, int row_size, int col_size, float alpha, float beta,
                              const float* input_data_global1, const float* input_data_global2,
                              float* answer_global, float* workspace_data_global, void* stream);

/* outerProduct_gpu.h:94-115: the per-path products whose path sums are
 * the outer products.  PointwiseProduct: out[p] = in1[p] in2[p];
 * SymmetricUpperTriangular: in (rows, paths) -> out (rows, rows, paths) with
 * out[r][c][p] = in[r][p] in[c][p] for c >= r (the lower triangle is not
 * written); MultipleFirst: out[r][p] = in1[r][p] in2[p].  blocks and
 * threads are accepted for signature parity and ignored. */
int PointwiseProduct_gpu(int blocks, int threads, int paths, const float* in1_global,
                         const float* in2_global, float* out_global, void* stream);
int PointwiseProductSymmetricUpperTriangular_gpu(int blocks, int threads, int paths, int rows,
                                                 const float* in_global, float* out_global,
                                                 void* stream);
int PointwiseProductMultipleFirst_gpu(int blocks, int threads, int paths, const float* in1_global,
                                      const float* in2_global, float* out_global, int rows,
                                      void* stream);

/* New: Bermudan payer swaption path quantities fused with the LMM evolution
 * (same model as lmm_price_fwd).  exercise_steps (HOST, increasing) are
 * evolution steps; per date e: basis_variables[e][0] = co-terminal swap rate
 * SR(t_e), [e][1] = f_{t_e}(T_{t_e}); exercise_values[e] = max((SR - K) A, 0)
 * with A the co-terminal annuity (values at T_{t_e}); numeraires[e] = spot
 * numeraire B_{t_e}.  Replaces the evolve + discount ratios + annuities +
 * swap rates + extraction + exercise value chain of LS_test.cu:680-780. */
int bermudan_paths_gpu(int paths, int rates, int factors, const float* variates,
                       const float* initial_rates, const float* taus, const float* displacements,
                       const float* pseudo_roots, int n_ex, const int* exercise_steps, float strike,
                       float* basis_variables, float* exercise_values, float* numeraires,
                       void* stream);

/* New: the fused LS first pass (backward induction; one regression per date,
 * generateRegressionCoefficientsViaLSMultiQuadratic_gold at depth 0,
 * gold/MultiLS_regression.cpp:20-420).  Per date the normal equations of
 * the quadratic basis in the normalised variables x~ = (x - var_shift) *
 * var_scale (HOST arrays (E, max_vars), NULL for 0 / 1) and of the target
 * cash + deflated next value are accumulated on MFMA (FP32) over all paths
 * or, with itm_only, over paths whose exercise value is positive, summed in
 * FP64 and solved by FP64 Cholesky.  coefficients (device, E x
 * basis_count(max_vars)) apply to x~; price_out (device double) is the
 * first-pass price.  vars_each_step is a HOST array; cash_flows may be NULL
 * (no flows); workspace holds ls_workspace(paths) bytes.  max_vars <= 4.
 * Runs one launch per date and one for the price; with the environment
 * variable KOO_LS_LAUNCH_PER_DATE=0 the whole pass is one persistent launch
 * (bases of at most 7 functions, blocks all resident), with the same
 * results bit for bit. */
int ls_workspace(int paths, size_t* bytes);
int ls_first_pass_gpu(int n_ex, const float* basis_variables, const int* vars_each_step,
                      int max_vars, const float* cash_flows, const float* exercise_values,
                      const float* numeraires, int paths, int use_cross, int itm_only,
                      const float* var_shift, const float* var_scale, void* workspace,
                      double* coefficients, double* price_out, void* stream);

/* No counterpart in the reference.  ls_first_pass_gpu for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): no workspace argument; the pass uses the
 * library's own device array (16 MiB, one per GPU, held until the process
 * ends), so ls_workspace(paths) must not exceed 16 MiB (up to 3 723 200
 * paths on a 256-CU device; larger sizes are rejected with KOO_ERR_INVALID
 * before anything runs).  Calls on one GPU share that array: calls that overlap on a GPU (two
 * streams, two host threads) race and must not be issued.  Same results as
 * ls_first_pass_gpu.  Calls no allocating routine. */
int ls_first_pass_noalloc_gpu(int n_ex, const float* basis_variables, const int* vars_each_step,
                              int max_vars, const float* cash_flows, const float* exercise_values,
                              const float* numeraires, int paths, int use_cross, int itm_only,
                              const float* var_shift, const float* var_scale,
                              double* coefficients, double* price_out, void* stream);

/* generateMultiRegressionCoefficientsViaLS_gpu_flexisolve (LS_main_cula.h:
 * 565-838; the caller in the reference's test is
 * generateRegressionCoefficientsViaLSMultiquadratic_flexi_main,
 * LSmulti_test.cu:1690-1712) with RegressionSelectorFraction(lower_frac,
 * upper_frac, sd_guess, multiplier) (selector 0, RegressionSelectorConcrete.cu:
 * 102-233, as its gold RegressionSelector_gold_Fraction) or
 * RegressionSelectorStandardDeviations(sd_guess) (selector 1, :19-90; the
 * band is open and sd_guess standard deviations wide; lower_frac, upper_frac
 * and multiplier are then unused): the multi-regression LS first pass, as the gold
 * generateRegressionCoefficientsViaLSMultiQuadratic_gold computes it
 * (gold/MultiLS_regression.cpp:20-420).  Per exercise date, up to
 * max_regressions nested regressions: the first over every path, each later
 * one over the paths whose net continuation value lies in the band the
 * selector chose on the previous one ([-c, c] holding between lower_frac and
 * upper_frac of them, or (-c, c) with c = sd_guess standard deviations),
 * while more than lower_path_cutoff paths remain.
 * Variables are standardised per regression (normalise != 0) or raw; the
 * normal equations are solved like the gold's SVD with tolerance 1e-8.
 * FP64 throughout.  Inputs as ls_first_pass_gpu (basis_variables
 * (E, max_vars, P), cash_flows / exercise_values / numeraires (E, P), all
 * device FP32; vars_each_step HOST).  Outputs (device double, zero where a
 * regression was not reached): coefficients (E, max_regressions,
 * basis_count(max_vars)), lower_cuts / upper_cuts (E, max_regressions),
 * means / sds (E, max_regressions, max_vars; may be NULL), price_out (the
 * first-pass price).  workspace: ls_multi_workspace(paths) bytes, 256-byte
 * aligned.  The reference returns the price and fills float coefficients;
 * here everything is a device output, and the pass never synchronises. */
int ls_multi_workspace(int paths, size_t* bytes);
int ls_multi_first_pass_gpu(int n_ex, const float* basis_variables, const int* vars_each_step,
                            int max_vars, const float* cash_flows, const float* exercise_values,
                            const float* numeraires, int paths, int use_cross, int normalise,
                            int max_regressions, int lower_path_cutoff, int selector,
                            double lower_frac, double upper_frac, double sd_guess,
                            double multiplier, void* workspace, double* coefficients,
                            double* lower_cuts, double* upper_cuts, double* means, double* sds,
                            double* price_out, void* stream);

/* New: price under frozen coefficients (SecondPassPriceLSAMultiUsingAggregatedFlows_gold,
 * gold/MultiLS_regression.h:237-289, depth 0). */
int ls_second_pass_gpu(int n_ex, const float* basis_variables, const int* vars_each_step,
                       int max_vars, const float* cash_flows, const float* exercise_values,
                       const float* numeraires, int paths, int use_cross, int itm_only,
                       const float* var_shift, const float* var_scale, const double* coefficients,
                       void* workspace, double* price_out, void* stream);

/* New: Gram matrix of [basis rows; target] over the paths whose mask is
 * non-zero (mask and target may be NULL), (n x n) FP64 with
 * n = n_basis + (target != NULL) <= 16; X^T X and X^T y of the LS normal
 * equations in one MFMA contraction.  workspace: ls_workspace(paths) bytes. */
int ls_normal_eq_gpu(int paths, int n_basis, const float* basis_functions, const float* target,
                     const float* mask, void* workspace, double* gram_out, void* stream);

/* New: FP64 Cholesky solve of A x = b (n < 16, device arrays, A row-major);
 * info (device, may be NULL) = 0, or 1 when A is not positive definite
 * (x = 0).  Replaces the host SVD of LinearSolverConcrete_main.cu:127-188. */
int cholesky_solve_f64_gpu(int n, const double* A, const double* b, double* x, int* info,
                           void* stream);

/* =====================================================================
 * Multi-asset Black-Scholes paths (basket products).
 * =================================================================== */

/* Correlate_gpu.h:10-19: out(s, r, p) = sum_f A(s, r, f) in(s, p, f) for
 * r >= alive[s] (device array); in is (steps, paths, factors), out
 * (steps, out_dimensions, paths), A advances A_offset_per_step floats per step
 * (0: one matrix for all steps). */
int correlated_paths_multistep_GPU(const float* input_data_device, float* output_data_device,
                                   const float* A_device, const int* alive_device,
                                   int A_offset_per_step, int factors, int out_dimensions,
                                   int number_paths, int steps, void* stream);

/* multid_path_gen_BS_gpu.h:10-16: out(s, r, p) = exp(logRates[r] +
 * sum_{u<=s} increments(u, r, p)), (steps, rates, paths). */
int multi_dim_BS_path_generator_kernel(const float* logRates_device,
                                       const float* correlatedBrownianIncrements_device,
                                       float* outputPaths_device, int paths, int rates, int steps,
                                       void* stream);

/* weighted_average_gpu.h:16-19: out(s, p) = sum_j weights[j] in(s, j, p);
 * weights is a device array. */
int basketWeightings_gpu(const float* inputPaths, float* outputPaths, const float* weights,
                         int paths, int numberOfStocks, int numberSteps, void* stream);

/* New: arithmetic-average basket Asian call with pathwise adjoint Greeks in one
 * pass.  x_a(t_{s+1}) = x_a(t_s) + (rate - vol_a^2/2) dt + (C z_s)_a with
 * C = scaled_chol (assets x assets, row-major, C[a][f] = vol_a sqrt(dt)
 * L[a][f], L the lower Cholesky factor of the correlation, zero above the
 * diagonal) and z the bridged variates (steps, assets, paths); payoff
 * e^{-rate T} (sum_{s=1..S} sum_a w_a e^{x_a(t_s)} / S - strike)^+.
 * out (device, FP64): [price, dV/dS_a(0) (assets), dV/dvol_a (assets)] means
 * over paths, the vegas with L held fixed.  All arrays are device arrays;
 * paths a multiple of 4, variates 16-byte aligned, assets <= 64; partials
 * hold basket_asian_workspace() bytes. */
int basket_asian_workspace(int paths, int assets, size_t* partial_bytes);
int basket_asian_gpu(int paths, int assets, int steps, const float* variates,
                     const float* log_spots, const float* vols, const float* scaled_chol,
                     const float* weights, float rate, float dt, float strike, double* partials,
                     double* out, void* stream);

/* =====================================================================
 * Bermudan payer swaption under a frozen exercise strategy: price and
 * pathwise adjoint Greeks (new; the frozen strategy is the reference's
 * second pass, cashFlowGen_earlyExProduct_gpu.h:28-172, and the adjoint is
 * the reference's open task doc/todo.txt:10).
 *
 * Same model and variates as lmm_price_fwd.  The path exercises at the first
 * exercise step whose exercise value ev = (SR - K) A is positive (when
 * itm_only) and exceeds cont = beta_e . phi(x~), phi the quadratic basis with
 * cross term of x~ = ((SR - shift[e][0]) scale[e][0], (f_t - shift[e][1])
 * scale[e][1]) -- the coefficients ls_first_pass_gpu returns.  price_out is
 * the mean of ev / B at exercise; greeks_out[0..N) = dV/df_i(0),
 * [N..2N) = dV/dsigma_i (A proportional to sigma_i), with the exercise
 * decisions held fixed.  coefficients is a DEVICE array (E x 6 doubles, as
 * ls_first_pass_gpu writes them); exercise_steps (increasing, <= 40 dates),
 * var_shift and var_scale (E x 2, NULL for 0 / 1) are HOST arrays; buffers
 * come from bermudan_adjoint_workspace().  The backward reads the path
 * buffers and the checkpoints the forward wrote (the checkpoint buffer also
 * carries the model's step table, written by the forward).  Nothing
 * synchronises, so the whole first pass
 * plus adjoint can be captured in a graph.
 * =================================================================== */
int bermudan_adjoint_workspace(int paths, int rates, int factors, int checkpoint_every,
                               size_t* checkpoint_bytes, size_t* scratch_bytes,
                               size_t* path_bytes, size_t* partial_bytes);
int bermudan_frozen_fwd(int paths, int rates, int factors, int checkpoint_every,
                        const float* variates, const float* initial_rates, const float* taus,
                        const float* displacements, const float* pseudo_roots, float strike,
                        int n_ex, const int* exercise_steps, const double* coefficients,
                        const double* var_shift, const double* var_scale, int itm_only,
                        float* checkpoints, float* path_buffers, double* partials,
                        double* price_out, void* stream);
int bermudan_frozen_bwd(int paths, int rates, int factors, int checkpoint_every,
                        const float* variates, const float* initial_rates, const float* taus,
                        const float* displacements, const float* pseudo_roots, const float* vols,
                        float strike, const float* checkpoints, const float* path_buffers,
                        float* segment_scratch, double* partials, double* greeks_out,
                        void* stream);

/* =====================================================================
 * Portfolios of Bermudan swaptions (new): the frozen forward and the
 * adjoint backward of every product in one launch each.  A product is
 * described as for bermudan_frozen_fwd/_bwd (its own buffers from
 * bermudan_adjoint_workspace; partials shared by its forward and backward);
 * the batch takes factors <= 5 and rates <= 45.  prepare() writes the
 * descriptor table (bermudan_portfolio_descriptor_bytes(n) bytes of device
 * memory) with a synchronous copy, ordering the products longest first,
 * and fills the plan; fwd()/bwd() then launch without synchronising or
 * allocating, so the whole portfolio can be captured in one graph.  Each
 * product's price_out (1 double) and greeks_out (2 x rates doubles) are as
 * in the single-product calls.  Replaces the reference's per-product loop
 * over cashFlowGeneratorEE_kernel second passes
 * (cashFlowGen_earlyExProduct_gpu.h:28-172), which has no adjoint.
 * =================================================================== */
typedef struct koo_berm_product {
  int paths, rates, factors, checkpoint_every;
  const float* variates;      /* (S,F,P) */
  const float* initial_rates; /* N */
  const float* taus;          /* N */
  const float* displacements; /* N */
  const float* pseudo_roots;  /* (S,N,F) */
  const float* vols;          /* N */
  float strike;
  int n_ex;
  const int* exercise_steps;  /* host, increasing */
  const double* coefficients; /* device, E x 6 */
  const double* var_shift;    /* host, E x 2, or NULL */
  const double* var_scale;    /* host, E x 2, or NULL */
  int itm_only;
  float* checkpoints;
  float* path_buffers;
  float* segment_scratch;
  double* partials;
  double* price_out;  /* device */
  double* greeks_out; /* device */
} koo_berm_product;

/* Products are grouped by rate count into up to 4 classes (<= 16, 24, 32
 * and more rates), one launch each, so a block reserves LDS for its class's
 * largest curve rather than the portfolio's. */
typedef struct koo_portfolio_plan {
  int products, blocks, max_rates; /* the 256-lane kernels */
  int ls_blocks, max_ex;           /* the 1024-lane LS kernels (LS batch only) */
  int classes;
  int class_first_block[4], class_blocks[4], class_max_rates[4];
} koo_portfolio_plan;

size_t bermudan_portfolio_descriptor_bytes(int n_products);
int bermudan_portfolio_prepare(int n_products, const koo_berm_product* products,
                               void* descriptors_device, koo_portfolio_plan* plan);
int bermudan_portfolio_fwd(const koo_portfolio_plan* plan, const void* descriptors_device,
                           void* stream);
int bermudan_portfolio_bwd(const koo_portfolio_plan* plan, const void* descriptors_device,
                           void* stream);

/* Portfolio path quantities and LS first passes (new): bermudan_paths_gpu
 * for every product in one launch, then the backward induction of every
 * product date by date, one launch per date rank (round r handles date
 * E_q - 1 - r of product q; a product finishes with its date-0 decision in
 * round E_q), then the first-pass prices.  Per product as in
 * bermudan_paths_gpu + ls_first_pass_gpu with two basis variables per date
 * (co-terminal swap rate, first forward), the quadratic basis with cross
 * term, in-the-money regression when itm_only; paths x rates within the
 * batch limits (factors <= 5, rates <= 50, dates <= 40).  Coefficients and
 * prices are bit-identical to the single-product calls. */
typedef struct koo_ls_product {
  int paths, rates, factors;
  const float* variates;      /* (S,F,P) */
  const float* initial_rates; /* N */
  const float* taus;          /* N */
  const float* displacements; /* N */
  const float* pseudo_roots;  /* (S,N,F) */
  float strike;
  int n_ex;
  const int* exercise_steps; /* host, increasing */
  const float* var_shift;    /* host, E x 2, or NULL */
  const float* var_scale;    /* host, E x 2, or NULL */
  int itm_only;
  float* basis_variables; /* (E,2,P) out */
  float* exercise_values; /* (E,P) out */
  float* numeraires;      /* (E,P) out */
  void* workspace;        /* ls_workspace(paths) bytes */
  double* coefficients;   /* device, E x 6 out */
  double* price_out;      /* device, first-pass price */
  int ls_blocks;          /* 1024-lane blocks of each LS date launch for this product,
                             at most one per 1024 paths and per CU; 0: the default share
                             (about two blocks per CU over the portfolio) */
} koo_ls_product;

size_t bermudan_ls_portfolio_descriptor_bytes(int n_products);
int bermudan_ls_portfolio_prepare(int n_products, const koo_ls_product* products,
                                  void* descriptors_device, koo_portfolio_plan* plan);
int bermudan_ls_portfolio_run(const koo_portfolio_plan* plan, const void* descriptors_device,
                              void* stream);

/* =====================================================================
 * A portfolio of Bermudan swaptions in one persistent launch (new; no
 * counterpart in the reference, which prices product after product):
 * for every product the fused Sobol -> scramble -> normal -> Brownian bridge
 * generation (as sobol_bridge_gpu into its variates; skipped when gen is
 * NULL and the variates are given), the path quantities and LS first pass
 * (as bermudan_paths_gpu + ls_first_pass_gpu), and the frozen forward and
 * adjoint backward under the first pass's coefficients (as
 * bermudan_frozen_fwd/_bwd, checkpoint interval per product), all in one
 * kernel whose resident blocks are split into one team per product.  Each
 * product is described twice, as for the batched calls (ls[i] and adj[i]
 * must agree on sizes, model arrays, strike, dates, itm_only and
 * coefficients; its partials are the backward's exchange words, so they are
 * overwritten), with factors <= 5, rates <= 45, dates <= 40.  Outputs per
 * product: variates (with gen), basis variables, exercise values,
 * numeraires, coefficients, first-pass price, frozen price and 2N Greeks.
 * prepare() sizes the teams in proportion to each product's work (work[i],
 * any positive unit, e.g. block-time measured by an earlier run; NULL: a
 * model of rate updates per phase) and writes the descriptors into
 * device_buffer (bermudan_portfolio_persistent_bytes bytes) with
 * synchronous copies; after a run the buffer holds each block's run time in
 * 100 MHz ticks at ticks_offset (uint64 per block, grid entries); run()
 * launches three kernels (arming, error word, the portfolio) and neither
 * synchronises nor allocates, so it can be captured in a graph.  The
 * frozen Greeks equal bermudan_frozen_fwd/_bwd on the same variates and
 * coefficients bit for bit; the first pass sums its partials over the
 * team's blocks, so its coefficients agree with ls_first_pass_gpu to
 * rounding.  A wait that gives up leaves NaN prices.
 * =================================================================== */
typedef struct koo_gen_product {
  unsigned path_offset;          /* Sobol points path_offset + 1, ... */
  int number_scrambled;          /* dimensions scrambled */
  const unsigned* directions;    /* device, dims x 32 */
  const unsigned* scramblers;    /* device, number_scrambled */
  const int* allocation_indices; /* device, F x S (bridge_allocate_diagonal) */
} koo_gen_product;

typedef struct koo_persistent_plan {
  int products, grid, lds_bytes, tab_offset, blocks_per_cu;
  size_t err_offset, ticks_offset;
  int team_first[512], team_blocks[512]; /* per product, in the order given */
} koo_persistent_plan;

size_t bermudan_portfolio_persistent_bytes(int n_products, const int* paths);
int bermudan_portfolio_persistent_prepare(int n_products, const koo_gen_product* gen,
                                          const koo_ls_product* ls, const koo_berm_product* adj,
                                          const double* work, void* device_buffer,
                                          koo_persistent_plan* plan);
int bermudan_portfolio_persistent_run(const koo_persistent_plan* plan, void* device_buffer,
                                      void* stream);

#ifdef __cplusplus
}
#endif
#endif /* KOO_C_ABI_H */
