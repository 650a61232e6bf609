/*
 * koo_oracle.h -- FP64 CPU oracle for the LMM Monte Carlo pricing path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls
 * this code: it is loaded by the tests, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py, and only as the checker.  Every function is a
 * plain-C restatement of a reference gold routine or kernel (cited per
 * function in koo_oracle.c), computed in double precision.
 *
 * Layout conventions (the reference's, "path-minor"):
 *   Sobol integers / normals       [D][P]          D = steps*factors
 *   bridged increments             (S,F,P)
 *   correlated increments          (S,N,P)
 *   rates, log-rates               (S,N,P)
 *   discount ratios                (S,N+1,P)
 *   pseudo-roots A                 (S,N,F)
 *   fixed drifts                   (S,N)
 *   spot-measure numeraires        (S,P)   (S == N)
 */
#ifndef KOO_ORACLE_H
#define KOO_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- stage (a): Sobol, scramble, inverse normal ---------------------- */
int orc_sobol_directions(int n_dims, const uint32_t *degree, const uint32_t *a_bits,
                         const uint32_t *m /* [n_dims][max_m] */, int max_m,
                         uint32_t *dirs /* [n_dims][32] */);
void orc_sobol_ints(int n_vectors, int n_dims, uint32_t path_offset,
                    const uint32_t *dirs, uint32_t *out /* [n_dims][n_vectors] */);
void orc_scramble(uint32_t *data, const uint32_t *scramblers, int number_dim, int paths);
float orc_invnorm_sb_ref_f32(uint32_t x);  /* reference FP32 formula, verbatim semantics */
double orc_invnorm_u32(uint32_t x);          /* accurate FP64 quantile of the product's u(x) */
double orc_invnorm(double u);                /* accurate FP64 quantile, 0<u<1 */
void orc_invnorm_u32_array(const uint32_t *in, double *out, int64_t n);

/* ---- stage (b): bridge, correlate, evolve ----------------------------- */
void orc_bridge_allocate_diagonal(int factors, int steps, int *indices /* [F][S] */);
void orc_bridge_tables(int steps, int *bridge_index, int *left_index, int *right_index,
                       double *left_weight, double *right_weight, double *std_dev);
void orc_bridge_transform(int steps, int factors, int paths, const int *indices,
                          const double *in /* [D][P] */, double *out /* (S,F,P) */);
void orc_bridge_pow2(int log2, int vectors, const double *in /* [S][n] */, double *out);
void orc_bridge_reorder(const double *in /* [D][P] */, double *out /* (S,P,F) */, int log2,
                        int factors, int paths, int ordering);
void orc_cashflows_ee(int cross, double *f1, double *f2, const double *aux, const double *fd,
                      const int *id, int paths, int steps, int n_ex, const unsigned char *is_ex,
                      const double *r1, const double *sv, int sv_max, double *est,
                      int *ex_index);
void orc_correlate_drift(const double *in /* (S,P,F) */, double *out /* (S,N,P) */,
                         const double *A, const int *alive, const double *drift_fixed,
                         int a_offset_per_step, int factors, int out_dims, int paths, int steps);
void orc_spot_drifts(const double *A_step /* (N,F) */, const double *rates, const double *taus,
                     const double *displacements, int factors, int n_rates, int alive,
                     double *drifts);
void orc_lmm_pc_evolve(const double *initial_rates, const double *initial_log_rates,
                       const double *taus, const double *variates, const double *A,
                       const double *initial_drifts, const double *fixed_drifts,
                       const double *displacements, const int *alive, int paths, int factors,
                       int steps, int rates, int transposed_variates,
                       double *rates_out, double *log_rates_out, double *discounts_out);
void orc_lmm_steps(const double *initial_log_rates, const double *taus, const double *incr,
                   const double *A, const double *initial_drifts, const double *displacements,
                   const int *alive, int paths, int factors, int steps, int rates, int euler,
                   double *rates_out, double *log_rates_out);
void orc_discount_ratios(const double *rates, const double *taus, const int *alive, int paths,
                         int steps, int n_rates, double *discounts);
void orc_coterminal_annuities(const double *discounts, const double *taus, const int *alive,
                              int paths, int steps, int n_rates, double *annuities);
void orc_coterminal_swap_rates(const double *discounts, const double *annuities, const int *alive,
                               int paths, int steps, int n_rates, double *swap_rates);
void orc_swap_rate(const double *discounts, const double *taus, int start_index, int end_index,
                   int paths, int step_in, int step_out, int n_rates, double *swap_rate);
void orc_spot_numeraires(const double *discounts, int paths, int n_rates, double *numeraires);
void orc_forward_extraction(const double *forwards, const int *indices, int paths, int steps,
                            int n_rates, double *out);

/* ---- stage (c): cash flows, discounting, reductions ------------------- */
void orc_swap_flows(double *flows1, double *flows2, const double *aux, int paths, int steps,
                    const double *rates1);
void orc_tarn_flows(double *flows1, double *flows2, const double *aux, int paths, int steps,
                    const double *rates1, const double *rates2);
void orc_caplet_flows(double *flows1, double *flows2, const double *aux, int paths, int steps,
                      const double *rates1);
void orc_cash_flow_indices_weights(const double *rate_times, const double *payment_times,
                                   int n_times, int *first_index, int *second_index,
                                   double *thetas);
void orc_cash_flow_discounting(const int *first_index, const int *second_index,
                               const double *thetas, const double *discounts,
                               const double *flows, const double *numeraires, int paths,
                               int steps, double *discounted, double *summed);
void orc_cash_flow_discounting_partial(int flows_per_path, const int *gen_indices,
                                       const int *first_index, const int *second_index,
                                       const double *thetas, const int *target_indices,
                                       const double *discounts, const double *flows,
                                       const double *numeraires, int paths, int steps,
                                       double *discounted);
double orc_sum_compensated(const double *x, int64_t n);

/* ---- host set-up: correlation, PCA pseudo-roots ----------------------- */
int orc_sym_eigen(int n, const double *a /* n*n */, double *evals, double *evecs /* n*n, col=vec */);
void orc_exp_corr(const double *times, int n, double L, double beta, double *corr);
int orc_pseudo_sqrt(int n, const double *cov, int factors, int normalise, double *out /* n*F */);
int orc_flat_vol_pseudo_roots(const double *rate_times /* N+1 */, const double *evol_times /* S */,
                              const double *vols, int n_rates, int steps, int factors, double L,
                              double beta, double *A /* (S,N,F) */);
int orc_factor_pseudo_roots(const double *rate_times /* N+1 */, const double *vols, int n_rates,
                            int factors, double L, double beta, double *A /* (S,N,F) */,
                            double *B /* (N,F) or NULL */);

/* ---- fused pipeline: caplet portfolio price and pathwise adjoint ------ */
typedef struct {
  int n_rates, factors, paths;
  const double *f0, *taus, *disp, *A, *sigma; /* A (S,N,F), S = N */
  double strike;
} orc_caplet_model;

void orc_caplet_price_paths(const orc_caplet_model *m, const double *z /* (S,F,P) */,
                            double *price_per_path);
void orc_caplet_adjoint_paths(const orc_caplet_model *m, const double *z,
                              double *price_per_path, double *delta_sum /* N */,
                              double *vega_sum /* N */);
void orc_caplet_pipeline(const orc_caplet_model *m, const uint32_t *ints /* [D][P] */,
                         int do_adjoint, double *price, double *deltas, double *vegas);

/* ---- stage (d): Longstaff-Schwartz -------------------------------------- */
int orc_basis_count(int n_vars, int use_cross);
void orc_basis_functions(const double *vars, int n_vars, int use_cross, double *out);
int orc_cholesky_solve(int n, const double *A, const double *b, double *x);
int orc_sym_solve_truncated(int n, const double *A, const double *b, double tol, double *x);
int orc_sym_solve_cut(int n, const double *A, const double *b, double tol, double *x);
double orc_ls_first_pass(int n_ex, const double *basis_vars /* (E,maxVars,P) */,
                         const int *vars_each_step, int max_vars, const double *cash /* (E,P) */,
                         const double *ex_values /* (E,P) */, const double *numeraires /* (E,P) */,
                         int paths, int use_cross, int itm_only, double *coeffs /* (E,maxData) */);
long orc_select_fraction(const double *net, const int *pre, int *sel, int paths, long n_pre,
                         double lower_frac, double upper_frac, double sd_guess, double multiplier,
                         double *lower_cut, double *upper_cut);
long orc_select_sds(const double *net, const int *pre, int *sel, int paths, long n_pre,
                    double n_sds, double *lower_cut, double *upper_cut);
double orc_ls_multi_first_pass(int n_ex, const double *bv, const int *vars_each_step,
                               int max_vars, const double *cash, const double *ev,
                               const double *num, int paths, int use_cross, int normalise,
                               int max_regressions, int lower_path_cutoff, int selector,
                               double lower_frac, double upper_frac,
                               double sd_guess, double multiplier, double *coeffs, double *lower,
                               double *upper, double *means, double *sds);
double orc_ls_second_pass(int n_ex, const double *basis_vars, const int *vars_each_step,
                          int max_vars, const double *cash, const double *ex_values,
                          const double *numeraires, int paths, int use_cross, int itm_only,
                          const double *coeffs);
void orc_bermudan_paths(const orc_caplet_model *m, const double *z /* (S,F,P) */, int n_ex,
                        const int *ex_steps, double *bv /* (E,2,P) */, double *ev /* (E,P) */,
                        double *num /* (E,P) */);

/* ---- Bermudan swaption under a frozen strategy: price and adjoint ---- */
typedef struct {
  int n_ex;
  const int *ex_steps;
  const double *coeffs; /* (E, 6): quadratic cross basis in the normalised variables */
  const double *shift;  /* (E, 2) */
  const double *scale;  /* (E, 2) */
  int itm_only;
} orc_frozen_strategy;

void orc_bermudan_frozen_paths(const orc_caplet_model *m, const orc_frozen_strategy *fs,
                               const double *z /* (S,F,P) */, int do_adjoint,
                               double *value_per_path, int *exercise_step /* or NULL */,
                               double *delta_sum /* N */, double *vega_sum /* N */);
void orc_bermudan_given_paths(const orc_caplet_model *m, const orc_frozen_strategy *fs,
                              const double *z /* (S,F,P) */, const int *given /* P */,
                              int do_adjoint, double *value_per_path, double *delta_sum,
                              double *vega_sum);

/* ---- multi-dimensional Black-Scholes: basket Asian ------------------- */
typedef struct {
  int assets, steps, paths;
  const double *x0;      /* log spots */
  const double *vols;    /* per asset */
  const double *chol;    /* A x A lower Cholesky factor of the correlation */
  const double *weights; /* basket weights */
  double rate, dt, strike;
} orc_basket_model;

int orc_cholesky_lower(int n, const double *C, double *L);
void orc_correlated_paths(const double *in /* (S,P,F) */, double *out /* (S,N,P) */,
                          const double *A, const int *alive, int a_offset_per_step, int factors,
                          int out_dims, int paths, int steps);
void orc_multid_bs_paths(const double *log0, const double *incr /* (S,N,P) */,
                         const double *drifts /* (S,N) or NULL */, double *out /* (S,N,P) */,
                         int paths, int rates, int steps);
void orc_basket_weightings(const double *in /* (S,N,P) */, double *out /* (S,P) */,
                           const double *w, int paths, int stocks, int steps);
void orc_basket_asian_paths(const orc_basket_model *m, const double *z /* (S,A,P) */,
                            double *price_per_path, double *delta_sum, double *vega_sum);
/* OpenMP thread count used by the parallel oracle loops (CPU baseline timing). */
void orc_set_threads(int n);
int orc_get_threads(void);
#ifdef __cplusplus
}
#endif
#endif
